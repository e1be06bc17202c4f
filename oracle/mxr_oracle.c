/*
 * mxr_oracle.c -- FP64 CPU oracle for the radar hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP library is
 * compared against.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it; the product path (matx_amd/, include/) never does.
 *
 * Every routine is a plain restatement of the reference's arithmetic in double
 * precision.  Citations are to the MatX sources (paths relative to the MatX
 * repository root):
 *
 *   orc_fft / orc_dft ........ DFT definition with the cuFFT/FFTW sign and norm
 *                              rules: include/matx/transforms/fft/fft_cuda.h:126-188,
 *                              include/matx/core/operator_options.h:45-49
 *   orc_hamming .............. include/matx/generators/hamming.h:96-99
 *   orc_waveform_spectrum .... examples/simple_radar_pipeline.h:276-296
 *   orc_pulse_compress ....... examples/simple_radar_pipeline.h:298-300
 *   orc_fft_conv ............. examples/fft_conv.cu:90-108 (circular correlation
 *                              form ifft(fft(x,N)*conj(fft(w,N))))
 *   orc_conv1d ............... include/matx/transforms/conv.h:226-289 (host loop,
 *                              FULL/SAME/VALID start offsets) and
 *                              include/matx/kernels/conv.cuh:155-168
 *   orc_conv2d ............... include/matx/transforms/conv.h:325-399
 *   orc_tpc .................. examples/simple_radar_pipeline.h:321-328 (SAME) and
 *                              test/test_vectors/generators/01_radar.py:65-66 (VALID)
 *   orc_doppler .............. examples/simple_radar_pipeline.h:344-357 and
 *                              test/test_vectors/generators/01_radar.py:75-78
 *   orc_cfar ................. examples/simple_radar_pipeline.h:394-436,485-492,
 *                              calcDets body :86-98;
 *                              test/test_vectors/generators/01_radar.py:86-117
 *
 * How each routine is pinned (tests/test_oracle.py): numpy's FP64 FFT and an
 * O(n^2) DFT for orc_fft; the ConvCorr.cu known-answer tables and a brute-force
 * FP64 formula for conv1d/conv2d; scipy.signal.convolve2d for conv2d and the
 * CFAR sums; np.hamming for the window; closed-form CFAR norm counts; and the
 * reference test's 17-detections-per-channel invariant for the full model chain.
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef double complex cplx;
typedef int64_t i64;

#define ORC_PI 3.14159265358979323846264338327950288

static int is_pow2(i64 n) { return n > 0 && (n & (n - 1)) == 0; }

/* O(n^2) DFT.  sign = -1 forward, +1 inverse; no normalisation. */
void orc_dft(const cplx *x, cplx *y, i64 n, int inverse)
{
  const double sgn = inverse ? 1.0 : -1.0;
  for (i64 k = 0; k < n; k++) {
    cplx acc = 0;
    for (i64 t = 0; t < n; t++) {
      /* reduce the angle index modulo n before scaling to keep it exact */
      i64 m = (k * t) % n;
      double a = sgn * 2.0 * ORC_PI * (double)m / (double)n;
      acc += x[t] * (cos(a) + I * sin(a));
    }
    y[k] = acc;
  }
}

/* In-place iterative radix-2 Cooley-Tukey (bit reversal, then butterflies).
 * Power-of-two n only; any other n goes through orc_dft.  No normalisation:
 * forward uses e^{-2 pi i nk/N}, inverse e^{+2 pi i nk/N} (cuFFT/FFTW sign
 * convention). */
void orc_fft(cplx *x, i64 n, int inverse)
{
  if (n <= 1) return;
  if (!is_pow2(n)) {
    cplx *tmp = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    orc_dft(x, tmp, n, inverse);
    memcpy(x, tmp, sizeof(cplx) * (size_t)n);
    free(tmp);
    return;
  }
  for (i64 i = 1, j = 0; i < n; i++) {
    i64 bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) {
      cplx t = x[i];
      x[i] = x[j];
      x[j] = t;
    }
  }
  const double sgn = inverse ? 1.0 : -1.0;
  for (i64 len = 2; len <= n; len <<= 1) {
    const i64 half = len >> 1;
    for (i64 k = 0; k < half; k++) {
      /* every twiddle straight from cos/sin (no recurrences) */
      double a = sgn * 2.0 * ORC_PI * (double)k / (double)len;
      cplx w = cos(a) + I * sin(a);
      for (i64 s = 0; s < n; s += len) {
        cplx u = x[s + k];
        cplx v = x[s + k + half] * w;
        x[s + k] = u + v;
        x[s + k + half] = u - v;
      }
    }
  }
}

/* Batched FFT over `rows` contiguous rows of length n (row stride ld). */
void orc_fft_rows(cplx *x, i64 rows, i64 ld, i64 n, int inverse)
{
#pragma omp parallel for schedule(static)
  for (i64 r = 0; r < rows; r++) orc_fft(x + r * ld, n, inverse);
}

double orc_hamming(i64 i, i64 n)
{
  if (n <= 1) return 1.0;
  return 0.54 - 0.46 * cos(2.0 * ORC_PI * (double)i / (double)(n - 1));
}

/* Wc = conj(FFT_N(pad_N(w2))), w2 = w1/||w1||, w1 = w .* hamming(L).
 * window / normalize switch the two preparation steps (both on for the
 * pipeline; both off for plain fft_conv).  If w_out is non-NULL the prepared
 * time-domain waveform (w2) is written there (MatX mutates its waveform view
 * in place, simple_radar_pipeline.h:288-294). */
void orc_waveform_spectrum(const cplx *w, i64 L, i64 N, int window,
                           int normalize, cplx *Wc, cplx *w_out)
{
  cplx *buf = (cplx *)calloc((size_t)N, sizeof(cplx));
  i64 Lc = L < N ? L : N;
  for (i64 l = 0; l < Lc; l++) buf[l] = w[l] * (window ? orc_hamming(l, L) : 1.0);
  if (normalize) {
    double s = 0;
    for (i64 l = 0; l < Lc; l++) s += creal(buf[l]) * creal(buf[l]) + cimag(buf[l]) * cimag(buf[l]);
    double nrm = sqrt(s);
    for (i64 l = 0; l < Lc; l++) buf[l] /= nrm;
  }
  if (w_out) memcpy(w_out, buf, sizeof(cplx) * (size_t)Lc);
  orc_fft(buf, N, 0);
  for (i64 k = 0; k < N; k++) Wc[k] = conj(buf[k]);
  free(buf);
}

/* y[r][n] = (1/N) IFFT_N( FFT_N(pad_N(x[r][0..S))) .* Wc )[n], n < nout.
 * x rows have stride ldx, y rows stride ldy.  nout <= N. */
void orc_pulse_compress(const cplx *x, i64 rows, i64 ldx, i64 S,
                        const cplx *Wc, i64 N, i64 nout, cplx *y, i64 ldy)
{
#pragma omp parallel
  {
    cplx *buf = (cplx *)malloc(sizeof(cplx) * (size_t)N);
#pragma omp for schedule(static)
    for (i64 r = 0; r < rows; r++) {
      i64 Sc = S < N ? S : N;
      for (i64 i = 0; i < N; i++) buf[i] = i < Sc ? x[r * ldx + i] : 0;
      orc_fft(buf, N, 0);
      for (i64 k = 0; k < N; k++) buf[k] *= Wc[k];
      orc_fft(buf, N, 1);
      for (i64 i = 0; i < nout; i++) y[r * ldy + i] = buf[i] / (double)N;
    }
    free(buf);
  }
}

/* fft_conv: out = ifft(fft(x,N) .* conj(fft(w,N))), all N outputs. */
void orc_fft_conv(const cplx *x, i64 rows, i64 ldx, i64 S, const cplx *w,
                  i64 L, i64 N, cplx *y, i64 ldy)
{
  cplx *Wc = (cplx *)malloc(sizeof(cplx) * (size_t)N);
  orc_waveform_spectrum(w, L, N, 0, 0, Wc, NULL);
  orc_pulse_compress(x, rows, ldx, S, Wc, N, N, y, ldy);
  free(Wc);
}

/* Start offset into the FULL result for each mode (conv.h:253-258,
 * conv.cuh:155-168).  mode: 0 FULL, 1 SAME, 2 VALID. */
static i64 conv_start(i64 flen, int mode)
{
  if (mode == 1) return (flen & 1) ? (flen - 1) / 2 : flen / 2 - 1;
  if (mode == 2) return flen - 1;
  return 0;
}

i64 orc_conv1d_outlen(i64 n, i64 flen, int mode)
{
  if (mode == 0) return n + flen - 1;
  if (mode == 1) return n;
  return n - flen + 1;
}

/* Direct 1-D convolution of `rows` signals (row stride ldx / ldo). */
void orc_conv1d(const cplx *x, i64 rows, i64 ldx, i64 n, const cplx *f,
                i64 flen, int mode, cplx *out, i64 ldo)
{
  const i64 start = conv_start(flen, mode);
  const i64 olen = orc_conv1d_outlen(n, flen, mode);
#pragma omp parallel for schedule(static)
  for (i64 r = 0; r < rows; r++) {
    for (i64 o = 0; o < olen; o++) {
      const i64 full = o + start;
      cplx s = 0;
      for (i64 t = 0; t < flen; t++) {
        const i64 si = full - t;
        if (si >= 0 && si < n) s += x[r * ldx + si] * f[t];
      }
      out[r * ldo + o] = s;
    }
  }
}

/* Direct 2-D convolution over `batch` planes (H x W contiguous),
 * filter fh x fw, output plane oh x ow contiguous.  Same index rule as the
 * host loop in conv.h:340-380 (filter flipped). */
void orc_conv2d(const cplx *x, i64 batch, i64 H, i64 W, const cplx *f,
                i64 fh, i64 fw, int mode, cplx *out)
{
  i64 dy = 0, dx = 0, oh, ow;
  if (mode == 1) { dy = fh / 2; dx = fw / 2; oh = H; ow = W; }
  else if (mode == 0) { dy = fh - 1; dx = fw - 1; oh = H + fh - 1; ow = W + fw - 1; }
  else { oh = H - fh + 1; ow = W - fw + 1; }
#pragma omp parallel for collapse(2) schedule(static)
  for (i64 b = 0; b < batch; b++)
    for (i64 oy = 0; oy < oh; oy++)
      for (i64 ox = 0; ox < ow; ox++) {
        cplx s = 0;
        for (i64 fy = 0; fy < fh; fy++) {
          i64 iy = oy + fy - dy;
          if (iy < 0 || iy >= H) continue;
          for (i64 fx = 0; fx < fw; fx++) {
            i64 ix = ox + fx - dx;
            if (ix < 0 || ix >= W) continue;
            s += x[(b * H + iy) * W + ix] * f[(fh - 1 - fy) * fw + (fw - 1 - fx)];
          }
        }
        out[(b * oh + oy) * ow + ox] = s;
      }
}

/* Three-pulse canceller [1,-2,1] along pulses.
 * y: C x P x (ldy) complex, pulse p of channel c at y[(c*P + p)*ldy + r].
 * mode 1 (SAME, MatX stage): out rows 0..P-1, t[p] = y[p-1]-2y[p]+y[p+1],
 *   zero outside [0,P).  mode 2 (VALID, numpy model): out rows 0..P-3,
 *   t[i] = y[i]-2y[i+1]+y[i+2].  out: C x rows_out x R contiguous. */
void orc_tpc(const cplx *y, i64 C, i64 P, i64 R, i64 ldy, int mode, cplx *out)
{
  const i64 rows = mode == 2 ? P - 2 : P;
#pragma omp parallel for collapse(2) schedule(static)
  for (i64 c = 0; c < C; c++)
    for (i64 i = 0; i < rows; i++) {
      const i64 p = mode == 2 ? i + 1 : i;
      for (i64 r = 0; r < R; r++) {
        cplx a = p - 1 >= 0 ? y[(c * P + p - 1) * ldy + r] : 0;
        cplx b = y[(c * P + p) * ldy + r];
        cplx d = p + 1 < P ? y[(c * P + p + 1) * ldy + r] : 0;
        out[(c * rows + i) * R + r] = a - 2.0 * b + d;
      }
    }
}

/* Doppler: rows 0..win_rows-1 of t (C x rows x R) are multiplied by
 * hamming(win_rows); rows >= win_rows pass unwindowed; rows rows..Nd-1 are
 * zero; FFT of length Nd along the row axis.  D: C x Nd x R. */
void orc_doppler(const cplx *t, i64 C, i64 rows, i64 R, i64 win_rows, i64 Nd,
                 cplx *D)
{
#pragma omp parallel
  {
    cplx *buf = (cplx *)malloc(sizeof(cplx) * (size_t)Nd);
#pragma omp for collapse(2) schedule(static)
    for (i64 c = 0; c < C; c++)
      for (i64 r = 0; r < R; r++) {
        for (i64 p = 0; p < Nd; p++) {
          cplx v = 0;
          if (p < rows) {
            v = t[(c * rows + p) * R + r];
            if (p < win_rows) v *= orc_hamming(p, win_rows);
          }
          buf[p] = v;
        }
        orc_fft(buf, Nd, 0);
        for (i64 k = 0; k < Nd; k++) D[(c * Nd + k) * R + r] = buf[k];
      }
    free(buf);
  }
}

/* CA-CFAR on a power map pw (C x Ny x Nx).  mask (my x mx) is applied as
 * MatX does it: FULL conv2d then trim at offset (my/2, mx/2)
 * (simple_radar_pipeline.h:402-415), i.e.
 *   sum(y,x) = sum_{i,j} mask[i][j] * pw[y + my/2 - i][x + mx/2 - j]
 * with zero padding; norm is the same sum over ones.  ba = sum/norm,
 * alpha = norm*(pfa^(-1/norm)-1), det = pw > alpha*ba.  margin (optional)
 * receives (pw - alpha*ba)/(alpha*ba), the relative distance to threshold. */
void orc_cfar(const double *pw, i64 C, i64 Ny, i64 Nx, const double *mask,
              i64 my, i64 mx, double pfa, double *ba, double *norm,
              double *margin, int32_t *dets)
{
  const i64 cy = my / 2, cx = mx / 2;
#pragma omp parallel for collapse(2) schedule(static)
  for (i64 c = 0; c < C; c++)
    for (i64 y = 0; y < Ny; y++)
      for (i64 x = 0; x < Nx; x++) {
        double s = 0, nm = 0;
        for (i64 i = 0; i < my; i++) {
          i64 yy = y + cy - i;
          if (yy < 0 || yy >= Ny) continue;
          for (i64 j = 0; j < mx; j++) {
            i64 xx = x + cx - j;
            if (xx < 0 || xx >= Nx) continue;
            s += mask[i * mx + j] * pw[(c * Ny + yy) * Nx + xx];
            nm += mask[i * mx + j];
          }
        }
        const i64 o = (c * Ny + y) * Nx + x;
        const double b = s / nm;
        const double alpha = nm * (pow(pfa, -1.0 / nm) - 1.0);
        const double thr = alpha * b;
        if (ba) ba[o] = b;
        if (norm) norm[o] = nm;
        if (margin) margin[o] = (pw[o] - thr) / thr;
        dets[o] = pw[o] > thr ? 1 : 0;
      }
}

/* |D|^2 */
void orc_abs2(const cplx *D, i64 n, double *pw)
{
#pragma omp parallel for schedule(static)
  for (i64 i = 0; i < n; i++) pw[i] = creal(D[i]) * creal(D[i]) + cimag(D[i]) * cimag(D[i]);
}
