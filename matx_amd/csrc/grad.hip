// Row-wise correlation with one waveform per row, and the adjoints that a
// gradient through fft_conv needs (the differentiable form of
// y = ifft(fft(x, n) .* conj(fft(w, n))), examples/fft_conv.cu:90-108).
//
// Forward (conj_w = 1: correlation, 0: convolution), per row b:
//   y_b = IFFT_n(X_b .* W_b')          W' = conj(W) or W,  X = FFT_n(x_b), W = FFT_n(w_b)
// y is complex-linear in x and in conj(w) (correlation) or w (convolution).
// For a real loss with torch's gradient convention g = dL/dRe + i dL/dIm,
//   correlation:  gx = IFFT(GY .* W),        gw = IFFT(X .* conj(GY))
//   convolution:  gx = IFFT(GY .* conj(W)),  gw = IFFT(GY .* conj(X))
// with GY = FFT_n(gy), restricted to the supports of x and w.  Every one of
// these is one more row-wise correlation or convolution, so one kernel body
// serves the forward and both adjoints: a workgroup per row keeps two
// n-point spectra in LDS (n <= 8192: 2 x 68 KiB), runs the Stockham passes of
// the shared FFT engine, multiplies, inverts and stores the kept lags.
#include <algorithm>
#include <cmath>

#include "fft_engine.hpp"
#include "mxr_cache.hpp"

namespace mxr {

// One row operand: element i of row b at p[off(b) + i*s], i < len.  With
// wrap_n > 0 the operand is a lag sequence of a wrap_n-point circular
// correlation gathered into a shorter transform: buffer slot i < pos holds
// lag i, slot M - j (j <= neg) holds lag -j (element wrap_n - j), the rest
// zeros.
struct RowOp {
  const float2 *p;
  int64_t len, s;
  BatchIdx bi;  // sa: offsets of this operand per batch dim
  int64_t wrap_n;
  int64_t pos, neg;
};

struct XcArgs {
  RowOp a, b;        // out_b = IFFT(FFT(a_b) .* f(FFT(b_b)))
  int mode;          // 0: A.*conj(B), 1: A.*B, 3: |A|^2 (b unused), 4: A.*2Re(B)
  float2 *out;
  int64_t nout, os;  // kept lags, out element stride
  // full_n > 0: the M-point result stands for a full_n-point circular one
  // whose nonzero lags are 0..npos-1 and -1..-nneg (at full_n - j); lag i
  // of the output reads slot i, slot M - (full_n - i) or zero
  int64_t full_n, npos, nneg;
  BatchIdx obi;      // sa: out offsets per batch dim
  const float2 *tw;
  float scale;
};

template <int M>
struct XcCfg {
  static constexpr int NT = M / 16;
};

template <int M, bool INV>
__device__ __forceinline__ void row_fft(float2 *buf, const float2 *tw) {
  using LY = Layout<M, 1, false>;
  using SQ = Seq<M, false>;
  lds_passes<LY, SQ, M, XcCfg<M>::NT, 1, INV, 0, SQ::n>(buf, tw);
}

// Row b of the operand into LDS, zero past len.  A thread's 16 loads are
// issued together: unconditional loads (element 0 past len, zeroed by a mask)
// before any LDS store.
template <int M>
__device__ __forceinline__ void load_row(float2 *buf, const RowOp &op, int64_t row) {
  using LY = Layout<M, 1, false>;
  constexpr int NT = XcCfg<M>::NT, PER = M / NT;
  int64_t off, unused;
  batch_offsets(op.bi, row, &off, &unused);
  const float2 *src = op.p + off;
  float2 v[PER];
#pragma unroll
  for (int m = 0; m < PER; ++m) v[m] = make_float2(0.f, 0.f);
  if (op.len > 0) {
#pragma unroll
    for (int m = 0; m < PER; ++m) {
      const int i = threadIdx.x + m * NT;
      int64_t e = i;  // element of the operand row in slot i
      if (op.wrap_n > 0) e = i < op.pos ? i : (i >= M - op.neg ? op.wrap_n - (M - i) : -1);
      const bool ok = e >= 0 && e < op.len;
      v[m] = keep_if(src[(ok ? e : 0) * op.s], ok);
    }
  }
#pragma unroll
  for (int m = 0; m < PER; ++m) buf[LY::addr(0, threadIdx.x + m * NT)] = v[m];
}

template <int M>
__global__ __launch_bounds__(XcCfg<M>::NT) void xcorr_rows_kernel(XcArgs a) {
  using LY = Layout<M, 1, false>;
  constexpr int NT = XcCfg<M>::NT;
  __shared__ float2 A[LY::LDS_ELEMS];
  __shared__ float2 B[LY::LDS_ELEMS];
  const int64_t row = blockIdx.x;
  load_row<M>(A, a.a, row);
  if (a.mode != 3) load_row<M>(B, a.b, row);
  wg_barrier();
  row_fft<M, false>(A, a.tw);
  if (a.mode != 3) row_fft<M, false>(B, a.tw);
  for (int i = threadIdx.x; i < M; i += NT) {
    const float2 x = A[LY::addr(0, i)];
    float2 r;
    if (a.mode == 3) {
      r = make_float2(cabs2(x), 0.f);
    } else {
      const float2 y = B[LY::addr(0, i)];
      if (a.mode == 4) r = cscale(x, 2.0f * y.x);
      else r = a.mode == 0 ? cmulc(x, y) : cmul(x, y);
    }
    A[LY::addr(0, i)] = r;
  }
  wg_barrier();
  row_fft<M, true>(A, a.tw);
  int64_t off, unused;
  batch_offsets(a.obi, row, &off, &unused);
  float2 *dst = a.out + off;
  if (a.full_n > 0) {
    for (int64_t i = threadIdx.x; i < a.nout; i += NT) {
      float2 r = make_float2(0.f, 0.f);
      if (i < a.npos) r = A[LY::addr(0, (int)i)];
      else if (i >= a.full_n - a.nneg) r = A[LY::addr(0, (int)(M - (a.full_n - i)))];
      dst[i * a.os] = cscale(r, a.scale);
    }
    return;
  }
  for (int i = threadIdx.x; i < a.nout; i += NT) dst[(int64_t)i * a.os] = cscale(A[LY::addr(0, i)], a.scale);
}

static mxr_status xc_dispatch(int M, const XcArgs &a, int64_t rows, hipStream_t s) {
  if (rows == 0 || a.nout == 0) return MXR_SUCCESS;
  MXR_REQUIRE(rows <= 0x7fffffff, MXR_INVALID_SIZE, "row-wise correlation: too many rows");
  ProfScope ps("xcorr_rows_kernel", s);
  switch (M) {
#define MXR_XC(MM)                                                                            \
  case MM:                                                                                    \
    xcorr_rows_kernel<MM><<<dim3((unsigned)rows), dim3(XcCfg<MM>::NT), 0, s>>>(a);             \
    break;
    MXR_XC(1024)
    MXR_XC(2048)
    MXR_XC(4096)
    MXR_XC(8192)
#undef MXR_XC
    default:
      set_error(MXR_NOT_SUPPORTED, "row-wise correlation: size %d not in 1024..8192", M);
      return MXR_NOT_SUPPORTED;
  }
  MXR_HIP_TRY(hipGetLastError());
  return MXR_SUCCESS;
}

// Fills a RowOp for a [..., len] operand whose batch dims are `batch`.
static mxr_status row_op(const mxr_tensor *t, const mxr_tensor *batch, const char *fn,
                         const char *name, RowOp *op) {
  mxr_status st;
  if ((st = check_tensor(t, name, MXR_C64, 1, MXR_MAX_RANK))) return st;
  const int nb = batch->rank - 1;
  MXR_REQUIRE(t->rank == batch->rank, MXR_INVALID_DIM, "%s: %s has rank %d, expected %d", fn, name,
              t->rank, batch->rank);
  for (int d = 0; d < nb; ++d)
    MXR_REQUIRE(t->shape[d] == batch->shape[d], MXR_INVALID_SIZE, "%s: %s batch dim %d differs", fn,
                name, d);
  op->p = (const float2 *)t->data;
  op->len = t->shape[t->rank - 1];
  op->s = t->stride[t->rank - 1];
  op->bi.nd = nb;
  for (int d = 0; d < nb; ++d) {
    op->bi.shape[d] = t->shape[d];
    op->bi.sa[d] = t->stride[d];
    op->bi.sb[d] = 0;
  }
  return MXR_SUCCESS;
}

static int64_t batch_rows(const mxr_tensor *t) {
  int64_t r = 1;
  for (int d = 0; d < t->rank - 1; ++d) r *= t->shape[d];
  return r;
}

// The transform size a row correlation needs: an n-point circular
// correlation (or convolution) of rows whose linear result has at most
// `support` nonzero lags equals the linear result, so the smallest power of
// two >= support (at least 1024, the smallest kernel) gives the same lags.
static int64_t xc_size(int64_t support) {
  int64_t m = 1024;
  while (m < support) m <<= 1;
  return m;
}

static mxr_status xc_run(const char *fn, const mxr_tensor *out, const RowOp &A, const RowOp &B,
                         int mode, int64_t n, hipStream_t s) {
  MXR_REQUIRE(n >= 1024 && n <= 8192 && (n & (n - 1)) == 0, MXR_NOT_SUPPORTED,
              "%s: per-row transform size %lld must be a power of two in [1024, 8192]", fn,
              (long long)n);
  XcArgs a{};
  a.a = A;
  a.b = B;
  a.mode = mode;
  a.out = (float2 *)out->data;
  a.nout = out->shape[out->rank - 1];
  // Short rows: transform at the support instead of n (the same lags,
  // exactly; the lags outside the support are written as zeros).
  //  3 (autocorrelation): lags -(La-1) .. La-1;  0 (correlation): -(Lb-1) ..
  //  La-1;  1 (convolution): 0 .. La+Lb-2.  4 (the autocorrelation's
  //  adjoint, kept lags < La): only the lags -(La-1) .. La-1 of b reach the
  //  kept outputs, and M >= 2La - 1 keeps them from aliasing.
  const int64_t La = A.len, Lb = B.len;
  int64_t M = n;
  if (mode == 3 && La >= 1 && xc_size(2 * La - 1) < n) {
    M = xc_size(2 * La - 1);
    a.full_n = n, a.npos = La, a.nneg = La - 1;
  } else if ((mode == 0 || mode == 1) && La >= 1 && Lb >= 1 && xc_size(La + Lb - 1) < n) {
    M = xc_size(La + Lb - 1);
    a.full_n = n, a.npos = mode == 0 ? La : La + Lb - 1, a.nneg = mode == 0 ? Lb - 1 : 0;
  } else if (mode == 4 && La >= 1 && a.nout <= La && xc_size(2 * La - 1) < n) {
    M = xc_size(2 * La - 1);
    a.b.wrap_n = n, a.b.pos = La, a.b.neg = La - 1;
  }
  n = M;
  a.os = out->stride[out->rank - 1];
  a.obi.nd = out->rank - 1;
  for (int d = 0; d < out->rank - 1; ++d) {
    a.obi.shape[d] = out->shape[d];
    a.obi.sa[d] = out->stride[d];
  }
  a.tw = twiddles_device((int)n, false);
  MXR_REQUIRE(a.tw != nullptr, MXR_OUT_OF_MEMORY, "%s: twiddle table allocation failed", fn);
  a.scale = (float)(1.0 / (double)n);
  return xc_dispatch((int)n, a, batch_rows(out), s);
}

static bool same_view(const mxr_tensor *a, const mxr_tensor *b) {
  if (a->data != b->data || a->rank != b->rank) return false;
  for (int d = 0; d < a->rank; ++d)
    if (a->shape[d] != b->shape[d] || a->stride[d] != b->stride[d]) return false;
  return true;
}

// Forward with one waveform per row (called by mxr_fft_conv when w has
// batch dims).
mxr_status fft_conv_rows(const mxr_tensor *out, const mxr_tensor *x, const mxr_tensor *w,
                         int64_t n, int conj_w, hipStream_t s) {
  const char *fn = "fft_conv";
  mxr_status st;
  if ((st = check_tensor(out, "output", MXR_C64, 1, MXR_MAX_RANK))) return st;
  RowOp A{}, B{};
  if ((st = row_op(x, out, fn, "input", &A))) return st;
  if ((st = row_op(w, out, fn, "waveform", &B))) return st;
  MXR_REQUIRE(A.len <= n && B.len <= n && out->shape[out->rank - 1] <= n, MXR_INVALID_SIZE,
              "%s: rows longer than the transform size %lld", fn, (long long)n);
  MXR_REQUIRE(!views_overlap(*out, *x) && !views_overlap(*out, *w), MXR_ALIAS_ERROR,
              "%s: output overlaps an input", fn);
  const int mode = conj_w ? (same_view(x, w) ? 3 : 0) : 1;
  return xc_run(fn, out, A, B, mode, n, s);
}

static mxr_status fft_conv_adjoint_impl(const mxr_tensor *gx, const mxr_tensor *gw,
                                        const mxr_tensor *gy, const mxr_tensor *x,
                                        const mxr_tensor *w, int64_t n, int conj_w,
                                        hipStream_t s) {
  const char *fn = "fft_conv_adjoint";
  mxr_status st;
  if ((st = check_tensor(gy, "gy", MXR_C64, 1, MXR_MAX_RANK))) return st;
  MXR_REQUIRE(gx || gw, MXR_INVALID_PARAMETER, "%s: no gradient requested", fn);
  RowOp X{}, W{}, GY{};
  if ((st = row_op(x, gy, fn, "x", &X))) return st;
  if ((st = row_op(gy, gy, fn, "gy", &GY))) return st;
  MXR_REQUIRE(w->rank == gy->rank, MXR_NOT_SUPPORTED,
              "%s: the waveform needs one row per signal row (rank %d)", fn, gy->rank);
  if ((st = row_op(w, gy, fn, "w", &W))) return st;
  MXR_REQUIRE(X.len <= n && W.len <= n && GY.len <= n, MXR_INVALID_SIZE,
              "%s: rows longer than the transform size %lld", fn, (long long)n);
  if (gx) {
    if ((st = check_tensor(gx, "gx", MXR_C64, 1, MXR_MAX_RANK))) return st;
    for (int d = 0; d < gx->rank; ++d)
      MXR_REQUIRE(gx->rank == x->rank && gx->shape[d] == x->shape[d], MXR_INVALID_SIZE,
                  "%s: gx must have x's shape", fn);
    MXR_REQUIRE(!views_overlap(*gx, *gy) && !views_overlap(*gx, *x) && !views_overlap(*gx, *w),
                MXR_ALIAS_ERROR, "%s: gx overlaps an input", fn);
  }
  if (gw) {
    if ((st = check_tensor(gw, "gw", MXR_C64, 1, MXR_MAX_RANK))) return st;
    for (int d = 0; d < gw->rank; ++d)
      MXR_REQUIRE(gw->rank == w->rank && gw->shape[d] == w->shape[d], MXR_INVALID_SIZE,
                  "%s: gw must have w's shape", fn);
    MXR_REQUIRE(!views_overlap(*gw, *gy) && !views_overlap(*gw, *x) && !views_overlap(*gw, *w),
                MXR_ALIAS_ERROR, "%s: gw overlaps an input", fn);
    if (gx) MXR_REQUIRE(!views_overlap(*gx, *gw), MXR_ALIAS_ERROR, "%s: gx and gw overlap", fn);
  }
  // correlation: gx = IFFT(GY.*W), gw = IFFT(X.*conj(GY))
  // convolution: gx = IFFT(GY.*conj(W)), gw = IFFT(GY.*conj(X))
  if (gx && (st = xc_run(fn, gx, GY, W, conj_w ? 1 : 0, n, s))) return st;
  if (gw && (st = xc_run(fn, gw, conj_w ? X : GY, conj_w ? GY : X, 0, n, s))) return st;
  return MXR_SUCCESS;
}

// Gradient w.r.t. s of r = fft_conv(s, s, n) (a waveform's matched filter
// to itself): gx + gw = IFFT(GY .* S) + IFFT(S .* conj(GY)) = IFFT(S .* 2Re(GY)).
static mxr_status autocorr_adjoint_impl(const mxr_tensor *gs, const mxr_tensor *gy,
                                        const mxr_tensor *s, int64_t n, hipStream_t st_) {
  const char *fn = "autocorr_adjoint";
  mxr_status st;
  if ((st = check_tensor(gy, "gy", MXR_C64, 1, MXR_MAX_RANK))) return st;
  if ((st = check_tensor(gs, "gs", MXR_C64, 1, MXR_MAX_RANK))) return st;
  RowOp S{}, GY{};
  if ((st = row_op(s, gy, fn, "s", &S))) return st;
  if ((st = row_op(gy, gy, fn, "gy", &GY))) return st;
  MXR_REQUIRE(S.len <= n && GY.len <= n, MXR_INVALID_SIZE,
              "%s: rows longer than the transform size %lld", fn, (long long)n);
  MXR_REQUIRE(gs->rank == s->rank, MXR_INVALID_SIZE, "%s: gs must have s's shape", fn);
  for (int d = 0; d < gs->rank; ++d)
    MXR_REQUIRE(gs->shape[d] == s->shape[d], MXR_INVALID_SIZE, "%s: gs must have s's shape", fn);
  MXR_REQUIRE(!views_overlap(*gs, *gy) && !views_overlap(*gs, *s), MXR_ALIAS_ERROR,
              "%s: gs overlaps an input", fn);
  return xc_run(fn, gs, S, GY, 4, n, st_);
}

// ---------------------------------------------------- phase-only ISL head
// The waveform term of the design loss in two kernels: ISL_b of the n-point
// autocorrelation of s_b = exp(j phi_b), and its gradient w.r.t. phi.  The
// phases become samples on load and the autocorrelation never leaves LDS;
// the transform runs at the support (xc_size(2L - 1)) when that is shorter
// than n, which gives the same lags (the others are zero and add nothing to
// either sum).  Sums run in double, as in the stand-alone ISL head.  The
// adjoint forms gr = dISL/dr in place of r, then gs = IFFT(S .* 2Re(GR)) / M
// (the autocorrelation's adjoint) and gphi = Im(conj(s) gs) = cos(phi)
// Im(gs) - sin(phi) Re(gs) (the chain rule through s = exp(j phi), torch's
// g = dL/dRe + i dL/dIm convention).
struct PhIslArgs {
  const float *phi;
  int64_t ps_b, ps_l;  // phi [B, L]
  int L;
  float *isl;          // forward output [B]
  int64_t is_b;
  const float *gisl;   // adjoint input [B]
  int64_t gs_b;
  float *gphi;         // adjoint output [B, L]
  int64_t gps_b, gps_l;
  const float2 *tw;
};

template <int NT>
__device__ __forceinline__ double xc_block_sum(double v, double *red) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  __syncthreads();  // red is free
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int w = 0; w < NT / 64; ++w) t += red[w];
  return t;
}

// s = exp(j phi) of row `row` into buf, zero past L.  Pre: L >= 1.
template <int M>
__device__ __forceinline__ void load_phases(float2 *buf, const PhIslArgs &a, int64_t row) {
  using LY = Layout<M, 1, false>;
  constexpr int NT = XcCfg<M>::NT, PER = M / NT;
  const float *src = a.phi + row * a.ps_b;
  float ph[PER];
#pragma unroll
  for (int m = 0; m < PER; ++m) {
    const int i = threadIdx.x + m * NT;
    ph[m] = src[(int64_t)(i < a.L ? i : 0) * a.ps_l];
  }
#pragma unroll
  for (int m = 0; m < PER; ++m) {
    const int i = threadIdx.x + m * NT;
    float sn, cs;
    sincosf(ph[m], &sn, &cs);
    buf[LY::addr(0, i)] = i < a.L ? make_float2(cs, sn) : make_float2(0.f, 0.f);
  }
}

// |FFT(s)|^2 -> inverse: the unscaled autocorrelation in buf (spec, when
// given, keeps the spectrum S).  Returns sum |r|^2 and |r[0]|^2 (unscaled).
template <int M>
__device__ __forceinline__ void phase_autocorr(float2 *buf, float2 *spec, const PhIslArgs &a,
                                               int64_t row, double *red, double *S, double *P0) {
  using LY = Layout<M, 1, false>;
  constexpr int NT = XcCfg<M>::NT;
  load_phases<M>(buf, a, row);
  wg_barrier();
  row_fft<M, false>(buf, a.tw);
  for (int i = threadIdx.x; i < M; i += NT) {
    const float2 x = buf[LY::addr(0, i)];
    if (spec) spec[LY::addr(0, i)] = x;
    buf[LY::addr(0, i)] = make_float2(cabs2(x), 0.f);
  }
  wg_barrier();
  row_fft<M, true>(buf, a.tw);
  const float2 v0 = buf[LY::addr(0, 0)];
  double t = 0.0;
  for (int i = threadIdx.x; i < M; i += NT) {
    const float2 v = buf[LY::addr(0, i)];
    t += (double)v.x * v.x + (double)v.y * v.y;
  }
  *S = xc_block_sum<NT>(t, red);  // its barriers also order the reads of r[0] before any write
  *P0 = (double)v0.x * v0.x + (double)v0.y * v0.y;
}

template <int M>
__global__ __launch_bounds__(XcCfg<M>::NT) void phase_isl_kernel(PhIslArgs a) {
  using LY = Layout<M, 1, false>;
  constexpr int NT = XcCfg<M>::NT;
  __shared__ float2 A[LY::LDS_ELEMS];
  __shared__ double red[NT / 64];
  const int64_t row = blockIdx.x;
  double S, P0;
  phase_autocorr<M>(A, nullptr, a, row, red, &S, &P0);
  if (threadIdx.x == 0) a.isl[row * a.is_b] = (float)((S - P0) / P0);
}

template <int M>
__global__ __launch_bounds__(XcCfg<M>::NT) void phase_isl_adj_kernel(PhIslArgs a) {
  using LY = Layout<M, 1, false>;
  constexpr int NT = XcCfg<M>::NT;
  __shared__ float2 A[LY::LDS_ELEMS];
  __shared__ float2 B[LY::LDS_ELEMS];
  __shared__ double red[NT / 64];
  const int64_t row = blockIdx.x;
  double S, P0;
  phase_autocorr<M>(A, B, a, row, red, &S, &P0);
  // with r = r_raw / M: dISL/dr = 2 g r / P0 (l != 0), -2 g (S - P0) / P0^2 r
  // (l = 0), in terms of the unscaled sums: factors M / P0_raw and
  // M (S_raw - P0_raw) / P0_raw^2
  const double g = (double)a.gisl[row * a.gs_b];
  const float side = (float)(2.0 * g * M / P0);
  const float lag0 = (float)(-2.0 * g * (S - P0) * M / (P0 * P0));
  for (int i = threadIdx.x; i < M; i += NT) A[LY::addr(0, i)] = cscale(A[LY::addr(0, i)], i == 0 ? lag0 : side);
  wg_barrier();
  row_fft<M, false>(A, a.tw);
  for (int i = threadIdx.x; i < M; i += NT)
    A[LY::addr(0, i)] = cscale(B[LY::addr(0, i)], 2.0f * A[LY::addr(0, i)].x);
  wg_barrier();
  row_fft<M, true>(A, a.tw);
  const float inv = 1.0f / (float)M;
  const float *src = a.phi + row * a.ps_b;
  float *dst = a.gphi + row * a.gps_b;
  for (int i = threadIdx.x; i < a.L; i += NT) {
    const float2 gs = cscale(A[LY::addr(0, i)], inv);
    float sn, cs;
    sincosf(src[(int64_t)i * a.ps_l], &sn, &cs);
    dst[(int64_t)i * a.gps_l] = cs * gs.y - sn * gs.x;
  }
}

// [B, L] f32 phases: a row view with its strides.
static mxr_status phase_isl_args(const char *fn, const mxr_tensor *phi, int64_t n, PhIslArgs *a,
                                 int64_t *M) {
  mxr_status st;
  if ((st = check_tensor(phi, "phi", MXR_F32, 2, 2))) return st;
  MXR_REQUIRE(n >= 1024 && n <= 8192 && (n & (n - 1)) == 0, MXR_NOT_SUPPORTED,
              "%s: transform size %lld must be a power of two in [1024, 8192]", fn, (long long)n);
  MXR_REQUIRE(phi->shape[1] >= 1 && phi->shape[1] <= n, MXR_INVALID_SIZE,
              "%s: need 1 <= L <= n (L = %lld)", fn, (long long)phi->shape[1]);
  MXR_REQUIRE(phi->shape[0] <= 0x7fffffff, MXR_INVALID_SIZE, "%s: too many rows", fn);
  a->phi = (const float *)phi->data;
  a->ps_b = phi->stride[0];
  a->ps_l = phi->stride[1];
  a->L = (int)phi->shape[1];
  *M = std::min<int64_t>(n, xc_size(2 * phi->shape[1] - 1));
  a->tw = twiddles_device((int)*M, false);
  MXR_REQUIRE(a->tw != nullptr, MXR_OUT_OF_MEMORY, "%s: twiddle table allocation failed", fn);
  return MXR_SUCCESS;
}

#define MXR_PH_DISPATCH(KERNEL, M, rows, a, s)                                                    \
  switch (M) {                                                                                   \
    case 1024: KERNEL<1024><<<dim3((unsigned)(rows)), dim3(XcCfg<1024>::NT), 0, s>>>(a); break; \
    case 2048: KERNEL<2048><<<dim3((unsigned)(rows)), dim3(XcCfg<2048>::NT), 0, s>>>(a); break; \
    case 4096: KERNEL<4096><<<dim3((unsigned)(rows)), dim3(XcCfg<4096>::NT), 0, s>>>(a); break; \
    default: KERNEL<8192><<<dim3((unsigned)(rows)), dim3(XcCfg<8192>::NT), 0, s>>>(a); break;   \
  }

static mxr_status phase_isl_impl(const mxr_tensor *isl, const mxr_tensor *phi, int64_t n,
                                 hipStream_t s) {
  const char *fn = "phase_isl";
  PhIslArgs a{};
  int64_t M;
  mxr_status st;
  if ((st = phase_isl_args(fn, phi, n, &a, &M))) return st;
  if ((st = check_tensor(isl, "isl", MXR_F32, 1, 1))) return st;
  MXR_REQUIRE(isl->shape[0] == phi->shape[0], MXR_INVALID_SIZE, "%s: isl must be [B]", fn);
  MXR_REQUIRE(!views_overlap(*isl, *phi), MXR_ALIAS_ERROR, "%s: isl overlaps phi", fn);
  if (phi->shape[0] == 0) return MXR_SUCCESS;
  a.isl = (float *)isl->data;
  a.is_b = isl->stride[0];
  ProfScope ps("phase_isl_kernel", s);
  MXR_PH_DISPATCH(phase_isl_kernel, M, phi->shape[0], a, s)
  MXR_HIP_TRY(hipGetLastError());
  return MXR_SUCCESS;
}

static mxr_status phase_isl_adjoint_impl(const mxr_tensor *gphi, const mxr_tensor *gisl,
                                         const mxr_tensor *phi, int64_t n, hipStream_t s) {
  const char *fn = "phase_isl_adjoint";
  PhIslArgs a{};
  int64_t M;
  mxr_status st;
  if ((st = phase_isl_args(fn, phi, n, &a, &M))) return st;
  if ((st = check_tensor(gisl, "gisl", MXR_F32, 1, 1))) return st;
  if ((st = check_tensor(gphi, "gphi", MXR_F32, 2, 2))) return st;
  MXR_REQUIRE(gisl->shape[0] == phi->shape[0] && gphi->shape[0] == phi->shape[0] &&
                  gphi->shape[1] == phi->shape[1],
              MXR_INVALID_SIZE, "%s: gisl [B] and gphi [B, L] must match phi", fn);
  MXR_REQUIRE(!views_overlap(*gphi, *phi) && !views_overlap(*gphi, *gisl), MXR_ALIAS_ERROR,
              "%s: gphi overlaps an input", fn);
  if (phi->shape[0] == 0) return MXR_SUCCESS;
  a.gisl = (const float *)gisl->data;
  a.gs_b = gisl->stride[0];
  a.gphi = (float *)gphi->data;
  a.gps_b = gphi->stride[0];
  a.gps_l = gphi->stride[1];
  ProfScope ps("phase_isl_adj_kernel", s);
  MXR_PH_DISPATCH(phase_isl_adj_kernel, M, phi->shape[0], a, s)
  MXR_HIP_TRY(hipGetLastError());
  return MXR_SUCCESS;
}
#undef MXR_PH_DISPATCH

}  // namespace mxr

extern "C" mxr_status mxr_autocorr_adjoint(const mxr_tensor *gs, const mxr_tensor *gy,
                                           const mxr_tensor *s, int64_t n, hipStream_t stream) {
  return mxr::guarded("mxr_autocorr_adjoint",
                      [&] { return mxr::autocorr_adjoint_impl(gs, gy, s, n, stream); });
}

extern "C" mxr_status mxr_fft_conv_adjoint(const mxr_tensor *gx, const mxr_tensor *gw,
                                           const mxr_tensor *gy, const mxr_tensor *x,
                                           const mxr_tensor *w, int64_t n, int conj_w,
                                           hipStream_t stream) {
  return mxr::guarded("mxr_fft_conv_adjoint", [&] {
    return mxr::fft_conv_adjoint_impl(gx, gw, gy, x, w, n, conj_w, stream);
  });
}

extern "C" mxr_status mxr_phase_isl(const mxr_tensor *isl, const mxr_tensor *phi, int64_t n,
                                    hipStream_t stream) {
  return mxr::guarded("mxr_phase_isl", [&] { return mxr::phase_isl_impl(isl, phi, n, stream); });
}

extern "C" mxr_status mxr_phase_isl_adjoint(const mxr_tensor *gphi, const mxr_tensor *gisl,
                                            const mxr_tensor *phi, int64_t n, hipStream_t stream) {
  return mxr::guarded("mxr_phase_isl_adjoint",
                      [&] { return mxr::phase_isl_adjoint_impl(gphi, gisl, phi, n, stream); });
}
