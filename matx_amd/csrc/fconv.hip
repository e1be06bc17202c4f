// 1-D convolution and correlation by the FFT method (conv1d / corr with
// MATX_C_METHOD_FFT: include/matx/transforms/conv.h:51-172, corr.h:60-67).
// MatX transforms both operands at P = N + L - 1 points, multiplies, inverse
// transforms and slices the mode's window: FULL [0, P), SAME [s, s + N) with
// s = (L-1)/2 for odd L and L/2 - 1 for even L, VALID [L-1, N).  corr(x, h)
// is conv1d(x, reverse(conj(h))).
//
// Here the linear convolution is exact at any transform length >= P, so:
//  * P <= 16384: one workgroup per row zero-pads x to the power of two
//    M >= P in LDS, transforms, multiplies by the filter spectrum (computed
//    once per call by fconv_spec_kernel), inverse transforms and writes only
//    the mode's window.
//  * longer rows (c64, one filter for every row): output k >= L-1 of the
//    FULL convolution is lag k-(L-1) >= 0 of the correlation of x with
//    g = reverse(conj(h)) (for corr: g = h), which the overlap-save
//    pulse-compression kernels compute for rows of any length; the at most
//    L-1 outputs before it read only x[0, L-1) and come from the fused kernel
//    run on that prefix.
#include <algorithm>

#include "fft_engine.hpp"
#include "mxr_cache.hpp"
#include "mxr_internal.hpp"

namespace mxr {

mxr_status pc_common(const char *fn, const mxr_tensor *y, const mxr_tensor *x, const mxr_tensor *w,
                     int64_t n, int window, int normalize, int writeback, int conj_w,
                     hipStream_t stream);

// TB rows per 256-thread workgroup below 4096 points, one row of M/16
// threads from 4096 up (one radix-16 butterfly per thread per pass).
template <int M>
struct FcCfg {
  static constexpr int TB = M >= 4096 ? 1 : 4096 / M;
  static constexpr int NT = M >= 4096 ? M / 16 : 256;
};

struct FcArgs {
  const void *x;   // rows of N samples, c64 or f32
  int x_c;
  int64_t xlen;    // samples per row used (<= M)
  int64_t xs;      // stride along the axis
  const void *h;   // filter [L], c64 or f32
  int h_c;
  int64_t hlen, hs;
  int corr;        // transform reverse(conj(h)) instead of h
  BatchIdx bi;     // rows: sa = x offsets, sb = out offsets
  int64_t rows;
  void *out;       // c64, or f32 (real part) when x and h are real
  int out_c;
  int64_t os;
  int64_t o0, olen;  // window [o0, o0 + olen) of the FULL convolution
  float2 *spec;      // M-point filter spectrum, scaled by 1/M
  const float2 *tw;
};

__device__ __forceinline__ float2 ld_cf(const void *p, int cplx, int64_t i) {
  return cplx ? static_cast<const float2 *>(p)[i] : make_float2(static_cast<const float *>(p)[i], 0.f);
}

// Filter spectrum: FFT_M of h (or of reverse(conj(h))) zero-padded, / M.
template <int M>
__global__ __launch_bounds__(FcCfg<M>::NT) void fconv_spec_kernel(FcArgs a) {
  constexpr int NT = FcCfg<M>::NT;
  using LY = Layout<M, 1, false>;
  using SQ = Seq<M, false>;
  __shared__ float2 lds[LY::LDS_ELEMS];
  for (int i = threadIdx.x; i < M; i += NT) {
    float2 v = make_float2(0.f, 0.f);
    if (i < a.hlen) {
      if (a.corr) {
        v = ld_cf(a.h, a.h_c, (a.hlen - 1 - i) * a.hs);
        v.y = -v.y;
      } else {
        v = ld_cf(a.h, a.h_c, i * a.hs);
      }
    }
    lds[LY::addr(0, i)] = v;
  }
  __syncthreads();
  lds_passes<LY, SQ, M, NT, 1, false, 0, SQ::n>(lds, a.tw);
  const float s = 1.0f / (float)M;  // exact: a power of two
  for (int k = threadIdx.x; k < M; k += NT) a.spec[k] = cscale(lds[LY::addr(0, k)], s);
}

// TB rows per workgroup: x zero-padded to M, FFT, spectrum multiply, IFFT,
// the window [o0, o0 + olen) stored.
template <int M>
__global__ __launch_bounds__(FcCfg<M>::NT) void fconv_kernel(FcArgs a) {
  constexpr int TB = FcCfg<M>::TB;
  constexpr int NT = FcCfg<M>::NT;
  using LY = Layout<M, TB, false>;
  using SQ = Seq<M, false>;
  __shared__ float2 lds[LY::LDS_ELEMS];
  __shared__ int64_t base_in[TB], base_out[TB];

  const int64_t t0 = (int64_t)blockIdx.x * TB;
  for (int t = threadIdx.x; t < TB; t += NT) {
    int64_t oi = 0, oo = 0;
    if (t0 + t < a.rows) batch_offsets(a.bi, t0 + t, &oi, &oo);
    base_in[t] = oi;
    base_out[t] = oo;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < TB * M; e += NT) {
    const int t = e / M, i = e % M;
    float2 v = make_float2(0.f, 0.f);
    if (t0 + t < a.rows && i < a.xlen) v = ld_cf(a.x, a.x_c, base_in[t] + (int64_t)i * a.xs);
    lds[LY::addr(t, i)] = v;
  }
  __syncthreads();
  lds_passes<LY, SQ, M, NT, TB, false, 0, SQ::n>(lds, a.tw);
  for (int e = threadIdx.x; e < TB * M; e += NT) {
    const int t = e / M, k = e % M;
    lds[LY::addr(t, k)] = cmul(lds[LY::addr(t, k)], a.spec[k]);
  }
  __syncthreads();
  lds_passes<LY, SQ, M, NT, TB, true, 0, SQ::n>(lds, a.tw);
  const int olen = (int)a.olen, o0 = (int)a.o0;
  for (int e = threadIdx.x; e < TB * olen; e += NT) {
    const int t = e / olen, k = e % olen;
    if (t0 + t >= a.rows) continue;
    const float2 v = lds[LY::addr(t, o0 + k)];
    const int64_t o = base_out[t] + (int64_t)k * a.os;
    if (a.out_c) static_cast<float2 *>(a.out)[o] = v;
    else static_cast<float *>(a.out)[o] = v.x;
  }
}

__global__ void rev_conj_kernel(float2 *g, const float2 *h, int64_t L, int64_t hs) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= L) return;
  const float2 v = h[(L - 1 - i) * hs];
  g[i] = make_float2(v.x, -v.y);
}

// Largest transform of the fused path (one workgroup per row).
constexpr int64_t FCONV_MAX_M = 16384;

template <int M>
static mxr_status launch_fconv(FcArgs a, hipStream_t s) {
  using C = FcCfg<M>;
  {
    ProfScope ps("fconv_spec_kernel", s);
    fconv_spec_kernel<M><<<dim3(1), dim3(C::NT), 0, s>>>(a);
  }
  MXR_HIP_TRY(hipGetLastError());
  const int64_t grid = (a.rows + C::TB - 1) / C::TB;
  if (grid == 0 || a.olen == 0) return MXR_SUCCESS;
  MXR_REQUIRE(grid <= 0x7fffffff, MXR_INVALID_SIZE, "conv1d_fft: too many rows");
  {
    ProfScope ps("fconv_kernel", s);
    fconv_kernel<M><<<dim3((unsigned)grid), dim3(C::NT), 0, s>>>(a);
  }
  MXR_HIP_TRY(hipGetLastError());
  return MXR_SUCCESS;
}

// Fused path: rows of `in` (first xlen samples) convolved with filt, FULL
// outputs [o0, o0 + olen) written to `out`'s rows.
static mxr_status fconv_fused(const mxr_tensor *out, const mxr_tensor *in, int64_t xlen,
                              const mxr_tensor *filt, int corr, int64_t o0, int64_t olen,
                              hipStream_t s) {
  const int r = in->rank;
  const int64_t L = filt->shape[0];
  int64_t M = 8;
  while (M < xlen + L - 1) M <<= 1;
  MXR_REQUIRE(M <= FCONV_MAX_M && o0 + olen <= xlen + L - 1, MXR_INVALID_SIZE,
              "conv1d_fft: internal window error");
  FcArgs a{};
  a.x = in->data;
  a.x_c = in->dtype == MXR_C64;
  a.xlen = xlen;
  a.xs = in->stride[r - 1];
  a.h = filt->data;
  a.h_c = filt->dtype == MXR_C64;
  a.hlen = L;
  a.hs = filt->stride[0];
  a.corr = corr;
  a.bi.nd = 0;
  a.rows = 1;
  for (int d = 0; d < r - 1; ++d) {
    a.bi.shape[a.bi.nd] = in->shape[d];
    a.bi.sa[a.bi.nd] = in->stride[d];
    a.bi.sb[a.bi.nd] = out->stride[d];
    a.bi.nd++;
    a.rows *= in->shape[d];
  }
  a.out = out->data;
  a.out_c = out->dtype == MXR_C64;
  a.os = out->stride[r - 1];
  a.o0 = o0;
  a.olen = olen;
  a.spec = static_cast<float2 *>(stream_workspace("conv1d_fft_spec", sizeof(float2) * M, s));
  if (a.spec == nullptr) return workspace_failure("conv1d_fft");
  a.tw = twiddles_device((int)M, false);
  MXR_REQUIRE(a.tw, MXR_OUT_OF_MEMORY, "conv1d_fft: twiddle table allocation failed");
  switch (M) {
    case 8: return launch_fconv<8>(a, s);
    case 16: return launch_fconv<16>(a, s);
    case 32: return launch_fconv<32>(a, s);
    case 64: return launch_fconv<64>(a, s);
    case 128: return launch_fconv<128>(a, s);
    case 256: return launch_fconv<256>(a, s);
    case 512: return launch_fconv<512>(a, s);
    case 1024: return launch_fconv<1024>(a, s);
    case 2048: return launch_fconv<2048>(a, s);
    case 4096: return launch_fconv<4096>(a, s);
    case 8192: return launch_fconv<8192>(a, s);
    case 16384: return launch_fconv<16384>(a, s);
    default: break;
  }
  return MXR_NOT_SUPPORTED;
}

mxr_status conv1d_fft_impl(const mxr_tensor *out, const mxr_tensor *in, const mxr_tensor *filt,
                           int mode, int corr, hipStream_t s) {
  const char *fn = "conv1d_fft";
  mxr_status st;
  if ((st = check_tensor(in, "conv input", 0, 1, MXR_MAX_RANK))) return st;
  if ((st = check_tensor(out, "conv output", 0, 1, MXR_MAX_RANK))) return st;
  if ((st = check_tensor(filt, "conv filter", 0, 1, 1))) return st;
  for (const mxr_tensor *t : {in, out, filt})
    MXR_REQUIRE(t->dtype == MXR_C64 || t->dtype == MXR_F32, MXR_INVALID_TYPE,
                "%s: tensors must be c64 or f32", fn);
  MXR_REQUIRE(in->rank == out->rank, MXR_INVALID_DIM, "%s: input rank %d != output rank %d", fn,
              in->rank, out->rank);
  const bool want_c = in->dtype == MXR_C64 || filt->dtype == MXR_C64;
  MXR_REQUIRE((out->dtype == MXR_C64) == want_c, MXR_INVALID_TYPE,
              "%s: output type must be complex iff input or filter is", fn);
  MXR_REQUIRE(mode >= MXR_CONV_FULL && mode <= MXR_CONV_VALID, MXR_INVALID_PARAMETER,
              "%s: bad mode %d", fn, mode);
  MXR_REQUIRE(corr == 0 || corr == 1, MXR_INVALID_PARAMETER, "%s: corr must be 0 or 1", fn);
  const int r = in->rank;
  for (int d = 0; d < r - 1; ++d)
    MXR_REQUIRE(in->shape[d] == out->shape[d], MXR_INVALID_SIZE, "%s: batch dim %d differs", fn, d);
  const int64_t N = in->shape[r - 1], L = filt->shape[0];
  MXR_REQUIRE(N >= 1 && L >= 1, MXR_INVALID_SIZE, "%s: empty signal or filter", fn);
  MXR_REQUIRE(mode != MXR_CONV_VALID || N >= L, MXR_INVALID_SIZE,
              "%s: VALID needs a signal at least as long as the filter (%lld < %lld)", fn,
              (long long)N, (long long)L);
  const int64_t P = N + L - 1;
  const int64_t s0 = mode == MXR_CONV_FULL ? 0 : (mode == MXR_CONV_SAME ? ((L & 1) ? (L - 1) / 2 : L / 2 - 1) : L - 1);
  const int64_t olen = mode == MXR_CONV_FULL ? P : (mode == MXR_CONV_SAME ? N : N - L + 1);
  MXR_REQUIRE(out->shape[r - 1] == olen, MXR_INVALID_SIZE,
              "%s: output length %lld, expected %lld for this mode", fn,
              (long long)out->shape[r - 1], (long long)olen);
  MXR_REQUIRE(!views_overlap(*out, *in) && !views_overlap(*out, *filt), MXR_ALIAS_ERROR,
              "%s: output aliases an input", fn);
  if (numel(*out) == 0) return MXR_SUCCESS;
  if (P <= FCONV_MAX_M) return fconv_fused(out, in, N, filt, corr, s0, olen, s);

  // long rows: overlap-save correlation for the outputs at or after L-1,
  // the fused kernel on x[0, L-1) for the ones before
  MXR_REQUIRE(in->dtype == MXR_C64 && filt->dtype == MXR_C64, MXR_NOT_SUPPORTED,
              "%s: rows longer than %lld points (N + L - 1 = %lld) need c64 input and filter", fn,
              (long long)FCONV_MAX_M, (long long)P);
  MXR_REQUIRE(2 * L - 2 <= FCONV_MAX_M, MXR_NOT_SUPPORTED,
              "%s: filters longer than %lld taps need N + L - 1 <= %lld", fn,
              (long long)(FCONV_MAX_M / 2 + 1), (long long)FCONV_MAX_M);
  const int64_t head = std::max<int64_t>(0, L - 1 - s0);  // outputs before FULL index L-1
  if (head > 0) {
    mxr_tensor oh = *out;
    oh.shape[r - 1] = head;
    if ((st = fconv_fused(&oh, in, std::min(N, L - 1), filt, corr, s0, head, s))) return st;
  }
  mxr_tensor ot = *out;
  ot.shape[r - 1] = olen - head;
  ot.data = static_cast<float2 *>(out->data) + head * out->stride[r - 1];
  int64_t n = 1;
  while (n < P) n <<= 1;
  if (corr) return pc_common(fn, &ot, in, filt, n, 0, 0, 0, 1, s);
  // the correlation filter g = reverse(conj(h))
  float2 *g = static_cast<float2 *>(stream_workspace("conv1d_fft_rev", sizeof(float2) * L, s));
  if (g == nullptr) return workspace_failure(fn);
  rev_conj_kernel<<<dim3((unsigned)((L + 255) / 256)), dim3(256), 0, s>>>(
      g, static_cast<const float2 *>(filt->data), L, filt->stride[0]);
  MXR_HIP_TRY(hipGetLastError());
  mxr_tensor gt{};
  gt.data = g;
  gt.dtype = MXR_C64;
  gt.rank = 1;
  gt.shape[0] = L;
  gt.stride[0] = 1;
  return pc_common(fn, &ot, in, &gt, n, 0, 0, 0, 1, s);
}

}  // namespace mxr

extern "C" mxr_status mxr_conv1d_fft(const mxr_tensor *out, const mxr_tensor *in,
                                     const mxr_tensor *filt, int mode, int corr,
                                     hipStream_t stream) {
  return mxr::guarded("mxr_conv1d_fft",
                      [&] { return mxr::conv1d_fft_impl(out, in, filt, mode, corr, stream); });
}
