// Direct 1-D and 2-D convolution over strided views (replaces conv1d_impl /
// conv2d_impl with MATX_C_METHOD_DIRECT, include/matx/transforms/conv.h:
// 402-532 and 547-595).  One thread per output element, the filter staged in
// LDS; the output dimension with the smallest stride varies fastest across
// lanes so strided (permuted) views still read and write coalesced rows.
// Index rules: 1-D SAME start (L odd ? (L-1)/2 : L/2-1), VALID start L-1
// (include/matx/kernels/conv.cuh:155-168); 2-D out(oy,ox) =
// sum in(oy+fy-dy, ox+fx-dx) * f(fh-1-fy, fw-1-fx) with dy = fh-1 (FULL),
// fh/2 (SAME), 0 (VALID) (include/matx/transforms/conv.h:340-380).
#include <algorithm>

#include "mxr_internal.hpp"
#include "mxr_cache.hpp"

namespace mxr {

constexpr int CONV_MAX_RANK = 6;
constexpr int CONV_MAX_FILT = 4096;

struct ConvArgs {
  const void *in;
  void *out;
  const void *f;
  int in_c, out_c, f_c;  // complex flags
  int rank;              // of in/out
  int64_t oshape[CONV_MAX_RANK];
  int64_t ostr[CONV_MAX_RANK];
  int64_t istr[CONV_MAX_RANK];
  int order[CONV_MAX_RANK];  // dims from slowest to fastest for thread decomposition
  int64_t total;
  // 1-D: conv along dim rank-1; 2-D: along dims rank-2, rank-1
  int64_t in_n0, in_n1;     // input extents of the conv dims (n1 = last)
  int64_t fl0, fl1;         // filter extents
  int64_t off0, off1;       // dy, dx (2-D) or start (1-D, in off1)
  int64_t fs0, fs1;         // filter strides
  int two_d;
};

__device__ __forceinline__ float2 ld_c(const void *p, int64_t o, int is_c) {
  if (is_c) return reinterpret_cast<const float2 *>(p)[o];
  return make_float2(reinterpret_cast<const float *>(p)[o], 0.f);
}

__global__ __launch_bounds__(256) void conv_kernel(ConvArgs a) {
  __shared__ float2 sf[CONV_MAX_FILT];
  const int64_t nf = a.fl0 * a.fl1;
  for (int64_t i = threadIdx.x; i < nf; i += blockDim.x) {
    const int64_t f0 = i / a.fl1, f1 = i % a.fl1;
    sf[i] = ld_c(a.f, f0 * a.fs0 + f1 * a.fs1, a.f_c);
  }
  __syncthreads();
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= a.total) return;
  int64_t idx[CONV_MAX_RANK];
  int64_t rem = e;
  for (int k = a.rank - 1; k >= 0; --k) {
    const int d = a.order[k];
    idx[d] = rem % a.oshape[d];
    rem /= a.oshape[d];
  }
  int64_t obase = 0, ibase = 0;
  const int nconv = a.two_d ? 2 : 1;
  for (int d = 0; d < a.rank - nconv; ++d) {
    obase += idx[d] * a.ostr[d];
    ibase += idx[d] * a.istr[d];
  }
  const int d1 = a.rank - 1;
  float2 acc = make_float2(0.f, 0.f);
  if (!a.two_d) {
    obase += idx[d1] * a.ostr[d1];
    const int64_t full = idx[d1] + a.off1;
    for (int64_t t = 0; t < a.fl1; ++t) {
      const int64_t si = full - t;
      if (si < 0 || si >= a.in_n1) continue;
      const float2 x = ld_c(a.in, ibase + si * a.istr[d1], a.in_c);
      acc = make_float2(acc.x + (x.x * sf[t].x - x.y * sf[t].y), acc.y + (x.x * sf[t].y + x.y * sf[t].x));
    }
  } else {
    const int d0 = a.rank - 2;
    obase += idx[d0] * a.ostr[d0] + idx[d1] * a.ostr[d1];
    for (int64_t fy = 0; fy < a.fl0; ++fy) {
      const int64_t iy = idx[d0] + fy - a.off0;
      if (iy < 0 || iy >= a.in_n0) continue;
      for (int64_t fx = 0; fx < a.fl1; ++fx) {
        const int64_t ix = idx[d1] + fx - a.off1;
        if (ix < 0 || ix >= a.in_n1) continue;
        const float2 x = ld_c(a.in, ibase + iy * a.istr[d0] + ix * a.istr[d1], a.in_c);
        const float2 w = sf[(a.fl0 - 1 - fy) * a.fl1 + (a.fl1 - 1 - fx)];
        acc = make_float2(acc.x + (x.x * w.x - x.y * w.y), acc.y + (x.x * w.y + x.y * w.x));
      }
    }
  }
  if (a.out_c) reinterpret_cast<float2 *>(a.out)[obase] = acc;
  else reinterpret_cast<float *>(a.out)[obase] = acc.x;
}

static mxr_status conv_common(const char *fn, const mxr_tensor *out, const mxr_tensor *in,
                              const mxr_tensor *filt, int mode, bool two_d, hipStream_t stream) {
  mxr_status st;
  const int nconv = two_d ? 2 : 1;
  if ((st = check_tensor(in, "conv input", 0, nconv, CONV_MAX_RANK))) return st;
  if ((st = check_tensor(out, "conv output", 0, nconv, CONV_MAX_RANK))) return st;
  if ((st = check_tensor(filt, "conv filter", 0, nconv, nconv))) return st;
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
  const int r = in->rank;
  for (int d = 0; d < r - nconv; ++d)
    MXR_REQUIRE(in->shape[d] == out->shape[d], MXR_INVALID_SIZE, "%s: batch dim %d differs", fn, d);
  MXR_REQUIRE(numel(*filt) <= CONV_MAX_FILT, MXR_NOT_SUPPORTED,
              "%s: filter of %lld taps exceeds the direct method limit %d", fn,
              (long long)numel(*filt), CONV_MAX_FILT);
  MXR_REQUIRE(!views_overlap(*out, *in) && !views_overlap(*out, *filt), MXR_ALIAS_ERROR,
              "%s: output aliases an input", fn);
  ConvArgs a{};
  a.in = in->data;
  a.out = out->data;
  a.f = filt->data;
  a.in_c = in->dtype == MXR_C64;
  a.out_c = out->dtype == MXR_C64;
  a.f_c = filt->dtype == MXR_C64;
  a.rank = r;
  a.two_d = two_d;
  a.total = numel(*out);
  for (int d = 0; d < r; ++d) {
    a.oshape[d] = out->shape[d];
    a.ostr[d] = out->stride[d];
    a.istr[d] = in->stride[d];
  }
  auto expect = [&](int64_t n, int64_t L) -> int64_t {
    return mode == MXR_CONV_FULL ? n + L - 1 : (mode == MXR_CONV_SAME ? n : n - L + 1);
  };
  if (!two_d) {
    const int64_t n = in->shape[r - 1], L = filt->shape[0];
    MXR_REQUIRE(out->shape[r - 1] == expect(n, L), MXR_INVALID_SIZE,
                "%s: output length %lld, expected %lld for this mode", fn,
                (long long)out->shape[r - 1], (long long)expect(n, L));
    a.in_n0 = 1;
    a.in_n1 = n;
    a.fl0 = 1;
    a.fl1 = L;
    a.fs0 = 0;
    a.fs1 = filt->stride[0];
    a.off1 = mode == MXR_CONV_FULL ? 0 : (mode == MXR_CONV_SAME ? ((L & 1) ? (L - 1) / 2 : L / 2 - 1) : L - 1);
  } else {
    const int64_t H = in->shape[r - 2], W = in->shape[r - 1];
    const int64_t fh = filt->shape[0], fw = filt->shape[1];
    MXR_REQUIRE(out->shape[r - 2] == expect(H, fh) && out->shape[r - 1] == expect(W, fw),
                MXR_INVALID_SIZE, "%s: output shape does not match the mode", fn);
    a.in_n0 = H;
    a.in_n1 = W;
    a.fl0 = fh;
    a.fl1 = fw;
    a.fs0 = filt->stride[0];
    a.fs1 = filt->stride[1];
    a.off0 = mode == MXR_CONV_FULL ? fh - 1 : (mode == MXR_CONV_SAME ? fh / 2 : 0);
    a.off1 = mode == MXR_CONV_FULL ? fw - 1 : (mode == MXR_CONV_SAME ? fw / 2 : 0);
  }
  // thread order: the output dim with the smallest |stride| is fastest
  int ord[CONV_MAX_RANK];
  for (int d = 0; d < r; ++d) ord[d] = d;
  std::stable_sort(ord, ord + r, [&](int x, int y) {
    return std::llabs(out->stride[x]) > std::llabs(out->stride[y]);
  });
  for (int d = 0; d < r; ++d) a.order[d] = ord[d];
  if (a.total == 0) return MXR_SUCCESS;
  const int64_t blocks = (a.total + 255) / 256;
  MXR_REQUIRE(blocks <= 0x7fffffff, MXR_INVALID_SIZE, "%s: too large", fn);
  {
    ProfScope ps(two_d ? "conv2d_kernel" : "conv1d_kernel", stream);
    conv_kernel<<<dim3((unsigned)blocks), dim3(256), 0, stream>>>(a);
  }
  MXR_HIP_TRY(hipGetLastError());
  return MXR_SUCCESS;
}

}  // namespace mxr

extern "C" {

mxr_status mxr_conv1d(const mxr_tensor *out, const mxr_tensor *in, const mxr_tensor *filt, int mode,
                      hipStream_t stream) {
  return mxr::guarded("mxr_conv1d",
                      [&] { return mxr::conv_common("conv1d", out, in, filt, mode, false, stream); });
}

mxr_status mxr_conv2d(const mxr_tensor *out, const mxr_tensor *in, const mxr_tensor *filt, int mode,
                      hipStream_t stream) {
  return mxr::guarded("mxr_conv2d",
                      [&] { return mxr::conv_common("conv2d", out, in, filt, mode, true, stream); });
}

}  // extern "C"
