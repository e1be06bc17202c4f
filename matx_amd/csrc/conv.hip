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

// ------------------------------------------------ tiled real 2-D (CFAR)
// Real map, real filter (the CA-CFAR window sum conv2d(xPow, mask), and the
// cell counts conv2d(ones, mask)): a workgroup computes a 32 x 64 output
// tile of one batch index.  It stages the input rows oy0 - dy ..
// oy0 + 31 + fh - 1 - dy and columns ox0 - dx .. ox0 + 63 + fw - 1 - dx
// (zero outside the map) in LDS once; thread (tx, ty) then accumulates the
// 8 outputs of rows 8 ty .. 8 ty + 7 of column tx.  The filter height FH is
// a template parameter so that, per filter column, each of the 8 + FH - 1
// staged rows the thread touches is read from LDS once and feeds every
// (output row, filter row) pair it belongs to: 19.5 LDS reads per output for
// the 5 x 13 CFAR mask instead of one per tap (65), which leaves the kernel
// bound by its FMAs.  Tap weights are uniform and come through the scalar
// cache.
constexpr int C2_TH = 32, C2_TW = 64, C2_FHMAX = 16, C2_FWMAX = 64;
constexpr int C2_LW = C2_TW + C2_FWMAX;  // staged row pitch (floats)

struct Conv2TileArgs {
  const float *in;
  float *out;
  const float *f;      // filter, flipped on read: f[(fh-1-fy) fs0 + (fw-1-fx) fs1]
  int64_t fs0, fs1;
  int fh, fw, dy, dx;
  int H, W, Ho, Wo;
  int64_t is0, is1, os0, os1;  // strides of the conv dims
  int nb;                      // batch dims (rank - 2)
  int64_t bshape[CONV_MAX_RANK], bis[CONV_MAX_RANK], bos[CONV_MAX_RANK];
  int tiles_x;
};

template <int FH>
__global__ __launch_bounds__(256) void conv2d_tile_kernel(Conv2TileArgs a) {
  __shared__ float tile[(C2_TH + C2_FHMAX) * C2_LW];
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const int ox0 = (blockIdx.x % a.tiles_x) * C2_TW;
  const int oy0 = (blockIdx.x / a.tiles_x) * C2_TH;
  int64_t ib = 0, ob = 0, rem = blockIdx.y;
  for (int d = a.nb - 1; d >= 0; --d) {
    const int64_t i = rem % a.bshape[d];
    rem /= a.bshape[d];
    ib += i * a.bis[d];
    ob += i * a.bos[d];
  }
  const int rows = C2_TH + a.fh - 1, cols = C2_TW + a.fw - 1;
  const int iy0 = oy0 - a.dy, ix0 = ox0 - a.dx;
  // staged column by column: thread t loads column t % 128 of rows
  // t / 128, t / 128 + 2, ..., in chunks of 8 loads issued before any is
  // stored; rows outside the map read the column's first row and are zeroed
  // with a mask (a select against a load would become a branch around it
  // and serialise the loads); lanes past the staged width stay idle
  const int sc = threadIdx.x & (C2_LW - 1), sr0 = threadIdx.x / C2_LW;
  constexpr int RSTEP = 256 / C2_LW;
  const int ix = ix0 + sc;
  if (sc < cols && (a.H == 0 || a.W == 0)) {
    // empty map: nothing to load, the window sums are zero
    for (int r = sr0; r < rows; r += RSTEP) tile[r * C2_LW + sc] = 0.f;
  } else if (sc < cols) {
    const bool col_in = ix >= 0 && ix < a.W;
    const float *colp = a.in + ib + (col_in ? (int64_t)ix * a.is1 : 0);
    constexpr int NR = (C2_TH + C2_FHMAX + RSTEP - 1) / RSTEP;
#pragma unroll 1
    for (int k0 = 0; k0 < NR && sr0 + k0 * RSTEP < rows; k0 += 8) {
      float vv[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int iy = iy0 + sr0 + (k0 + j) * RSTEP;
        const bool in = col_in && iy >= 0 && iy < a.H;
        const float v = colp[in ? (int64_t)iy * a.is0 : 0];
        vv[j] = __int_as_float(__float_as_int(v) & (in ? -1 : 0));
      }
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int r = sr0 + (k0 + j) * RSTEP;
        if (r < rows) tile[r * C2_LW + sc] = vv[j];
      }
    }
  }
  __syncthreads();
  float acc[8];
#pragma unroll
  for (int r = 0; r < 8; ++r) acc[r] = 0.f;
  const float *tb = tile + (8 * ty) * C2_LW + tx;
  // out[r] += in[r + fy][fx] * f[FH-1-fy][fw-1-fx]: staged row ry = r + fy
  for (int fx = 0; fx < a.fw; ++fx) {
    float w[FH];
#pragma unroll
    for (int fy = 0; fy < FH; ++fy) w[fy] = a.f[(FH - 1 - fy) * a.fs0 + (a.fw - 1 - fx) * a.fs1];
    float v[8 + FH - 1];
#pragma unroll
    for (int ry = 0; ry < 8 + FH - 1; ++ry) v[ry] = tb[ry * C2_LW + fx];
#pragma unroll
    for (int ry = 0; ry < 8 + FH - 1; ++ry)
#pragma unroll
      for (int fy = 0; fy < FH; ++fy)
        if (ry - fy >= 0 && ry - fy < 8) acc[ry - fy] = fmaf(v[ry], w[fy], acc[ry - fy]);
  }
  const int ox = ox0 + tx;
  if (ox >= a.Wo) return;
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    const int oy = oy0 + 8 * ty + r;
    if (oy < a.Ho) a.out[ob + oy * a.os0 + ox * a.os1] = acc[r];
  }
}

// one instantiation per filter height 1 .. C2_FHMAX
template <int FH>
static void launch_conv2d_tile(const Conv2TileArgs &t, dim3 grid, hipStream_t s) {
  if (t.fh == FH) conv2d_tile_kernel<FH><<<grid, dim3(256), 0, s>>>(t);
  else if constexpr (FH < C2_FHMAX) launch_conv2d_tile<FH + 1>(t, grid, s);
}

// --------------------------------------------- 1-D along a strided axis
// conv1d along the last view dim when that dim is strided and a batch dim
// has unit stride in both views (the three-pulse canceller on permuted
// views, conv1d(inputView.Permute({0,2,1}), [1,-2,1], SAME)): lanes run
// along the unit-stride batch dim, so every load and store is a
// contiguous run; each thread computes a strip of 16 outputs along the conv
// axis (its taps' inputs are re-read from the L1).  The remaining batch
// dims come from blockIdx.z.
constexpr int C1_STRIP = 16;

struct Conv1StripArgs {
  const void *in;
  void *out;
  const void *f;
  int in_c, out_c, f_c;
  int64_t fl, fs, off;     // taps, filter stride, start offset (index rule of conv1d)
  int64_t n_in, n_out;     // conv-axis lengths
  int64_t is_c, os_c;      // conv-axis strides
  int64_t n_lane;          // extent of the unit-stride batch dim
  int nb;                  // other batch dims
  int64_t bshape[CONV_MAX_RANK], bis[CONV_MAX_RANK], bos[CONV_MAX_RANK];
};

__global__ __launch_bounds__(256) void conv1d_strip_kernel(Conv1StripArgs a) {
  const int64_t l = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t o0 = (int64_t)blockIdx.y * C1_STRIP;
  int64_t ib = 0, ob = 0, rem = blockIdx.z;
  for (int d = a.nb - 1; d >= 0; --d) {
    const int64_t i = rem % a.bshape[d];
    rem /= a.bshape[d];
    ib += i * a.bis[d];
    ob += i * a.bos[d];
  }
  if (l >= a.n_lane) return;
  ib += l;
  ob += l;
  for (int k = 0; k < C1_STRIP; ++k) {
    const int64_t o = o0 + k;
    if (o >= a.n_out) break;
    float2 acc = make_float2(0.f, 0.f);
    const int64_t full = o + a.off;
    for (int64_t t = 0; t < a.fl; ++t) {
      const int64_t si = full - t;
      if (si < 0 || si >= a.n_in) continue;
      const float2 x = ld_c(a.in, ib + si * a.is_c, a.in_c);
      const float2 w = ld_c(a.f, t * a.fs, a.f_c);
      acc = make_float2(acc.x + (x.x * w.x - x.y * w.y), acc.y + (x.x * w.y + x.y * w.x));
    }
    if (a.out_c) reinterpret_cast<float2 *>(a.out)[ob + o * a.os_c] = acc;
    else reinterpret_cast<float *>(a.out)[ob + o * a.os_c] = acc.x;
  }
}

// Short filters (fl <= 16), the canceller's case: the strip's whole input
// window (16 + fl - 1 samples) is loaded into registers first, every load
// issued before any arithmetic (positions outside the input read its first
// sample and are zeroed with a mask), then each output is summed from
// registers.  FLM bounds fl at compile time; taps u >= fl are skipped by a
// uniform branch.
template <int FLM, bool IC, bool FC>
__global__ __launch_bounds__(256) void conv1d_window_kernel(Conv1StripArgs a) {
  constexpr int NW = C1_STRIP + FLM - 1;
  const int64_t l = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t o0 = (int64_t)blockIdx.y * C1_STRIP;
  int64_t ib = 0, ob = 0, rem = blockIdx.z;
  for (int d = a.nb - 1; d >= 0; --d) {
    const int64_t i = rem % a.bshape[d];
    rem /= a.bshape[d];
    ib += i * a.bis[d];
    ob += i * a.bos[d];
  }
  if (l >= a.n_lane) return;
  ib += l;
  ob += l;
  const int fl = (int)a.fl;
  // window sample j is input position s0 + j; output o0 + k takes tap t
  // from window sample k + (fl - 1 - t)
  const int64_t s0 = o0 + a.off - (fl - 1);
  float2 xv[NW];
  bool ok[NW];
#pragma unroll
  for (int j = 0; j < NW; ++j) {
    const int64_t si = s0 + j;
    ok[j] = j < C1_STRIP + fl - 1 && si >= 0 && si < a.n_in;
    const int64_t at = ib + (ok[j] ? si : 0) * a.is_c;
    if constexpr (IC) xv[j] = reinterpret_cast<const float2 *>(a.in)[at];
    else xv[j] = make_float2(reinterpret_cast<const float *>(a.in)[at], 0.f);
  }
#pragma unroll
  for (int j = 0; j < NW; ++j) xv[j] = keep_if(xv[j], ok[j]);
  float2 w[FLM];
#pragma unroll
  for (int u = 0; u < FLM; ++u) {
    // window offset u carries tap t = fl - 1 - u
    const int64_t t = u < fl ? fl - 1 - u : 0;
    if constexpr (FC) w[u] = reinterpret_cast<const float2 *>(a.f)[t * a.fs];
    else w[u] = make_float2(reinterpret_cast<const float *>(a.f)[t * a.fs], 0.f);
  }
#pragma unroll
  for (int k = 0; k < C1_STRIP; ++k) {
    const int64_t o = o0 + k;
    if (o >= a.n_out) break;
    float2 acc = make_float2(0.f, 0.f);
#pragma unroll
    for (int u = 0; u < FLM; ++u) {
      if (u < fl) {
        const float2 x = xv[k + u];
        if constexpr (FC) acc = make_float2(acc.x + (x.x * w[u].x - x.y * w[u].y), acc.y + (x.x * w[u].y + x.y * w[u].x));
        else acc = make_float2(acc.x + x.x * w[u].x, acc.y + x.y * w[u].x);
      }
    }
    if (a.out_c) reinterpret_cast<float2 *>(a.out)[ob + o * a.os_c] = acc;
    else reinterpret_cast<float *>(a.out)[ob + o * a.os_c] = acc.x;
  }
}

template <int FLM>
static void launch_conv1d_window(const Conv1StripArgs &c, dim3 grid, hipStream_t s) {
  if (c.in_c && c.f_c) conv1d_window_kernel<FLM, true, true><<<grid, dim3(256), 0, s>>>(c);
  else if (c.in_c) conv1d_window_kernel<FLM, true, false><<<grid, dim3(256), 0, s>>>(c);
  else if (c.f_c) conv1d_window_kernel<FLM, false, true><<<grid, dim3(256), 0, s>>>(c);
  else conv1d_window_kernel<FLM, false, false><<<grid, dim3(256), 0, s>>>(c);
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
  // real 2-D maps with a window that fits the staged tile
  if (two_d && !a.in_c && !a.f_c && !a.out_c && a.fl0 <= C2_FHMAX && a.fl1 <= C2_FWMAX &&
      out->shape[r - 2] < (1 << 30) && out->shape[r - 1] < (1 << 30) &&
      in->shape[r - 2] < (1 << 30) && in->shape[r - 1] < (1 << 30)) {
    Conv2TileArgs t{};
    t.in = (const float *)in->data;
    t.out = (float *)out->data;
    t.f = (const float *)filt->data;
    t.fs0 = filt->stride[0];
    t.fs1 = filt->stride[1];
    t.fh = (int)a.fl0;
    t.fw = (int)a.fl1;
    t.dy = (int)a.off0;
    t.dx = (int)a.off1;
    t.H = (int)in->shape[r - 2];
    t.W = (int)in->shape[r - 1];
    t.Ho = (int)out->shape[r - 2];
    t.Wo = (int)out->shape[r - 1];
    t.is0 = in->stride[r - 2];
    t.is1 = in->stride[r - 1];
    t.os0 = out->stride[r - 2];
    t.os1 = out->stride[r - 1];
    t.nb = r - 2;
    int64_t nbatch = 1;
    for (int d = 0; d < r - 2; ++d) {
      t.bshape[d] = out->shape[d];
      t.bis[d] = in->stride[d];
      t.bos[d] = out->stride[d];
      nbatch *= out->shape[d];
    }
    t.tiles_x = (t.Wo + C2_TW - 1) / C2_TW;
    const int64_t tiles = (int64_t)t.tiles_x * ((t.Ho + C2_TH - 1) / C2_TH);
    MXR_REQUIRE(tiles <= 0x7fffffff && nbatch <= 65535, MXR_INVALID_SIZE, "%s: too large", fn);
    ProfScope ps("conv2d_tile_kernel", stream);
    launch_conv2d_tile<1>(t, dim3((unsigned)tiles, (unsigned)nbatch), stream);
    MXR_HIP_TRY(hipGetLastError());
    return MXR_SUCCESS;
  }
  // 1-D along a strided axis with a unit-stride batch dim
  if (!two_d && in->stride[r - 1] != 1 && out->stride[r - 1] != 1) {
    int lane = -1;
    for (int d = 0; d < r - 1; ++d)
      if (in->stride[d] == 1 && out->stride[d] == 1 && out->shape[d] > 1) lane = d;
    int64_t nbatch = 1;
    for (int d = 0; d < r - 1; ++d)
      if (d != lane) nbatch *= out->shape[d];
    const int64_t strips = (out->shape[r - 1] + C1_STRIP - 1) / C1_STRIP;
    if (lane >= 0 && nbatch <= 65535 && strips <= 65535) {
      Conv1StripArgs c{};
      c.in = in->data;
      c.out = out->data;
      c.f = filt->data;
      c.in_c = a.in_c;
      c.out_c = a.out_c;
      c.f_c = a.f_c;
      c.fl = a.fl1;
      c.fs = a.fs1;
      c.off = a.off1;
      c.n_in = in->shape[r - 1];
      c.n_out = out->shape[r - 1];
      c.is_c = in->stride[r - 1];
      c.os_c = out->stride[r - 1];
      c.n_lane = out->shape[lane];
      for (int d = 0; d < r - 1; ++d) {
        if (d == lane) continue;
        c.bshape[c.nb] = out->shape[d];
        c.bis[c.nb] = in->stride[d];
        c.bos[c.nb] = out->stride[d];
        ++c.nb;
      }
      const int64_t lanes = (c.n_lane + 255) / 256;
      MXR_REQUIRE(lanes <= 0x7fffffff, MXR_INVALID_SIZE, "%s: too large", fn);
      const dim3 grid((unsigned)lanes, (unsigned)strips, (unsigned)nbatch);
      if (c.fl <= 16) {
        ProfScope ps("conv1d_window_kernel", stream);
        if (c.fl <= 4) launch_conv1d_window<4>(c, grid, stream);
        else if (c.fl <= 8) launch_conv1d_window<8>(c, grid, stream);
        else launch_conv1d_window<16>(c, grid, stream);
      } else {
        ProfScope ps("conv1d_strip_kernel", stream);
        conv1d_strip_kernel<<<grid, dim3(256), 0, stream>>>(c);
      }
      MXR_HIP_TRY(hipGetLastError());
      return MXR_SUCCESS;
    }
  }
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
