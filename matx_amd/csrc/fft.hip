// Generic batched 1-D C2C FFT over arbitrary strided views
// (replaces fft_impl/ifft_impl, include/matx/transforms/fft/fft_cuda.h:680-774).
//
// One workgroup owns TB whole transforms in LDS: a gather phase reads them
// with the view's strides (zero-padding or truncating to n, the size rule of
// include/matx/transforms/fft/fft_common.h:39-140), the Stockham passes run
// LDS->LDS, and a scatter phase writes the output view with the norm scale
// folded in (fft_cuda.h:126-188 does it as a separate pass).  When the
// innermost batch dim has unit stride (an FFT along a strided axis, e.g. the
// pulse axis of a permuted view) the gather/scatter walk that dim fastest so
// HBM accesses stay coalesced.
#include <cmath>

#include "fft_engine.hpp"
#include "mxr_cache.hpp"

namespace mxr {

struct FftArgs {
  const float2 *in;
  float2 *out;
  int64_t nbatch;
  int64_t in_len;  // elements of the input axis used (min(extent, n))
  int64_t istride, ostride;
  BatchIdx bi;     // batch dims: sa = input strides, sb = output strides
  int batch_fast;  // gather/scatter with lanes over the batch index
  float scale;
  const float2 *tw;
};

template <int N>
struct GenericCfg {
  static constexpr int TB = N >= 4096 ? 1 : 4096 / N;
  static constexpr int NT = N >= 4096 ? N / 16 : 256;
};

template <int N, bool INV>
__global__ __launch_bounds__(GenericCfg<N>::NT) void fft_generic_kernel(FftArgs a) {
  constexpr int TB = GenericCfg<N>::TB;
  constexpr int NT = GenericCfg<N>::NT;
  using LY = Layout<N, TB, false>;
  using SQ = Seq<N, false>;
  __shared__ float2 lds[LY::LDS_ELEMS];
  __shared__ int64_t base_in[TB], base_out[TB];

  const int64_t t0 = (int64_t)blockIdx.x * TB;
  for (int t = threadIdx.x; t < TB; t += NT) {
    int64_t oi = 0, oo = 0;
    if (t0 + t < a.nbatch) batch_offsets(a.bi, t0 + t, &oi, &oo);
    base_in[t] = oi;
    base_out[t] = oo;
  }
  __syncthreads();

  // gather
  for (int e = threadIdx.x; e < TB * N; e += NT) {
    int t, i;
    if (a.batch_fast) {
      t = e % TB;
      i = e / TB;
    } else {
      t = e / N;
      i = e % N;
    }
    float2 v = make_float2(0.f, 0.f);
    if (t0 + t < a.nbatch && i < a.in_len) v = a.in[base_in[t] + (int64_t)i * a.istride];
    lds[LY::addr(t, i)] = v;
  }
  __syncthreads();

  lds_passes<LY, SQ, N, NT, TB, INV, 0, SQ::n>(lds, a.tw);

  // scatter
  for (int e = threadIdx.x; e < TB * N; e += NT) {
    int t, i;
    if (a.batch_fast) {
      t = e % TB;
      i = e / TB;
    } else {
      t = e / N;
      i = e % N;
    }
    if (t0 + t < a.nbatch) a.out[base_out[t] + (int64_t)i * a.ostride] = cscale(lds[LY::addr(t, i)], a.scale);
  }
}

// n == 1: the transform is the identity (times the scale)
__global__ void fft_n1_kernel(FftArgs a) {
  int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= a.nbatch) return;
  int64_t oi, oo;
  batch_offsets(a.bi, b, &oi, &oo);
  float2 v = a.in_len > 0 ? a.in[oi] : make_float2(0.f, 0.f);
  a.out[oo] = cscale(v, a.scale);
}

template <int N>
static mxr_status launch_generic(const FftArgs &a, bool inv, hipStream_t s) {
  using C = GenericCfg<N>;
  const int64_t grid = (a.nbatch + C::TB - 1) / C::TB;
  if (grid == 0) return MXR_SUCCESS;
  MXR_REQUIRE(grid <= 0x7fffffff, MXR_INVALID_SIZE, "fft: batch too large");
  ProfScope ps("fft_generic_kernel", s);
  if (inv) fft_generic_kernel<N, true><<<dim3((unsigned)grid), dim3(C::NT), 0, s>>>(a);
  else fft_generic_kernel<N, false><<<dim3((unsigned)grid), dim3(C::NT), 0, s>>>(a);
  MXR_HIP_TRY(hipGetLastError());
  return MXR_SUCCESS;
}

static mxr_status dispatch_generic(int64_t n, const FftArgs &a, bool inv, hipStream_t s) {
  switch (n) {
    case 2: return launch_generic<2>(a, inv, s);
    case 4: return launch_generic<4>(a, inv, s);
    case 8: return launch_generic<8>(a, inv, s);
    case 16: return launch_generic<16>(a, inv, s);
    case 32: return launch_generic<32>(a, inv, s);
    case 64: return launch_generic<64>(a, inv, s);
    case 128: return launch_generic<128>(a, inv, s);
    case 256: return launch_generic<256>(a, inv, s);
    case 512: return launch_generic<512>(a, inv, s);
    case 1024: return launch_generic<1024>(a, inv, s);
    case 2048: return launch_generic<2048>(a, inv, s);
    case 4096: return launch_generic<4096>(a, inv, s);
    case 8192: return launch_generic<8192>(a, inv, s);
    case 16384: return launch_generic<16384>(a, inv, s);
    default: break;
  }
  set_error(MXR_NOT_SUPPORTED, "fft: size %lld is not a power of two in [1, 16384]", (long long)n);
  return MXR_NOT_SUPPORTED;
}

mxr_status fft_c2c_impl(const mxr_tensor *out, const mxr_tensor *in, int64_t n, int dir, int norm,
                        int axis, hipStream_t stream) {
  mxr_status st;
  if ((st = check_tensor(in, "fft input", MXR_C64, 1, MXR_MAX_RANK))) return st;
  if ((st = check_tensor(out, "fft output", MXR_C64, 1, MXR_MAX_RANK))) return st;
  MXR_REQUIRE(in->rank == out->rank, MXR_INVALID_DIM, "fft: input rank %d != output rank %d",
              in->rank, out->rank);
  const int rank = in->rank;
  if (axis < 0) axis += rank;
  MXR_REQUIRE(axis >= 0 && axis < rank, MXR_INVALID_DIM, "fft: axis out of range");
  MXR_REQUIRE(dir == MXR_FFT_FORWARD || dir == MXR_FFT_INVERSE, MXR_INVALID_PARAMETER,
              "fft: bad direction %d", dir);
  MXR_REQUIRE(norm >= MXR_NORM_BACKWARD && norm <= MXR_NORM_ORTHO, MXR_INVALID_PARAMETER,
              "fft: bad norm %d", norm);
  if (n == 0) n = in->shape[axis];
  // MatX never pads implicitly: the output extent along the axis is the
  // transform size (include/matx/operators/fft.h:111-199, FFT.cu:661-719).
  MXR_REQUIRE(out->shape[axis] == n, MXR_INVALID_SIZE,
              "fft: output size %lld along axis %d does not match transform size %lld",
              (long long)out->shape[axis], axis, (long long)n);
  for (int d = 0; d < rank; ++d) {
    if (d == axis) continue;
    MXR_REQUIRE(in->shape[d] == out->shape[d], MXR_INVALID_SIZE,
                "fft: batch dim %d differs (%lld vs %lld)", d, (long long)in->shape[d],
                (long long)out->shape[d]);
  }
  MXR_REQUIRE(n >= 1 && n <= 16384 && (n & (n - 1)) == 0, MXR_NOT_SUPPORTED,
              "fft: size %lld is not a power of two in [1, 16384]", (long long)n);
  if (in->data != out->data && views_overlap(*in, *out)) {
    set_error(MXR_ALIAS_ERROR, "fft: input and output overlap without being the same view");
    return MXR_ALIAS_ERROR;
  }

  FftArgs a{};
  a.in = (const float2 *)in->data;
  a.out = (float2 *)out->data;
  a.in_len = in->shape[axis] < n ? in->shape[axis] : n;
  a.istride = in->stride[axis];
  a.ostride = out->stride[axis];
  a.bi.nd = 0;
  a.nbatch = 1;
  int inner = -1;
  for (int d = 0; d < rank; ++d) {
    if (d == axis) continue;
    a.bi.shape[a.bi.nd] = in->shape[d];
    a.bi.sa[a.bi.nd] = in->stride[d];
    a.bi.sb[a.bi.nd] = out->stride[d];
    a.bi.nd++;
    a.nbatch *= in->shape[d];
    inner = d;
  }
  if (a.nbatch == 0) return MXR_SUCCESS;
  // batch-fast gather when the innermost batch dim is contiguous and the
  // transform axis is not
  a.batch_fast = (inner >= 0 && in->stride[inner] == 1 && a.istride != 1) ? 1 : 0;
  const double dn = (double)n;
  const bool inv = dir == MXR_FFT_INVERSE;
  if (norm == MXR_NORM_ORTHO) a.scale = (float)(1.0 / std::sqrt(dn));
  else if ((norm == MXR_NORM_BACKWARD) == inv) a.scale = (float)(1.0 / dn);
  else a.scale = 1.0f;
  if (n == 1) {
    const int64_t blocks = (a.nbatch + 255) / 256;
    fft_n1_kernel<<<dim3((unsigned)blocks), dim3(256), 0, stream>>>(a);
    MXR_HIP_TRY(hipGetLastError());
    return MXR_SUCCESS;
  }
  a.tw = twiddles_device((int)n, false);
  MXR_REQUIRE(a.tw != nullptr, MXR_OUT_OF_MEMORY, "fft: twiddle table allocation failed");
  return dispatch_generic(n, a, inv, stream);
}

}  // namespace mxr

extern "C" mxr_status mxr_fft_c2c(const mxr_tensor *out, const mxr_tensor *in, int64_t n, int dir,
                                  int norm, int axis, hipStream_t stream) {
  return mxr::guarded("mxr_fft_c2c",
                      [&] { return mxr::fft_c2c_impl(out, in, n, dir, norm, axis, stream); });
}
