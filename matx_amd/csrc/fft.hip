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
#include <complex>
#include <map>
#include <mutex>
#include <utility>
#include <vector>

#include "fft_engine.hpp"
#include "mxr_cache.hpp"

namespace mxr {

// Real-data forms (rfft / irfft, include/matx/operators/fft.h:580-740): R2C
// reads f32 samples (imaginary part zero) and stores bins 0..n/2; C2R reads
// bins 0..n/2, extends them as X[n-k] = conj(X[k]) and stores the real part
// (the imaginary parts of bins 0 and n/2 drop out, as in cuFFT and numpy).
enum FftKind { FFT_C2C = 0, FFT_R2C = 1, FFT_C2R = 2 };

struct FftArgs {
  const float2 *in;  // f32 samples for R2C
  float2 *out;       // f32 samples for C2R
  int64_t nbatch;
  int kind;
  int64_t n;
  int64_t out_len; // outputs stored per transform (n, or n/2+1 for R2C)
  int64_t in_len;  // elements of the input axis used (min(extent, n); n/2+1 for C2R)
  int64_t istride, ostride;
  BatchIdx bi;     // batch dims: sa = input strides, sb = output strides
  int batch_fast;  // gather/scatter with lanes over the batch index
  float scale;
  const float2 *tw;
};

__device__ __forceinline__ float2 fft_gather(const FftArgs &a, int64_t base, int64_t i) {
  if (a.kind == FFT_C2R) {
    const bool mirror = i > a.n / 2;
    const int64_t h = mirror ? a.n - i : i;
    if (h >= a.in_len) return make_float2(0.f, 0.f);
    float2 v = a.in[base + h * a.istride];
    if (mirror) v.y = -v.y;
    return v;
  }
  if (i >= a.in_len) return make_float2(0.f, 0.f);
  if (a.kind == FFT_R2C)
    return make_float2(reinterpret_cast<const float *>(a.in)[base + i * a.istride], 0.f);
  return a.in[base + i * a.istride];
}

__device__ __forceinline__ void fft_scatter(const FftArgs &a, int64_t base, int64_t k, float2 v) {
  if (a.kind == FFT_C2R) reinterpret_cast<float *>(a.out)[base + k * a.ostride] = v.x;
  else a.out[base + k * a.ostride] = v;
}

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
    if (t0 + t < a.nbatch) v = fft_gather(a, base_in[t], i);
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
    if (t0 + t < a.nbatch && i < a.out_len) fft_scatter(a, base_out[t], i, cscale(lds[LY::addr(t, i)], a.scale));
  }
}

// n == 1: the transform is the identity (times the scale)
__global__ void fft_n1_kernel(FftArgs a) {
  int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= a.nbatch) return;
  int64_t oi, oo;
  batch_offsets(a.bi, b, &oi, &oo);
  fft_scatter(a, oo, 0, cscale(fft_gather(a, oi, 0), a.scale));
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

// ------------------------------------------------- other sizes: Bluestein
// An n-point DFT of any n as a circular convolution of power-of-two length
// M >= 2n - 1 (Bluestein's chirp-z identity nk = (n^2 + k^2 - (k - n)^2)/2):
//   X[k] = c[k] * sum_j (x[j] c[j]) h[k - j],  c[j] = exp(-i pi j^2 / n),
//   h[m] = conj(c[|m|]),
// with the convolution done as IFFT_M(FFT_M(x.c) . H), H = FFT_M(h_ext)/M
// and h_ext[m] = h[m] (m < n), h_ext[M - m] = h[m] (0 < m < n), zero between.
// The inverse DFT is conj(DFT(conj(x))): same tables, conjugated in and out.
// One workgroup holds TB whole M-point sequences in LDS; gather, forward
// passes, spectrum multiply, inverse passes and scatter all stay there.
struct CztArgs {
  FftArgs f;
  int n;               // transform size
  const float2 *chirp; // c[0..n)
  const float2 *spec;  // H[0..M), scaled by 1/M
};

template <int M, bool INV>
__global__ __launch_bounds__(GenericCfg<M>::NT) void czt_kernel(CztArgs z) {
  constexpr int TB = GenericCfg<M>::TB;
  constexpr int NT = GenericCfg<M>::NT;
  using LY = Layout<M, TB, false>;
  using SQ = Seq<M, false>;
  __shared__ float2 lds[LY::LDS_ELEMS];
  __shared__ int64_t base_in[TB], base_out[TB];
  const FftArgs &a = z.f;
  const int n = z.n;

  const int64_t t0 = (int64_t)blockIdx.x * TB;
  for (int t = threadIdx.x; t < TB; t += NT) {
    int64_t oi = 0, oo = 0;
    if (t0 + t < a.nbatch) batch_offsets(a.bi, t0 + t, &oi, &oo);
    base_in[t] = oi;
    base_out[t] = oo;
  }
  __syncthreads();

  // gather x.c into the first n slots (x conjugated for the inverse; C2R
  // input extended to n first), zeros to M
  for (int e = threadIdx.x; e < TB * M; e += NT) {
    int t, i;
    if (a.batch_fast) {
      t = e % TB;
      i = e / TB;
    } else {
      t = e / M;
      i = e % M;
    }
    float2 v = make_float2(0.f, 0.f);
    if (t0 + t < a.nbatch && i < n) {
      float2 x = fft_gather(a, base_in[t], i);
      if (INV) x.y = -x.y;
      v = cmul(x, z.chirp[i]);
    }
    lds[LY::addr(t, i)] = v;
  }
  __syncthreads();
  lds_passes<LY, SQ, M, NT, TB, false, 0, SQ::n>(lds, a.tw);
  for (int e = threadIdx.x; e < TB * M; e += NT) {
    const int t = e / M, k = e % M;
    lds[LY::addr(t, k)] = cmul(lds[LY::addr(t, k)], z.spec[k]);
  }
  __syncthreads();
  lds_passes<LY, SQ, M, NT, TB, true, 0, SQ::n>(lds, a.tw);

  // scatter c.y for k < out_len (<= n) with the norm scale (conjugated for
  // the inverse)
  const int nout = (int)a.out_len;
  for (int e = threadIdx.x; e < TB * nout; e += NT) {
    int t, k;
    if (a.batch_fast) {
      t = e % TB;
      k = e / TB;
    } else {
      t = e / nout;
      k = e % nout;
    }
    if (t0 + t >= a.nbatch) continue;
    float2 y = cscale(cmul(lds[LY::addr(t, k)], z.chirp[k]), a.scale);
    if (INV) y.y = -y.y;
    fft_scatter(a, base_out[t], k, y);
  }
}

namespace {
// In-place radix-2 DFT in double, exp(-2 pi i jk / m), m a power of two.
void dft_pow2_double(std::vector<std::complex<double>> &v) {
  const size_t m = v.size();
  for (size_t i = 1, j = 0; i < m; ++i) {
    size_t bit = m >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) std::swap(v[i], v[j]);
  }
  const double two_pi = 6.283185307179586476925286766559;
  for (size_t len = 2; len <= m; len <<= 1) {
    for (size_t i = 0; i < m; i += len) {
      for (size_t k = 0; k < len / 2; ++k) {
        const double ang = -two_pi * (double)k / (double)len;
        const std::complex<double> w(std::cos(ang), std::sin(ang));
        const std::complex<double> u = v[i + k], t = w * v[i + k + len / 2];
        v[i + k] = u + t;
        v[i + k + len / 2] = u - t;
      }
    }
  }
}

std::mutex g_czt_mu;
std::map<std::pair<int, int>, float2 *> g_czt;  // (device, n) -> [chirp n | spec M]

void clear_czt() {
  std::lock_guard<std::mutex> g(g_czt_mu);
  for (auto &kv : g_czt) {
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(kv.first.first);
    (void)hipFree(kv.second);
    (void)hipSetDevice(cur);
  }
  g_czt.clear();
}
CacheClearRegistrar g_reg_czt(clear_czt);
}  // namespace

// Chirp and spectrum tables of an n-point transform (M-point convolution),
// built in double and rounded once; the chirp angle pi (j^2 mod 2n) / n is
// reduced in integers first.  Device pointer owned by the library.
static const float2 *czt_tables(int n, int M) {
  const auto key = std::make_pair(device_id(), n);
  std::lock_guard<std::mutex> g(g_czt_mu);
  auto it = g_czt.find(key);
  if (it != g_czt.end()) return it->second;
  const double pi = 3.14159265358979323846264338327950288;
  std::vector<std::complex<double>> c((size_t)n), h((size_t)M, 0.0);
  for (int64_t j = 0; j < n; ++j) {
    const int64_t e = (j * j) % (2 * (int64_t)n);
    const double ang = -pi * (double)e / (double)n;
    c[(size_t)j] = std::complex<double>(std::cos(ang), std::sin(ang));
  }
  h[0] = std::conj(c[0]);
  for (int j = 1; j < n; ++j) h[(size_t)j] = h[(size_t)(M - j)] = std::conj(c[(size_t)j]);
  dft_pow2_double(h);
  std::vector<float2> host((size_t)(n + M));
  for (int j = 0; j < n; ++j)
    host[(size_t)j] = make_float2((float)c[(size_t)j].real(), (float)c[(size_t)j].imag());
  for (int k = 0; k < M; ++k)
    host[(size_t)(n + k)] = make_float2((float)(h[(size_t)k].real() / M), (float)(h[(size_t)k].imag() / M));
  float2 *d = nullptr;
  if (hipMalloc(&d, sizeof(float2) * host.size()) != hipSuccess) return nullptr;
  // a synchronous copy: tables are built once per (device, size), before use
  if (hipMemcpy(d, host.data(), sizeof(float2) * host.size(), hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(d);
    return nullptr;
  }
  g_czt[key] = d;
  return d;
}

template <int M>
static mxr_status launch_czt(const CztArgs &z, bool inv, hipStream_t s) {
  using C = GenericCfg<M>;
  const int64_t grid = (z.f.nbatch + C::TB - 1) / C::TB;
  if (grid == 0) return MXR_SUCCESS;
  MXR_REQUIRE(grid <= 0x7fffffff, MXR_INVALID_SIZE, "fft: batch too large");
  ProfScope ps("czt_kernel", s);
  if (inv) czt_kernel<M, true><<<dim3((unsigned)grid), dim3(C::NT), 0, s>>>(z);
  else czt_kernel<M, false><<<dim3((unsigned)grid), dim3(C::NT), 0, s>>>(z);
  MXR_HIP_TRY(hipGetLastError());
  return MXR_SUCCESS;
}

// Largest size the Bluestein path takes (its convolution fits one 16384-point
// workgroup transform).
constexpr int64_t CZT_MAX_N = 8192;

static mxr_status dispatch_czt(int64_t n, FftArgs a, bool inv, hipStream_t s) {
  int M = 1;
  while (M < 2 * n - 1) M <<= 1;
  CztArgs z{};
  z.f = a;
  z.n = (int)n;
  z.f.tw = twiddles_device(M, false);
  const float2 *tab = czt_tables((int)n, M);
  MXR_REQUIRE(z.f.tw != nullptr && tab != nullptr, MXR_OUT_OF_MEMORY,
              "fft: table allocation failed for size %lld", (long long)n);
  z.chirp = tab;
  z.spec = tab + n;
  switch (M) {
    case 8: return launch_czt<8>(z, inv, s);
    case 16: return launch_czt<16>(z, inv, s);
    case 32: return launch_czt<32>(z, inv, s);
    case 64: return launch_czt<64>(z, inv, s);
    case 128: return launch_czt<128>(z, inv, s);
    case 256: return launch_czt<256>(z, inv, s);
    case 512: return launch_czt<512>(z, inv, s);
    case 1024: return launch_czt<1024>(z, inv, s);
    case 2048: return launch_czt<2048>(z, inv, s);
    case 4096: return launch_czt<4096>(z, inv, s);
    case 8192: return launch_czt<8192>(z, inv, s);
    case 16384: return launch_czt<16384>(z, inv, s);
    default: break;
  }
  set_error(MXR_NOT_SUPPORTED, "fft: no Bluestein kernel for size %lld", (long long)n);
  return MXR_NOT_SUPPORTED;
}

static mxr_status fft_impl(const char *fn, const mxr_tensor *out, const mxr_tensor *in, int64_t n,
                           int dir, int norm, int axis, int kind, hipStream_t stream) {
  mxr_status st;
  const int in_t = kind == FFT_R2C ? MXR_F32 : MXR_C64;
  const int out_t = kind == FFT_C2R ? MXR_F32 : MXR_C64;
  if ((st = check_tensor(in, "fft input", in_t, 1, MXR_MAX_RANK))) return st;
  if ((st = check_tensor(out, "fft output", out_t, 1, MXR_MAX_RANK))) return st;
  MXR_REQUIRE(in->rank == out->rank, MXR_INVALID_DIM, "%s: input rank %d != output rank %d", fn,
              in->rank, out->rank);
  const int rank = in->rank;
  if (axis < 0) axis += rank;
  MXR_REQUIRE(axis >= 0 && axis < rank, MXR_INVALID_DIM, "%s: axis out of range", fn);
  MXR_REQUIRE(dir == MXR_FFT_FORWARD || dir == MXR_FFT_INVERSE, MXR_INVALID_PARAMETER,
              "%s: bad direction %d", fn, dir);
  MXR_REQUIRE(norm >= MXR_NORM_BACKWARD && norm <= MXR_NORM_ORTHO, MXR_INVALID_PARAMETER,
              "%s: bad norm %d", fn, norm);
  // transform size: C2C and R2C default to the input length, C2R to
  // 2 (N - 1) for N input bins (include/matx/operators/fft.h:111-199)
  if (n == 0) n = kind == FFT_C2R ? 2 * (in->shape[axis] - 1) : in->shape[axis];
  // MatX never pads implicitly: the output extent along the axis is the
  // transform size, or n/2 + 1 bins for R2C (fft.h:111-199, FFT.cu:661-719).
  const int64_t out_len = kind == FFT_R2C ? n / 2 + 1 : n;
  MXR_REQUIRE(out->shape[axis] == out_len, MXR_INVALID_SIZE,
              "%s: output size %lld along axis %d does not match the %lld outputs of a "
              "%lld-point transform", fn, (long long)out->shape[axis], axis, (long long)out_len,
              (long long)n);
  for (int d = 0; d < rank; ++d) {
    if (d == axis) continue;
    MXR_REQUIRE(in->shape[d] == out->shape[d], MXR_INVALID_SIZE,
                "%s: batch dim %d differs (%lld vs %lld)", fn, d, (long long)in->shape[d],
                (long long)out->shape[d]);
  }
  const bool pow2 = n >= 1 && (n & (n - 1)) == 0;
  MXR_REQUIRE((pow2 && n <= 16384) || (n >= 1 && n <= CZT_MAX_N), MXR_NOT_SUPPORTED,
              "%s: size %lld unsupported (powers of two up to 16384, other sizes up to %lld)", fn,
              (long long)n, (long long)CZT_MAX_N);
  if (views_overlap(*in, *out)) {
    // a transform may overwrite its own input view (every element is read
    // before any is written), not a different view of the same memory
    const bool same = kind == FFT_C2C && in->data == out->data;
    MXR_REQUIRE(same, MXR_ALIAS_ERROR, "%s: input and output overlap%s", fn,
                kind == FFT_C2C ? " without being the same view" : "");
  }

  FftArgs a{};
  a.in = (const float2 *)in->data;
  a.out = (float2 *)out->data;
  a.kind = kind;
  a.n = n;
  a.out_len = out_len;
  const int64_t used = kind == FFT_C2R ? n / 2 + 1 : n;
  a.in_len = in->shape[axis] < used ? in->shape[axis] : used;
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
  if (!pow2) return dispatch_czt(n, a, inv, stream);
  a.tw = twiddles_device((int)n, false);
  MXR_REQUIRE(a.tw != nullptr, MXR_OUT_OF_MEMORY, "%s: twiddle table allocation failed", fn);
  return dispatch_generic(n, a, inv, stream);
}

mxr_status fft_c2c_impl(const mxr_tensor *out, const mxr_tensor *in, int64_t n, int dir, int norm,
                        int axis, hipStream_t stream) {
  return fft_impl("fft", out, in, n, dir, norm, axis, FFT_C2C, stream);
}

}  // namespace mxr

extern "C" mxr_status mxr_fft_r2c(const mxr_tensor *out, const mxr_tensor *in, int64_t n, int norm,
                                  int axis, hipStream_t stream) {
  return mxr::guarded("mxr_fft_r2c", [&] {
    return mxr::fft_impl("rfft", out, in, n, MXR_FFT_FORWARD, norm, axis, mxr::FFT_R2C, stream);
  });
}

extern "C" mxr_status mxr_fft_c2r(const mxr_tensor *out, const mxr_tensor *in, int64_t n, int norm,
                                  int axis, hipStream_t stream) {
  return mxr::guarded("mxr_fft_c2r", [&] {
    return mxr::fft_impl("irfft", out, in, n, MXR_FFT_INVERSE, norm, axis, mxr::FFT_C2R, stream);
  });
}

extern "C" mxr_status mxr_fft_c2c(const mxr_tensor *out, const mxr_tensor *in, int64_t n, int dir,
                                  int norm, int axis, hipStream_t stream) {
  return mxr::guarded("mxr_fft_c2c",
                      [&] { return mxr::fft_c2c_impl(out, in, n, dir, norm, axis, stream); });
}
