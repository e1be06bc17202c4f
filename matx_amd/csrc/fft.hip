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
#include <type_traits>
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

// fft_gather for loops that keep many loads in flight: every load is issued
// unconditionally (positions past the input's extent, and ok = false, read
// element 0 of the row and give zero), so a thread's loads carry no control
// flow between them and issue back to back.  KIND is a.kind, dispatched
// once outside the loop by with_kind.  Pre: a.in_len >= 1.
template <int KIND>
__device__ __forceinline__ float2 fft_load(const FftArgs &a, int64_t base, int64_t i, bool ok) {
  // the zero fill and the conjugation are bit operations on the loaded
  // words, not selects: a select between a load and a constant is turned
  // into a branch around the load, which serialises the loads again
  auto keep = [](float x, int mask) { return __int_as_float(__float_as_int(x) & mask); };
  if constexpr (KIND == FFT_C2R) {
    const bool mirror = i > a.n / 2;
    const int64_t h = mirror ? a.n - i : i;
    const bool in = ok && h < a.in_len;
    const int mk = in ? -1 : 0;
    const float2 v = a.in[base + (in ? h : 0) * a.istride];
    const float y = __int_as_float(__float_as_int(v.y) ^ (mirror ? (int)0x80000000 : 0));
    return make_float2(keep(v.x, mk), keep(y, mk));
  } else if constexpr (KIND == FFT_R2C) {
    const bool in = ok && i < a.in_len;
    const float r = reinterpret_cast<const float *>(a.in)[base + (in ? i : 0) * a.istride];
    return make_float2(keep(r, in ? -1 : 0), 0.f);
  } else {
    const bool in = ok && i < a.in_len;
    const int mk = in ? -1 : 0;
    const float2 v = a.in[base + (in ? i : 0) * a.istride];
    return make_float2(keep(v.x, mk), keep(v.y, mk));
  }
}

template <class F>
__device__ __forceinline__ void with_kind(int kind, F &&body) {
  if (kind == FFT_C2R) body(std::integral_constant<int, FFT_C2R>{});
  else if (kind == FFT_R2C) body(std::integral_constant<int, FFT_R2C>{});
  else body(std::integral_constant<int, FFT_C2C>{});
}

template <int N>
struct GenericCfg {
  static constexpr int TB = N >= 4096 ? 1 : 4096 / N;
  static constexpr int NT = N >= 4096 ? N / 16 : 256;
};

// COL: element i of transform t at i * TB + t (batch_fast calls, where the
// workgroup's transforms are neighbours in memory and the gather and
// scatter run lanes across them: consecutive lanes touch consecutive LDS
// words, where the transform-major layout puts them a padded row apart).
template <int N, bool INV, bool COL>
__global__ __launch_bounds__(GenericCfg<N>::NT) void fft_generic_kernel(FftArgs a) {
  constexpr int TB = GenericCfg<N>::TB;
  constexpr int NT = GenericCfg<N>::NT;
  using LY = Layout<N, TB, COL>;
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

  // gather: each thread's PER elements are loaded before any is stored, so
  // its loads are in flight together
  constexpr int PER = TB * N / NT;
  static_assert(PER * NT == TB * N, "whole elements per thread");
  {
    float2 gv[PER];
#pragma unroll
    for (int m = 0; m < PER; ++m) gv[m] = make_float2(0.f, 0.f);
    if (a.in_len > 0)
      with_kind(a.kind, [&](auto kt) {
        constexpr int KIND = decltype(kt)::value;
#pragma unroll
        for (int m = 0; m < PER; ++m) {
          const int e = threadIdx.x + m * NT;
          const int t = a.batch_fast ? e % TB : e / N;
          const int i = a.batch_fast ? e / TB : e % N;
          gv[m] = fft_load<KIND>(a, base_in[t], i, t0 + t < a.nbatch);
        }
      });
#pragma unroll
    for (int m = 0; m < PER; ++m) {
      const int e = threadIdx.x + m * NT;
      const int t = a.batch_fast ? e % TB : e / N;
      const int i = a.batch_fast ? e / TB : e % N;
      lds[LY::addr(t, i)] = gv[m];
    }
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
  if constexpr (C::TB > 1) {
    if (a.batch_fast) {
      if (inv) fft_generic_kernel<N, true, true><<<dim3((unsigned)grid), dim3(C::NT), 0, s>>>(a);
      else fft_generic_kernel<N, false, true><<<dim3((unsigned)grid), dim3(C::NT), 0, s>>>(a);
      MXR_HIP_TRY(hipGetLastError());
      return MXR_SUCCESS;
    }
  }
  if (inv) fft_generic_kernel<N, true, false><<<dim3((unsigned)grid), dim3(C::NT), 0, s>>>(a);
  else fft_generic_kernel<N, false, false><<<dim3((unsigned)grid), dim3(C::NT), 0, s>>>(a);
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
  // (each thread's loads are issued together, in chunks of CH elements,
  // before the values are combined and stored)
  constexpr int PER = TB * M / NT;
  constexpr int CH = PER < 8 ? PER : 8;
  static_assert(PER * NT == TB * M && PER % CH == 0, "whole chunks per thread");
#pragma unroll 1
  for (int m0 = 0; m0 < PER; m0 += CH) {
    float2 xv[CH], cv[CH];
#pragma unroll
    for (int j = 0; j < CH; ++j) xv[j] = make_float2(0.f, 0.f);
    if (a.in_len > 0)
      with_kind(a.kind, [&](auto kt) {
        constexpr int KIND = decltype(kt)::value;
#pragma unroll
        for (int j = 0; j < CH; ++j) {
          const int e = threadIdx.x + (m0 + j) * NT;
          const int t = a.batch_fast ? e % TB : e / M;
          const int i = a.batch_fast ? e / TB : e % M;
          xv[j] = fft_load<KIND>(a, base_in[t], i, t0 + t < a.nbatch && i < n);
        }
      });
#pragma unroll
    for (int j = 0; j < CH; ++j) {
      const int e = threadIdx.x + (m0 + j) * NT;
      const int i = a.batch_fast ? e / TB : e % M;
      cv[j] = z.chirp[i < n ? i : 0];
    }
#pragma unroll
    for (int j = 0; j < CH; ++j) {
      const int e = threadIdx.x + (m0 + j) * NT;
      const int t = a.batch_fast ? e % TB : e / M;
      const int i = a.batch_fast ? e / TB : e % M;
      float2 x = xv[j];
      if (INV) x.y = -x.y;
      lds[LY::addr(t, i)] = cmul(x, cv[j]);
    }
  }
  __syncthreads();
  lds_passes<LY, SQ, M, NT, TB, false, 0, SQ::n>(lds, a.tw);
#pragma unroll 1
  for (int m0 = 0; m0 < PER; m0 += CH) {
    float2 sv[CH];
#pragma unroll
    for (int j = 0; j < CH; ++j) sv[j] = z.spec[(threadIdx.x + (m0 + j) * NT) % M];
#pragma unroll
    for (int j = 0; j < CH; ++j) {
      const int e = threadIdx.x + (m0 + j) * NT;
      const int t = e / M, k = e % M;
      lds[LY::addr(t, k)] = cmul(lds[LY::addr(t, k)], sv[j]);
    }
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

// ------------------------------------------- long transforms: two passes
// n = N1 * N2 > 16384 (powers of two up to 2^24) as two passes of short
// transforms over the n-point sequence viewed as a matrix (the "four-step"
// factorisation with the transposes folded into the passes' addressing):
//   input index  i = n1 + N1 n2,  output index k = N2 k1 + k2,
//   X[N2 k1 + k2] = sum_n1 W_N1^(n1 k1) W_n^(n1 k2) sum_n2 x[n1 + N1 n2] W_N2^(n2 k2).
// Pass 1: for each n1, an N2-point transform over n2 (input stride N1),
//         times W_n^(n1 k2), stored row n1 of a work matrix [N1][N2].
// Pass 2: for each k2, an N1-point transform over n1 (work stride N2),
//         stored at output index N2 k1 + k2 with the norm scale.
// A workgroup owns TB neighbouring columns (n1 in pass 1, k2 in pass 2):
// pass 1 gathers TB consecutive input samples per n2 and stores whole work
// rows; pass 2 gathers TB consecutive work elements per n1 and stores TB
// consecutive outputs per k1, so every HBM access is a contiguous run.
// W_n^m (m = n1 k2 < n) is the product of two table entries,
// lo[m mod 2^s] * hi[m >> s], each built in double and rounded once.
//
// Sizes that are not powers of two above the single-workgroup Bluestein
// range use the same identity as czt_kernel with M = next_pow2(2n - 1)
// points through these two passes: pass 1 of the forward gathers x.c (zero
// beyond n), pass 2 of the forward stores the product with the chirp
// spectrum, pass 2 of the inverse stores c.y.
enum BigMode {
  BIG_VIEW_TO_WORK = 0,   // pass 1 of a plain transform
  BIG_WORK_TO_VIEW = 1,   // pass 2 of a plain transform
  BIG_CHIRP_TO_WORK = 2,  // Bluestein: pass 1 of the forward, gathering x.c
  BIG_WORK_SPEC = 3,      // Bluestein: pass 2 of the forward, times the spectrum
  BIG_WORK_TO_WORK = 4,   // Bluestein: pass 1 of the inverse
  BIG_WORK_CHIRP = 5,     // Bluestein: pass 2 of the inverse, c.y to the view
};

struct BigArgs {
  FftArgs f;               // the caller's views (gather: pass 1, scatter: pass 2)
  const float2 *src;       // work matrix read by pass 2, or by a Bluestein inverse pass 1
  float2 *dst;             // work matrix written by pass 1, or the Bluestein spectrum product
  int64_t n;               // N1 * N2
  int64_t N1, N2;
  const float2 *tw_lo, *tw_hi;  // W_n^m = tw_lo[m & lo_mask] * tw_hi[m >> lo_bits]
  int lo_bits;
  const float2 *tw_pass;   // Stockham table of the short transform
  int64_t czt_n;           // Bluestein: the user's transform size
  const float2 *chirp, *spec;
  int conj_io;             // Bluestein inverse DFT: conjugate in and out
  int64_t cols;            // columns per transform (N1 in pass 1, N2 in pass 2)
};

template <int N>
struct BigCfg {
  static constexpr int TB = (8192 / N) > 64 ? 64 : ((8192 / N) < 2 ? 2 : 8192 / N);
  static constexpr int NT = 256;
};

template <int N, bool INV, int MODE>
__global__ __launch_bounds__(256) void fft_big_kernel(BigArgs a) {
  constexpr int TB = BigCfg<N>::TB;
  constexpr int NT = BigCfg<N>::NT;
  constexpr bool PASS1 = MODE == BIG_VIEW_TO_WORK || MODE == BIG_CHIRP_TO_WORK || MODE == BIG_WORK_TO_WORK;
  using LY = Layout<N, TB, true>;
  using SQ = Seq<N, false>;
  __shared__ float2 lds[LY::LDS_ELEMS];
  const FftArgs &f = a.f;
  const int64_t nblk = a.cols / TB;
  const int64_t b = blockIdx.x / nblk;           // transform (batch) index
  const int64_t c0 = (blockIdx.x % nblk) * TB;   // first column of this workgroup
  int64_t base_in = 0, base_out = 0;
  batch_offsets(f.bi, b, &base_in, &base_out);
  const int64_t wbase = b * a.n;

  // gather: element i of column c0 + t, lanes over t (consecutive
  // addresses); each thread's loads are issued in chunks of CH before the
  // values are stored
  constexpr int PER = TB * N / NT;
  constexpr int CH = PER < 8 ? PER : 8;
  static_assert(PER * NT == TB * N && PER % CH == 0, "whole chunks per thread");
#pragma unroll 1
  for (int m0 = 0; m0 < PER; m0 += CH) {
    float2 gv[CH];
    if constexpr (MODE == BIG_VIEW_TO_WORK || MODE == BIG_CHIRP_TO_WORK) {
      float2 cv[CH];
#pragma unroll
      for (int j = 0; j < CH; ++j) gv[j] = make_float2(0.f, 0.f);
      if (f.in_len > 0)
        with_kind(f.kind, [&](auto kt) {
          constexpr int KIND = decltype(kt)::value;
#pragma unroll
          for (int j = 0; j < CH; ++j) {
            const int e = threadIdx.x + (m0 + j) * NT;
            const int64_t g = c0 + e % TB + a.N1 * (e / TB);
            gv[j] = fft_load<KIND>(f, base_in, g, MODE == BIG_VIEW_TO_WORK || g < a.czt_n);
          }
        });
      if constexpr (MODE == BIG_CHIRP_TO_WORK) {
#pragma unroll
        for (int j = 0; j < CH; ++j) {
          const int e = threadIdx.x + (m0 + j) * NT;
          const int64_t g = c0 + e % TB + a.N1 * (e / TB);
          cv[j] = a.chirp[g < a.czt_n ? g : 0];
        }
#pragma unroll
        for (int j = 0; j < CH; ++j) {
          float2 x = gv[j];
          if (a.conj_io) x.y = -x.y;
          gv[j] = cmul(x, cv[j]);
        }
      }
    } else {
#pragma unroll
      for (int j = 0; j < CH; ++j) {
        const int e = threadIdx.x + (m0 + j) * NT;
        const int64_t c = c0 + e % TB;
        const int i = e / TB;
        if constexpr (MODE == BIG_WORK_TO_WORK) gv[j] = a.src[wbase + c + a.N1 * i];
        else gv[j] = a.src[wbase + i * a.N2 + c];  // pass 2: row i, column c of [N1][N2]
      }
    }
#pragma unroll
    for (int j = 0; j < CH; ++j) {
      const int e = threadIdx.x + (m0 + j) * NT;
      lds[LY::addr(e % TB, e / TB)] = gv[j];
    }
  }
  __syncthreads();
  lds_passes<LY, SQ, N, NT, TB, INV, 0, SQ::n>(lds, a.tw_pass);

  if constexpr (PASS1) {
    // row c of the work matrix, times W_n^(c k); lanes over k (the table
    // entries of a chunk are loaded before its products are stored)
  #pragma unroll 1
  for (int m0 = 0; m0 < PER; m0 += CH) {
      float2 lo[CH], hi[CH];
#pragma unroll
      for (int j = 0; j < CH; ++j) {
        const int e = threadIdx.x + (m0 + j) * NT;
        const int64_t m = (c0 + e / N) * (e % N);
        lo[j] = a.tw_lo[m & ((1LL << a.lo_bits) - 1)];
        hi[j] = a.tw_hi[m >> a.lo_bits];
      }
#pragma unroll
      for (int j = 0; j < CH; ++j) {
        const int e = threadIdx.x + (m0 + j) * NT;
        const int t = e / N, k = e % N;
        float2 w = cmul(lo[j], hi[j]);
        if (INV) w.y = -w.y;
        a.dst[wbase + (c0 + t) * a.N2 + k] = cmul(lds[LY::addr(t, k)], w);
      }
    }
  } else {
    // output index N2 k + c; lanes over t (a chunk's spectrum or chirp
    // entries are loaded before its outputs are stored)
  #pragma unroll 1
  for (int m0 = 0; m0 < PER; m0 += CH) {
      float2 fac[CH];
      if constexpr (MODE == BIG_WORK_SPEC || MODE == BIG_WORK_CHIRP) {
#pragma unroll
        for (int j = 0; j < CH; ++j) {
          const int e = threadIdx.x + (m0 + j) * NT;
          const int64_t kg = a.N2 * (e / TB) + c0 + e % TB;
          if constexpr (MODE == BIG_WORK_SPEC) fac[j] = a.spec[kg];
          else fac[j] = kg < f.out_len ? a.chirp[kg] : make_float2(0.f, 0.f);
        }
      }
#pragma unroll
      for (int j = 0; j < CH; ++j) {
        const int e = threadIdx.x + (m0 + j) * NT;
        const int t = e % TB, k = e / TB;
        const int64_t kg = a.N2 * k + c0 + t;
        float2 v = lds[LY::addr(t, k)];
        if constexpr (MODE == BIG_WORK_SPEC) {
          a.dst[wbase + kg] = cmul(v, fac[j]);
        } else if constexpr (MODE == BIG_WORK_CHIRP) {
          if (kg < f.out_len) {
            float2 y = cscale(cmul(v, fac[j]), f.scale);
            if (a.conj_io) y.y = -y.y;
            fft_scatter(f, base_out, kg, y);
          }
        } else {
          if (kg < f.out_len) fft_scatter(f, base_out, kg, cscale(v, f.scale));
        }
      }
    }
  }
}

namespace {
std::mutex g_big_mu;
std::map<std::pair<int, int64_t>, float2 *> g_big_tw;  // (device, n) -> [lo | hi]

void clear_big_tw() {
  std::lock_guard<std::mutex> g(g_big_mu);
  int cur = 0;
  (void)hipGetDevice(&cur);
  for (auto &kv : g_big_tw) {
    (void)hipSetDevice(kv.first.first);
    (void)hipFree(kv.second);
  }
  (void)hipSetDevice(cur);
  g_big_tw.clear();
}
CacheClearRegistrar g_reg_big(clear_big_tw);

int ilog2(int64_t v) {
  int r = 0;
  while ((1LL << r) < v) ++r;
  return r;
}
}  // namespace

// W_n^m split tables for m < n: lo[j] = W^j (j < 2^s), hi[j] = W^(j 2^s)
// (j < n / 2^s), s = ceil(log2(n) / 2), exponents reduced in integers.
static const float2 *big_twiddles(int64_t n, int *lo_bits) {
  const int lb = (ilog2(n) + 1) / 2;
  *lo_bits = lb;
  const auto key = std::make_pair(device_id(), n);
  std::lock_guard<std::mutex> g(g_big_mu);
  auto it = g_big_tw.find(key);
  if (it != g_big_tw.end()) return it->second;
  const int64_t nlo = 1LL << lb, nhi = n >> lb;
  std::vector<float2> host((size_t)(nlo + nhi));
  const double two_pi = 6.283185307179586476925286766559;
  auto w = [&](int64_t m) {
    const double ang = -two_pi * (double)(m % n) / (double)n;
    return make_float2((float)std::cos(ang), (float)std::sin(ang));
  };
  for (int64_t j = 0; j < nlo; ++j) host[(size_t)j] = w(j);
  for (int64_t j = 0; j < nhi; ++j) host[(size_t)(nlo + j)] = w(j << lb);
  float2 *d = nullptr;
  if (hipMalloc(&d, sizeof(float2) * host.size()) != hipSuccess) return nullptr;
  if (hipMemcpy(d, host.data(), sizeof(float2) * host.size(), hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(d);
    return nullptr;
  }
  g_big_tw[key] = d;
  return d;
}

// Largest transform of the two-pass path, and its split: N1 <= N2, both
// powers of two of at most 4096 points.
constexpr int64_t BIG_MAX_N = 1LL << 24;

template <int N, bool INV, int MODE>
static mxr_status launch_big_one(const BigArgs &a, hipStream_t s) {
  constexpr int TB = BigCfg<N>::TB;
  BigArgs b = a;
  b.tw_pass = twiddles_device(N, false);
  MXR_REQUIRE(b.tw_pass != nullptr, MXR_OUT_OF_MEMORY, "fft: twiddle table allocation failed");
  const int64_t grid = a.f.nbatch * (b.cols / TB);
  MXR_REQUIRE(grid <= 0x7fffffff, MXR_INVALID_SIZE, "fft: batch too large");
  ProfScope ps("fft_big_kernel", s);
  fft_big_kernel<N, INV, MODE><<<dim3((unsigned)grid), dim3(256), 0, s>>>(b);
  MXR_HIP_TRY(hipGetLastError());
  return MXR_SUCCESS;
}

template <bool INV, int MODE>
static mxr_status launch_big(int64_t N, const BigArgs &a, hipStream_t s) {
  switch (N) {
    case 8: return launch_big_one<8, INV, MODE>(a, s);
    case 16: return launch_big_one<16, INV, MODE>(a, s);
    case 32: return launch_big_one<32, INV, MODE>(a, s);
    case 64: return launch_big_one<64, INV, MODE>(a, s);
    case 128: return launch_big_one<128, INV, MODE>(a, s);
    case 256: return launch_big_one<256, INV, MODE>(a, s);
    case 512: return launch_big_one<512, INV, MODE>(a, s);
    case 1024: return launch_big_one<1024, INV, MODE>(a, s);
    case 2048: return launch_big_one<2048, INV, MODE>(a, s);
    case 4096: return launch_big_one<4096, INV, MODE>(a, s);
    default: break;
  }
  set_error(MXR_NOT_SUPPORTED, "fft: no two-pass kernel for a %lld-point pass", (long long)N);
  return MXR_NOT_SUPPORTED;
}

// Sets n, N1, N2 and the twiddle tables of an m-point two-pass transform.
static mxr_status big_setup(BigArgs &b, int64_t m) {
  const int lg = ilog2(m);
  b.n = m;
  b.N1 = 1LL << (lg / 2);
  b.N2 = m / b.N1;
  const float2 *t = big_twiddles(m, &b.lo_bits);
  MXR_REQUIRE(t != nullptr, MXR_OUT_OF_MEMORY, "fft: twiddle table allocation failed");
  b.tw_lo = t;
  b.tw_hi = t + (1LL << b.lo_bits);
  return MXR_SUCCESS;
}

// A plain transform of power-of-two n in (16384, 2^24]: two passes through a
// per-stream work matrix of nbatch * n elements.
static mxr_status dispatch_big(int64_t n, const FftArgs &f, bool inv, hipStream_t s) {
  BigArgs b{};
  b.f = f;
  mxr_status st = big_setup(b, n);
  if (st) return st;
  float2 *work = (float2 *)stream_workspace("fft_big", sizeof(float2) * (size_t)(f.nbatch * n), s);
  if (!work) return workspace_failure("fft");
  b.dst = work;
  b.src = work;
  b.cols = b.N1;
  st = inv ? launch_big<true, BIG_VIEW_TO_WORK>(b.N2, b, s) : launch_big<false, BIG_VIEW_TO_WORK>(b.N2, b, s);
  if (st) return st;
  b.cols = b.N2;
  return inv ? launch_big<true, BIG_WORK_TO_VIEW>(b.N1, b, s) : launch_big<false, BIG_WORK_TO_VIEW>(b.N1, b, s);
}

// Chirp and spectrum tables of an n-point Bluestein transform with an
// M-point convolution (czt_tables), for M above the one-workgroup range.
static const float2 *czt_tables(int n, int M);

// Any size n in (8192, 2^23] that is not a power of two: Bluestein through
// the two-pass M-point transforms, M = next_pow2(2n - 1) <= 2^24.
static mxr_status dispatch_czt_big(int64_t n, const FftArgs &f, bool inv, hipStream_t s) {
  int64_t M = 1;
  while (M < 2 * n - 1) M <<= 1;
  BigArgs b{};
  b.f = f;
  mxr_status st = big_setup(b, M);
  if (st) return st;
  const float2 *tab = czt_tables((int)n, (int)M);
  MXR_REQUIRE(tab != nullptr, MXR_OUT_OF_MEMORY, "fft: table allocation failed for size %lld",
              (long long)n);
  b.czt_n = n;
  b.chirp = tab;
  b.spec = tab + n;
  b.conj_io = inv ? 1 : 0;
  const size_t bytes = sizeof(float2) * (size_t)(f.nbatch * M);
  float2 *w1 = (float2 *)stream_workspace("fft_big", bytes, s);
  float2 *w2 = w1 ? (float2 *)stream_workspace("fft_big_czt", bytes, s) : nullptr;
  if (!w1 || !w2) return workspace_failure("fft");
  // forward M-point transform of x.c, times the chirp spectrum (1/M in it)
  b.dst = w1;
  b.cols = b.N1;
  if ((st = launch_big<false, BIG_CHIRP_TO_WORK>(b.N2, b, s))) return st;
  b.src = w1;
  b.dst = w2;
  b.cols = b.N2;
  if ((st = launch_big<false, BIG_WORK_SPEC>(b.N1, b, s))) return st;
  // inverse M-point transform, then c.y for the first out_len outputs
  b.src = w2;
  b.dst = w1;
  b.cols = b.N1;
  if ((st = launch_big<true, BIG_WORK_TO_WORK>(b.N2, b, s))) return st;
  b.src = w1;
  b.cols = b.N2;
  return launch_big<true, BIG_WORK_CHIRP>(b.N1, b, s);
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
  MXR_REQUIRE((pow2 && n <= BIG_MAX_N) || (n >= 1 && n <= BIG_MAX_N / 2), MXR_NOT_SUPPORTED,
              "%s: size %lld unsupported (powers of two up to %lld, other sizes up to %lld)", fn,
              (long long)n, (long long)BIG_MAX_N, (long long)(BIG_MAX_N / 2));
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
  if (!pow2) return n <= CZT_MAX_N ? dispatch_czt(n, a, inv, stream) : dispatch_czt_big(n, a, inv, stream);
  if (n > 16384) return dispatch_big(n, a, inv, stream);
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
