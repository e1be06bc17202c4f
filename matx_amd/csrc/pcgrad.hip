// Waveform gradient of pulse compression, summed over every pulse of a cube:
// the backward of y_r = corr(x_r, w2), w2 = normalise(window .* w)
// (examples/simple_radar_pipeline.h:276-300) with respect to w, for waveform
// design over a whole cube (no MatX counterpart: MatX has no autodiff).
//
//   g2[l] = sum_r sum_{m < nout} x_r[(m + l) mod n] conj(gy_r[m])      l < L
//
// is a cross-correlation of every row with its lag gradient.  It runs over
// the forward plan's overlap-save blocks: block b of row r correlates the M
// inputs x_r[bV + i] with the V lag gradients gy_r[bV + i] (zero past the
// block), and for l < L, i < V no index wraps (i + l < V + L - 1 <= M), so
//   g2 = (1/M) F^H( sum_{r,b} FFT_M(x_rb) .* conj(FFT_M(gy_rb)) )[0, L).
// One persistent grid walks the (row, block) items; a workgroup keeps its
// share of the cross-spectrum sum in registers (each thread owns the bins
// that the last Stockham pass leaves in its registers, the same bins for
// every block), writes it once, and two small kernels sum the per-workgroup
// spectra in a fixed order (in double) and apply the inverse transform and
// the chain rule of the window and the normalisation:
//   w1 = h .* w,  nu = ||w1||,  gw = h .* (g2/nu - w1 Re<g2, w1>/nu^3).
// The spectra of x and gy are recomputed from the cube (nothing of the
// forward's transforms is stored).
#include "fft_engine.hpp"
#include "mxr_cache.hpp"
#include "pcgrad.hpp"

namespace mxr {

template <int M>
struct WgT {
  static constexpr int NT = M / 16;
  using LY = Layout<M, 1, false>;
  using SF = Seq<M, false>;
  static constexpr int K = SF::n;
  static constexpr int R0 = SF::template R<0>;
  static constexpr int RL = SF::template R<K - 1>;
  static constexpr int LSL = SF::template LS<K - 1>;
  using P0 = Pass<LY, R0, M, NT, 1>;
  using PL = Pass<LY, RL, M, NT, 1>;
  static_assert(K >= 3, "gradient kernel needs at least three passes");
};

struct WgArgs {
  const float2 *x, *gy;
  BatchIdx bi;  // row -> offsets (sa: x, sb: gy)
  int64_t items, nblk;
  int n, seff, nout, V;
  float2 *partial;  // [gridDim.x][M]
  const float2 *twf;
};

// Forward M-point transform of ld(i), i < M, ending in registers (the last
// pass computed, not stored).  Pre: the workgroup's previous LDS reads are
// done.  Post: LDS is free once every thread has passed a barrier.
template <int M, class LoadF>
__device__ __forceinline__ void fwd_to_regs(typename WgT<M>::PL &pl, LoadF &&ld,
                                            const float2 *twf, float2 *lds) {
  using T = WgT<M>;
  {
    typename T::P0 ps;
    ps.load(ld);
    ps.template compute<1, false>(nullptr);
    ps.template store<1>([&](int t, int i, float2 v) { lds[T::LY::addr(t, i)] = v; });
  }
  wg_barrier();
  lds_passes<typename T::LY, typename T::SF, M, T::NT, 1, false, 1, T::K - 1>(lds, twf);
  pl.load([&](int t, int i) { return lds[T::LY::addr(t, i)]; });
  pl.template compute<T::LSL, false>(twf + T::SF::template TWOFF<T::K - 1>);
}

template <int M>
__global__ __launch_bounds__(WgT<M>::NT) void pc_wgrad_kernel(WgArgs a) {
  using T = WgT<M>;
  constexpr int BPT = T::PL::BPT;
  __shared__ float2 lds[T::LY::LDS_ELEMS];
  float2 acc[BPT][T::RL];
#pragma unroll
  for (int u = 0; u < BPT; ++u)
#pragma unroll
    for (int q = 0; q < T::RL; ++q) acc[u][q] = make_float2(0.f, 0.f);

  for (int64_t it = blockIdx.x; it < a.items; it += gridDim.x) {
    if (it != (int64_t)blockIdx.x) wg_barrier();  // previous item's LDS reads are done
    const int64_t row = it / a.nblk;
    const int lo = (int)(it % a.nblk) * a.V;
    int64_t ox, oy;
    batch_offsets(a.bi, row, &ox, &oy);
    const float2 *xr = a.x + ox;
    const float2 *gr = a.gy + oy;
    const int n32 = a.n, seff = a.seff;
    const int lim = (a.nout - lo) < a.V ? (a.nout - lo) : a.V;
    typename T::PL px;
    fwd_to_regs<M>(px, [&](int, int i) {
      int gi = lo + i;
      if (gi >= n32) gi -= n32;
      return gi < seff ? xr[gi] : make_float2(0.f, 0.f);
    }, a.twf, lds);
    wg_barrier();  // every read of the x spectrum's last input is done
    typename T::PL pg;
    fwd_to_regs<M>(pg, [&](int, int i) { return i < lim ? gr[lo + i] : make_float2(0.f, 0.f); },
                   a.twf, lds);
#pragma unroll
    for (int u = 0; u < BPT; ++u) {
      if (!T::PL::active(u)) continue;
#pragma unroll
      for (int q = 0; q < T::RL; ++q) {
        const float2 X = px.v[u][q], G = pg.v[u][q];
        acc[u][q].x += X.x * G.x + X.y * G.y;  // X conj(G)
        acc[u][q].y += X.y * G.x - X.x * G.y;
      }
    }
  }
  // the last pass leaves bin j + q*LSL of butterfly j = tid + u*NT in v[u][q]
  float2 *pp = a.partial + (int64_t)blockIdx.x * M;
#pragma unroll
  for (int u = 0; u < BPT; ++u) {
    if (!T::PL::active(u)) continue;
    const int j = threadIdx.x + u * T::NT;
#pragma unroll
    for (int q = 0; q < T::RL; ++q) pp[j + q * T::LSL] = acc[u][q];
  }
}

// M = 8192: the transforms of the 8192-point pulse-compression kernel
// (radix 2 x 16 x 16 x 16, 512 threads, the radix-2 pass folded into the
// load, power-base twiddles for Ls = 2, 32, 512), so the x spectrum, the gy
// spectrum and the running product take 16 bins each per thread (bins
// j + 512q).  Those three live sets need more than the 128 VGPRs of two
// workgroups per CU, so it runs one (up to 256 VGPRs, no spills).
struct Wg8192Args {
  WgArgs a;
  const float4 *bA, *bB, *bC;
};

// Forward 8192-point transform of ld(i); bins j + 512q end in v[q].  Pre:
// the workgroup's previous LDS reads are done.  Post: LDS is free (the last
// exchange ends with a barrier).
template <class LoadF>
__device__ __forceinline__ void fwd8192_regs(float2 (&v)[16], LoadF &&ld, const float4 *pA,
                                             const float4 *pB, const float4 *pC, float2 *lds) {
  constexpr int NT = 512;
  using LY = Layout<8192, 1, false>;
  const int j = threadIdx.x;
  auto xchg = [&](int ls_out) {
    const int k = j & (ls_out - 1), base = (j - k) * 16 + k;
#pragma unroll
    for (int q = 0; q < 16; ++q) lds[LY::addr(0, base + q * ls_out)] = v[q];
    wg_barrier();
#pragma unroll
    for (int m = 0; m < 16; ++m) v[m] = lds[LY::addr(0, j + m * NT)];
    wg_barrier();
  };
  {
    const int jh = j >> 1;
    const bool odd = (j & 1) != 0;
    float2 a[16], b[16];
#pragma unroll
    for (int m = 0; m < 16; ++m) {
      a[m] = ld(jh + 256 * m);
      b[m] = ld(jh + 256 * m + 4096);
    }
#pragma unroll
    for (int m = 0; m < 16; ++m) v[m] = odd ? csub(a[m], b[m]) : cadd(a[m], b[m]);
  }
  float4 t0 = pA[0], t1 = pA[1];
  tw_from_base(v, t0, t1, false);  // Ls = 2
  dft16<false>(v);
  t0 = pB[0];
  t1 = pB[1];
  xchg(2);
  tw_from_base(v, t0, t1, false);  // Ls = 32
  dft16<false>(v);
  t0 = pC[0];
  t1 = pC[1];
  xchg(32);
  tw_from_base(v, t0, t1, false);  // Ls = 512
  dft16<false>(v);
}

__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(2)))
void wgrad8192_kernel(Wg8192Args p) {
  using LY = Layout<8192, 1, false>;
  __shared__ float2 lds[LY::LDS_ELEMS];
  const WgArgs &a = p.a;
  const int j = threadIdx.x;
  const float4 *pA = p.bA + (j & 1) * 2, *pB = p.bB + (j & 31) * 2, *pC = p.bC + j * 2;
  float2 acc[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) acc[q] = make_float2(0.f, 0.f);
  for (int64_t it = blockIdx.x; it < a.items; it += gridDim.x) {
    // (the previous item's last exchange ended with a barrier)
    const int64_t row = it / a.nblk;
    const int lo = (int)(it % a.nblk) * a.V;
    int64_t ox, oy;
    batch_offsets(a.bi, row, &ox, &oy);
    const float2 *xr = a.x + ox;
    const float2 *gr = a.gy + oy;
    const int n32 = a.n, seff = a.seff;
    const int lim = (a.nout - lo) < a.V ? (a.nout - lo) : a.V;
    float2 X[16], Gv[16];
    fwd8192_regs(X, [&](int i) {
      int gi = lo + i;
      if (gi >= n32) gi -= n32;
      return gi < seff ? xr[gi] : make_float2(0.f, 0.f);
    }, pA, pB, pC, lds);
    fwd8192_regs(Gv, [&](int i) { return i < lim ? gr[lo + i] : make_float2(0.f, 0.f); }, pA, pB,
                 pC, lds);
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      acc[q].x += X[q].x * Gv[q].x + X[q].y * Gv[q].y;  // X conj(G)
      acc[q].y += X[q].y * Gv[q].x - X[q].x * Gv[q].y;
    }
  }
  float2 *pp = a.partial + (int64_t)blockIdx.x * 8192;
#pragma unroll
  for (int q = 0; q < 16; ++q) pp[j + 512 * q] = acc[q];
}

// spec[k] = sum over workgroups g of partial[g][k], in order, in double.
__global__ __launch_bounds__(64) void wgrad_reduce_kernel(const float2 *partial, int G, int M,
                                                          float2 *spec) {
  const int k = blockIdx.x * 64 + threadIdx.x;
  if (k >= M) return;
  double re = 0.0, im = 0.0;
#pragma unroll 8
  for (int g = 0; g < G; ++g) {
    const float2 v = partial[(int64_t)g * M + k];
    re += (double)v.x;
    im += (double)v.y;
  }
  spec[k] = make_float2((float)re, (float)im);
}

// g2 = (1/M) F^H spec (spec == nullptr: g2 = 0), then the window and
// normalisation chain rule into gw.  One workgroup.
template <int M>
__global__ __launch_bounds__(WgT<M>::NT) void wgrad_finish_kernel(const float2 *spec,
                                                                  const float2 *w, int64_t ws,
                                                                  float2 *gw, int64_t gws, int L,
                                                                  int window, int normalize,
                                                                  const float2 *twf) {
  using T = WgT<M>;
  constexpr int NWV = T::NT / 64;
  __shared__ float2 lds[T::LY::LDS_ELEMS];
  __shared__ double red[2][NWV];
  const float sc = 1.0f / (float)M;  // exact: a power of two
  {
    typename T::P0 ps;
    ps.load([&](int, int i) { return spec ? cscale(spec[i], sc) : make_float2(0.f, 0.f); });
    ps.template compute<1, true>(nullptr);
    ps.template store<1>([&](int t, int i, float2 v) { lds[T::LY::addr(t, i)] = v; });
  }
  wg_barrier();
  // the forward sequence with conjugate twiddles is the unnormalised inverse
  lds_passes<typename T::LY, typename T::SF, M, T::NT, 1, true, 1, T::K>(lds, twf);

  double nu2 = 0.0, s = 0.0;
  for (int l = threadIdx.x; l < L; l += T::NT) {
    const float h = window ? hamming_f32(l, L) : 1.0f;
    const float2 w1 = cscale(w[(int64_t)l * ws], h);
    const float2 g = lds[T::LY::addr(0, l)];
    nu2 += (double)w1.x * w1.x + (double)w1.y * w1.y;
    s += (double)g.x * w1.x + (double)g.y * w1.y;  // Re(conj(g2) w1)
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    nu2 += __shfl_xor(nu2, o);
    s += __shfl_xor(s, o);
  }
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = nu2;
    red[1][threadIdx.x >> 6] = s;
  }
  __syncthreads();
  nu2 = 0.0;
  s = 0.0;
#pragma unroll
  for (int q = 0; q < NWV; ++q) {
    nu2 += red[0][q];
    s += red[1][q];
  }
  float ca = 1.0f, cb = 0.0f;
  if (normalize && nu2 > 0.0) {
    const double nu = sqrt(nu2);
    ca = (float)(1.0 / nu);
    cb = (float)(s / (nu2 * nu));
  }
  for (int l = threadIdx.x; l < L; l += T::NT) {
    const float h = window ? hamming_f32(l, L) : 1.0f;
    const float2 w1 = cscale(w[(int64_t)l * ws], h);
    const float2 g = lds[T::LY::addr(0, l)];
    const float2 r = make_float2(ca * g.x - cb * w1.x, ca * g.y - cb * w1.y);
    gw[(int64_t)l * gws] = cscale(r, h);
  }
}

template <int M>
static mxr_status wgrad_run(const WgradJob &j, hipStream_t s) {
  using T = WgT<M>;
  const float2 *twf = twiddles_device(M, false);
  MXR_REQUIRE(twf != nullptr, MXR_OUT_OF_MEMORY, "pc waveform gradient: twiddle table allocation failed");
  MXR_REQUIRE(j.L <= M && j.V >= 1, MXR_INVALID_SIZE, "pc waveform gradient: inconsistent block plan");
  const int64_t items = j.rows * j.nblk;
  float2 *spec = nullptr;
  if (items > 0) {
    int per_cu = 1, cus = 0;
    hipError_t oe;
    if constexpr (M == 8192)
      oe = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, wgrad8192_kernel, 512, 0);
    else
      oe = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, pc_wgrad_kernel<M>, T::NT, 0);
    if (oe != hipSuccess || per_cu < 1) per_cu = 1;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device_id()) !=
            hipSuccess || cus < 1)
      cus = 256;
    const int64_t grid = items < (int64_t)per_cu * cus ? items : (int64_t)per_cu * cus;
    float2 *ws = (float2 *)stream_workspace("pc_wgrad", sizeof(float2) * (size_t)(grid + 1) * M, s);
    if (ws == nullptr) return workspace_failure("pc waveform gradient");
    spec = ws + grid * M;
    WgArgs a{};
    a.x = j.x;
    a.gy = j.gy;
    a.bi = j.bi;
    a.items = items;
    a.nblk = j.nblk;
    a.n = j.n;
    a.seff = j.seff;
    a.nout = j.nout;
    a.V = (int)j.V;
    a.partial = ws;
    a.twf = twf;
    if constexpr (M == 8192) {
      Wg8192Args p{a, pow_base_device(8192, 2), pow_base_device(8192, 32), pow_base_device(8192, 512)};
      MXR_REQUIRE(p.bA && p.bB && p.bC, MXR_OUT_OF_MEMORY,
                  "pc waveform gradient: twiddle table allocation failed");
      ProfScope ps("wgrad8192_kernel", s);
      wgrad8192_kernel<<<dim3((unsigned)grid), dim3(512), 0, s>>>(p);
    } else {
      ProfScope ps("pc_wgrad_kernel", s);
      pc_wgrad_kernel<M><<<dim3((unsigned)grid), dim3(T::NT), 0, s>>>(a);
    }
    MXR_HIP_TRY(hipGetLastError());
    {
      ProfScope ps("wgrad_reduce_kernel", s);
      wgrad_reduce_kernel<<<dim3((M + 63) / 64), dim3(64), 0, s>>>(ws, (int)grid, M, spec);
    }
    MXR_HIP_TRY(hipGetLastError());
  }
  {
    ProfScope ps("wgrad_finish_kernel", s);
    wgrad_finish_kernel<M><<<dim3(1), dim3(T::NT), 0, s>>>(spec, j.w, j.wstride, j.gw, j.gwstride,
                                                          j.L, j.window, j.normalize, twf);
  }
  MXR_HIP_TRY(hipGetLastError());
  return MXR_SUCCESS;
}

mxr_status pc_waveform_grad_run(const WgradJob &j, hipStream_t s) {
  switch (j.M) {
    case 1024: return wgrad_run<1024>(j, s);
    case 2048: return wgrad_run<2048>(j, s);
    case 4096: return wgrad_run<4096>(j, s);
    case 8192: return wgrad_run<8192>(j, s);
    case 16384: return wgrad_run<16384>(j, s);
    default: break;
  }
  set_error(MXR_NOT_SUPPORTED, "pc waveform gradient: block size %d unsupported", j.M);
  return MXR_NOT_SUPPORTED;
}

}  // namespace mxr
