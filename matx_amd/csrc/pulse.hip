// Pulse compression and fft_conv: matched filtering of every pulse (row)
// against one waveform, i.e. y = (1/n) ifft(fft(x, n) .* conj(fft(w, n))),
// the PulseCompression stage of examples/simple_radar_pipeline.h:273-301 and
// the fft_conv idiom of examples/fft_conv.cu:90-108.
//
// MI355X design.  The lags a radar keeps (r < R = S-L+1, or all n for a full
// circular fft_conv) are produced by overlap-save over blocks of M <= n points:
// block b reads x[(bV + i) mod n], i < M, correlates it circularly with the
// M-point spectrum conj(fft(w, M)) and keeps its first V = M-L+1 lags, which
// equal lags bV..bV+V-1 of the n-point circular correlation exactly (no lag
// inside a block wraps).  A planner picks M to minimise blocks*M*log2(M)
// (4096 for a 9000-sample pulse and a 1000-sample waveform, 36% fewer flops
// than one 16384-point transform), and M = n reproduces the plain n-point
// form.  Each block is one fused kernel pass: load (zero-padding implicit),
// forward Stockham passes, conj-multiply (the waveform spectrum carries the
// 1/M of the inverse, an exact power-of-two scale), inverse passes, store of
// the kept lags.  The last forward pass and the first
// inverse pass share one radix, so they run back to back in registers.
#include <cmath>
#include <map>
#include <mutex>
#include <tuple>

#include "fft_engine.hpp"
#include "mxr_cache.hpp"
#include "mxr_trace.hpp"
#include "pcgrad.hpp"
#include "ragged.hpp"

namespace mxr {

struct PcArgs {
  const float2 *x;
  float2 *y;
  int64_t rows;
  BatchIdx bi;      // row index -> offsets (sa: x, sb: y)
  int64_t seff;     // valid input samples per row (min(S, n))
  int64_t n;        // circular period of the correlation
  int64_t nout;     // lags kept per row
  int64_t V;        // kept lags per block
  int64_t nblk;     // blocks per row
  int64_t bpw;      // blocks per workgroup (nblk when y aliases x)
  int64_t groups;   // total workgroups
  const float2 *wc; // M-point waveform spectrum, scaled by 1/M
  const float2 *twf, *twi;
};

template <int M>
struct PcCfg {
  static constexpr int NT = M / 16;
};

// One overlap-save block of one row: x_r[(lo + i) mod n], i < M (zero past
// seff) correlated with the M-point spectrum wc; lags lo .. lo+min(V, nout-lo)
// stored to y_r.
struct OlsBlock {
  const float2 *xr;
  float2 *yr;
  int lo, n, seff, nout, V;
  const float2 *wc;  // M-point spectrum, scaled by 1/M
};

// Body shared by the per-call kernel and the ragged-batch kernels.  Pre: the
// workgroup's previous use of lds is finished (barrier passed).
template <int M>
__device__ __forceinline__ void ols_block(const OlsBlock &ob, const float2 *twf,
                                          const float2 *twi, float2 *lds) {
  constexpr int NT = PcCfg<M>::NT;
  using LY = Layout<M, 1, false>;
  using SF = Seq<M, false>;
  using SI = Seq<M, true>;
  constexpr int K = SF::n;
  constexpr int R0 = SF::template R<0>;
  constexpr int RL = SF::template R<K - 1>;
  constexpr int LSL = SF::template LS<K - 1>;
  static_assert(K >= 2, "pc kernel needs at least two passes");
  const float2 *__restrict__ xr = ob.xr;
  float2 *__restrict__ yr = ob.yr;
  const int lo = ob.lo, n32 = ob.n, seff32 = ob.seff;

  // forward pass 0 straight from HBM (zero padding and wrap implicit)
  {
    Pass<LY, R0, M, NT, 1> ps;
    ps.load([&](int, int i) {
      int gi = lo + i;
      if (gi >= n32) gi -= n32;
      return keep_if(xr[gi < seff32 ? gi : 0], gi < seff32);
    });
    ps.template compute<1, false>(nullptr);
    ps.template store<1>([&](int t, int i, float2 v) { lds[LY::addr(t, i)] = v; });
  }
  wg_barrier();
  lds_passes<LY, SF, M, NT, 1, false, 1, K - 1>(lds, twf);

  // last forward pass, conj-multiply, first inverse pass: one register pass
  {
    Pass<LY, RL, M, NT, 1> ps;
    ps.load([&](int t, int i) { return lds[LY::addr(t, i)]; });
    wg_barrier();
    ps.template compute<LSL, false>(twf + SF::template TWOFF<K - 1>);
#pragma unroll
    for (int u = 0; u < ps.BPT; ++u) {
      if (!ps.active(u)) continue;
      const int j = threadIdx.x + u * NT;  // k == j in the last pass
#pragma unroll
      for (int q = 0; q < RL; ++q) ps.v[u][q] = cmul(ps.v[u][q], ob.wc[j + q * LSL]);
    }
    ps.template compute<1, true>(nullptr);
    ps.template store<1>([&](int t, int i, float2 v) { lds[LY::addr(t, i)] = v; });
  }
  wg_barrier();
  lds_passes<LY, SI, M, NT, 1, true, 1, K - 1>(lds, twi);

  // last inverse pass straight to HBM
  {
    constexpr int RI = SI::template R<K - 1>;
    constexpr int LSI = SI::template LS<K - 1>;
    Pass<LY, RI, M, NT, 1> ps;
    ps.load([&](int t, int i) { return lds[LY::addr(t, i)]; });
    ps.template compute<LSI, true>(twi + SI::template TWOFF<K - 1>);
    const int lim = (ob.nout - lo) < ob.V ? (ob.nout - lo) : ob.V;
    ps.template store<LSI>([&](int, int i, float2 v) {
      if (i < lim) yr[lo + i] = v;  // the spectrum carries the 1/M
    });
  }
}

// TB = 4096/M small blocks per 256-thread workgroup (M <= 2048): transform
// t runs the block obs[t] (kept in LDS).  Same passes as ols_block; an
// inactive transform has seff = 0 (reads nothing) and nout = lo (stores
// nothing).
template <int M>
__device__ __forceinline__ void ols_multi(const OlsBlock *obs, const float2 *twf,
                                          const float2 *twi, float2 *lds) {
  constexpr int NT = 256, TB = 4096 / M;
  using LY = Layout<M, TB, false>;
  using SF = Seq<M, false>;
  using SI = Seq<M, true>;
  constexpr int K = SF::n;
  constexpr int R0 = SF::template R<0>;
  constexpr int RL = SF::template R<K - 1>;
  constexpr int LSL = SF::template LS<K - 1>;
  constexpr int NBL = M / RL;
  static_assert(K >= 2 && R0 == 16, "small-block kernel expects a radix-16 first pass");
  {
    Pass<LY, R0, M, NT, TB> ps;
    ps.load([&](int t, int i) {
      const OlsBlock &ob = obs[t];
      int gi = ob.lo + i;
      if (gi >= ob.n) gi -= ob.n;
      // sample 0 of the row (an idle transform's row is a real one) and a
      // mask past seff: unconditional loads
      return keep_if(ob.xr[gi < ob.seff ? gi : 0], gi < ob.seff);
    });
    ps.template compute<1, false>(nullptr);
    ps.template store<1>([&](int t, int i, float2 v) { lds[LY::addr(t, i)] = v; });
  }
  wg_barrier();
  lds_passes<LY, SF, M, NT, TB, false, 1, K - 1>(lds, twf);
  {
    Pass<LY, RL, M, NT, TB> ps;
    ps.load([&](int t, int i) { return lds[LY::addr(t, i)]; });
    wg_barrier();
    ps.template compute<LSL, false>(twf + SF::template TWOFF<K - 1>);
#pragma unroll
    for (int u = 0; u < ps.BPT; ++u) {
      if (!ps.active(u)) continue;
      const int g = threadIdx.x + u * NT;
      const int t = g / NBL, j = g % NBL;  // k == j in the last pass
      const float2 *wc = obs[t].wc;
#pragma unroll
      for (int q = 0; q < RL; ++q) ps.v[u][q] = cmul(ps.v[u][q], wc[j + q * LSL]);
    }
    ps.template compute<1, true>(nullptr);
    ps.template store<1>([&](int t, int i, float2 v) { lds[LY::addr(t, i)] = v; });
  }
  wg_barrier();
  lds_passes<LY, SI, M, NT, TB, true, 1, K - 1>(lds, twi);
  {
    constexpr int RI = SI::template R<K - 1>;
    constexpr int LSI = SI::template LS<K - 1>;
    Pass<LY, RI, M, NT, TB> ps;
    ps.load([&](int t, int i) { return lds[LY::addr(t, i)]; });
    ps.template compute<LSI, true>(twi + SI::template TWOFF<K - 1>);
    ps.template store<LSI>([&](int t, int i, float2 v) {
      const OlsBlock &ob = obs[t];
      const int lim = (ob.nout - ob.lo) < ob.V ? (ob.nout - ob.lo) : ob.V;
      if (i < lim) ob.yr[ob.lo + i] = v;  // the spectrum carries the 1/M
    });
  }
}

// Per-call small blocks: 4096/M (row, block) groups per workgroup.
template <int M>
__global__ __launch_bounds__(256) void pc_ols_multi_kernel(PcArgs a) {
  constexpr int TB = 4096 / M;
  __shared__ float2 lds[Layout<M, TB, false>::LDS_ELEMS];
  __shared__ OlsBlock obs[TB];
  if (threadIdx.x < TB) {
    const int64_t g0 = (int64_t)blockIdx.x * TB;
    const int64_t g = g0 + threadIdx.x < a.groups ? g0 + threadIdx.x : g0;
    const int64_t row = g / a.nblk;
    int64_t ox, oy;
    batch_offsets(a.bi, row, &ox, &oy);
    OlsBlock ob{a.x + ox, a.y + oy, (int)((g % a.nblk) * a.V), (int)a.n, (int)a.seff,
                (int)a.nout, (int)a.V, a.wc};
    if (g0 + threadIdx.x >= a.groups) {  // past the last group: idle transform
      ob.seff = 0;
      ob.nout = ob.lo;
    }
    obs[threadIdx.x] = ob;
  }
  __syncthreads();
  ols_multi<M>(obs, a.twf, a.twi, lds);
}

// ROWLOOP: one workgroup walks blocks [b0, b1) of its row in order (needed
// when y aliases x: a block only overwrites lags below the next block's
// first input).  Otherwise each workgroup does exactly one block.
template <int M, bool ROWLOOP>
__global__ __launch_bounds__(PcCfg<M>::NT) __attribute__((amdgpu_waves_per_eu(4)))
void pc_ols_kernel(PcArgs a) {
  using LY = Layout<M, 1, false>;
  __shared__ float2 lds[LY::LDS_ELEMS];
  const int64_t g = xcd_remap(blockIdx.x, a.groups);
  const int64_t gpr = (a.nblk + a.bpw - 1) / a.bpw;
  const int64_t row = g / gpr;
  const int64_t b0 = (g % gpr) * a.bpw;
  const int64_t b1 = ROWLOOP ? (b0 + a.bpw < a.nblk ? b0 + a.bpw : a.nblk) : b0 + 1;
  int64_t ox, oy;
  batch_offsets(a.bi, row, &ox, &oy);
  OlsBlock ob{a.x + ox, a.y + oy, 0, (int)a.n, (int)a.seff, (int)a.nout, (int)a.V, a.wc};
  for (int64_t b = b0; b < b1; ++b) {
    if (b != b0) wg_barrier();  // previous block's LDS reads are done
    ob.lo = (int)(b * a.V);
    ols_block<M>(ob, a.twf, a.twi, lds);
  }
}

// Dedicated M = 4096 block kernel (plan radix 16 x 16 x 16, 256 threads, one
// radix-16 butterfly per thread per pass).  Twiddles are built in registers
// from four base powers W^k, W^2k, W^4k, W^8k per pass (one 32-byte row of a
// compact table, W = e^{-2 pi i / (16 Ls)}): W^{mk} for the other m is a
// product of at most four of them (at most three extra roundings).  The base
// rows are fetched before each LDS exchange, so their latency hides behind
// the exchange and its barriers instead of stalling the butterfly after it;
// the inverse passes reuse the forward rows conjugated.
struct Pc4096Args {
  PcArgs a;
  const float4 *base1;  // [16][2] float4: power-base rows of pass Ls = 16
  const float4 *base2;  // [256][2] float4: power-base rows of pass Ls = 256
};


// One M = 4096 overlap-save block (256 threads); lds holds 4096 padded
// elements.  Pre: the workgroup's previous use of lds is finished.
// Thread j owns butterfly j of every radix-16 pass: its twiddles of the
// Ls = 16 and Ls = 256 passes come from power-base rows (j & 15) and j
// (W^k, W^2k, W^4k, W^8k), fetched before each LDS exchange so their
// latency hides behind it (the inverse passes re-fetch them from cache and
// conjugate; keeping them live costs more registers than the reload).  A block that
// does not wrap reads x through a raw buffer resource ending at the row's
// last sample (the hardware returns zeros past it) and writes its kept lags
// through one ending at its last kept lag (stores past it are dropped), so
// the block needs no per-element tests.  spec(i) gives spectrum bin i (the
// block's 1/M included); thread j reads bins j + 256 q only.
// Loads samples m0 .. m0 + NV - 1 of thread j of block ob
// (x[lo + j + 256 m]) into v[0 .. NV).
template <int NV>
__device__ __forceinline__ void load_block4096(const OlsBlock &ob, float2 (&v)[NV], int m0 = 0) {
  constexpr int NT = 256;
  const int j = threadIdx.x;
  const float2 *__restrict__ xr = ob.xr;
  const int lo = ob.lo, n32 = ob.n, seff32 = ob.seff;
  if (lo + 4096 <= n32) {
    int avail = seff32 - lo;
    avail = avail < 0 ? 0 : (avail > 4096 ? 4096 : avail);
    const buf_rsrc rx = make_rsrc(xr + lo, avail * 8);
#pragma unroll
    for (int m = 0; m < NV; ++m) v[m] = buf_load2(rx, (j + (m0 + m) * NT) * 8);
  } else {  // the block wraps around the circular period n (full fft_conv)
#pragma unroll
    for (int m = 0; m < NV; ++m) {
      int gi = lo + j + (m0 + m) * NT;
      if (gi >= n32) gi -= n32;
      v[m] = gi < seff32 ? xr[gi] : make_float2(0.f, 0.f);
    }
  }
}

// The block's transforms and stores, from its samples in v (loaded by
// load_block4096).
// LATE_SPEC: the spectrum is written by other threads of the workgroup
// while the forward passes run (pc4096_wf_kernel); it is read after one more
// barrier, following the last forward pass, instead of being prefetched.
template <class Spec, bool LATE_SPEC = false>
__device__ __forceinline__ void filter_block4096(const OlsBlock &ob, float2 (&v)[16],
                                                 const float4 *base1, const float4 *base2,
                                                 float2 *lds, Spec spec) {
  constexpr int NT = 256;
  using LY = Layout<4096, 1, false>;
  const int j = threadIdx.x;
  const int lo = ob.lo;
  // Stockham exchange: write this pass's outputs, read the next pass's
  // inputs.  In the padded layout (one spare slot per 16) the outputs
  // base + q ls_out of butterfly j sit at pad(base) + q (ls_out = 1) or
  // pad(base) + 17 q (ls_out = 16), and input j + 256 m at pad(j) + 272 m:
  // one address per phase, the rest are instruction offsets.  (An unpadded
  // 32 KiB layout with an exclusive-or swizzle admits a fifth workgroup per
  // CU and measured 2 % slower.)
  const int rbase = lds_pad(j);
  auto xchg = [&](int ls_out) {
    const int k = j & (ls_out - 1), base = (j - k) * 16 + k;
    const int wbase = lds_pad(base), wstep = ls_out == 1 ? 1 : 17;
#pragma unroll
    for (int q = 0; q < 16; ++q) lds[wbase + q * wstep] = v[q];
    wg_barrier();
#pragma unroll
    for (int m = 0; m < 16; ++m) v[m] = lds[rbase + m * (NT + NT / 16)];
    wg_barrier();
  };
  const float4 *b1p = base1 + (j & 15) * 2;
  const float4 *b2p = base2 + j * 2;
  float4 t0, t1;

  // forward pass 0 (Ls = 1) on the samples as loaded
  dft16<false>(v);
  t0 = b1p[0];
  t1 = b1p[1];
  xchg(1);
  tw_from_base(v, t0, t1, false);  // forward pass 1 (Ls = 16)
  dft16<false>(v);
  t0 = b2p[0];
  t1 = b2p[1];
  // the spectrum bins this thread multiplies, fetched before the exchange
  // so that their cache latency hides behind it and the last forward pass
  float2 sp[16];
  if constexpr (!LATE_SPEC) {
#pragma unroll
    for (int q = 0; q < 16; ++q) sp[q] = spec(j + q * 256);
  }
  xchg(16);
  tw_from_base(v, t0, t1, false);  // forward pass 2 (Ls = 256)
  dft16<false>(v);
  if constexpr (LATE_SPEC) {
    wg_barrier();  // the spectrum is complete
#pragma unroll
    for (int q = 0; q < 16; ++q) sp[q] = spec(j + q * 256);
  }
#pragma unroll
  for (int q = 0; q < 16; ++q) v[q] = cmul(v[q], sp[q]);
  dft16<true>(v);            // inverse pass 0 (Ls = 1)
  t0 = b1p[0];
  t1 = b1p[1];
  xchg(1);
  tw_from_base(v, t0, t1, true);   // inverse pass 1 (Ls = 16)
  dft16<true>(v);
  t0 = b2p[0];
  t1 = b2p[1];
  xchg(16);
  tw_from_base(v, t0, t1, true);   // inverse pass 2 (Ls = 256)
  dft16<true>(v);
  // lags lo + j + 256 q below lo + lim are kept (the spectrum carries the 1/M)
  const int lim = (ob.nout - lo) < ob.V ? (ob.nout - lo) : ob.V;
  const buf_rsrc ry = make_rsrc(ob.yr + lo, lim * 8);
#pragma unroll
  for (int q = 0; q < 16; ++q) buf_store2(ry, (j + q * 256) * 8, v[q]);
}

// One block, loaded and filtered.  Pre: the workgroup's previous use of lds
// is finished.
template <class Spec>
__device__ __forceinline__ void ols_block4096(const OlsBlock ob, const float4 *base1,
                                              const float4 *base2, float2 *lds, Spec spec) {
  float2 v[16];
  load_block4096(ob, v, 0);
  filter_block4096(ob, v, base1, base2, lds, spec);
}

// The block's spectrum from its buffer (ob.wc).
__device__ __forceinline__ void ols_block4096(const OlsBlock ob, const float4 *base1,
                                              const float4 *base2, float2 *lds) {
  ols_block4096(ob, base1, base2, lds, [&](int i) { return ob.wc[i]; });
}

__device__ __forceinline__ OlsBlock pc4096_block(const PcArgs &a, int64_t g) {
  const int64_t row = g / a.nblk;
  int64_t ox, oy;
  batch_offsets(a.bi, row, &ox, &oy);
  return OlsBlock{a.x + ox, a.y + oy, (int)((g % a.nblk) * a.V), (int)a.n, (int)a.seff,
                  (int)a.nout, (int)a.V, a.wc};
}

// One workgroup per block.  (A persistent form that loads the next block's
// samples during the current block's inverse passes needs more registers
// than four workgroups per CU allow, and measured 11 % slower.)
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4)))
void pc4096_kernel(Pc4096Args pa) {
  using LY = Layout<4096, 1, false>;
  __shared__ float2 lds[LY::LDS_ELEMS];
  const PcArgs &a = pa.a;
  ols_block4096(pc4096_block(a, xcd_remap(blockIdx.x, a.groups)), pa.base1, pa.base2, lds);
}

// Small grids (fft_conv on a few hundred pulses, where one workgroup per
// pulse leaves most CUs idle): every workgroup transforms the zero-padded
// waveform itself instead of waiting for a one-workgroup spectrum launch.
// The workgroup has 512 threads: threads 0-255 load and filter the block
// (the forward passes, then the multiply and the inverse passes), threads
// 256-511 transform the waveform at the same time through their own LDS
// buffer and leave the spectrum there (bins j + 256 q of their thread j,
// where the block's thread j reads them after one more barrier).  The
// waveform threads then take part in the block's remaining barriers.  No
// window, normalisation or write-back (the fft_conv form).
struct Pc4096WfArgs {
  Pc4096Args p;
  const float2 *w;
  int64_t wstride;
  int L, conj;
  float out_scale;
};

__global__ __launch_bounds__(512) void pc4096_wf_kernel(Pc4096WfArgs q) {
  using LY = Layout<4096, 1, false>;
  __shared__ float2 lds[LY::LDS_ELEMS];   // the block's exchanges
  __shared__ float2 wsp[LY::LDS_ELEMS];   // the waveform's exchanges, then its spectrum
  constexpr int NT = 256;
  if (threadIdx.x < NT) {
    const PcArgs &a = q.p.a;
    const int64_t g = xcd_remap(blockIdx.x, a.groups);
    const int64_t row = g / a.nblk;
    int64_t ox, oy;
    batch_offsets(a.bi, row, &ox, &oy);
    const OlsBlock ob{a.x + ox, a.y + oy, (int)((g % a.nblk) * a.V), (int)a.n, (int)a.seff,
                      (int)a.nout, (int)a.V, nullptr};
    float2 v[16];
    load_block4096(ob, v, 0);
    // barriers: 4 in the forward exchanges, 1 for the spectrum, 4 in the
    // inverse exchanges
    auto spec = [&](int i) { return wsp[i]; };
    filter_block4096<decltype(spec), true>(ob, v, q.p.base1, q.p.base2, lds, spec);
    return;
  }
  // the waveform: the same three forward passes, exchanged through wsp
  const int j = threadIdx.x - NT;
  float2 v[16];
  auto xchg = [&](int ls_out) {
    const int k = j & (ls_out - 1), base = (j - k) * 16 + k;
#pragma unroll
    for (int m = 0; m < 16; ++m) wsp[LY::addr(0, base + m * ls_out)] = v[m];
    wg_barrier();
#pragma unroll
    for (int m = 0; m < 16; ++m) v[m] = wsp[LY::addr(0, j + m * NT)];
    wg_barrier();
  };
#pragma unroll
  for (int m = 0; m < 16; ++m) {
    const int i = j + m * NT;
    v[m] = keep_if(q.w[(int64_t)(i < q.L ? i : 0) * q.wstride], i < q.L);
  }
  const float4 *b1p = q.p.base1 + (j & 15) * 2;
  const float4 *b2p = q.p.base2 + j * 2;
  dft16<false>(v);
  float4 t0 = b1p[0], t1 = b1p[1];
  xchg(1);
  tw_from_base(v, t0, t1, false);
  dft16<false>(v);
  t0 = b2p[0];
  t1 = b2p[1];
  xchg(16);
  tw_from_base(v, t0, t1, false);
  dft16<false>(v);
  // the exchange above ended with a barrier after every read of wsp
#pragma unroll
  for (int m = 0; m < 16; ++m) {
    float2 x = cscale(v[m], q.out_scale);
    if (q.conj) x.y = -x.y;
    wsp[j + m * NT] = x;
  }
  wg_barrier();  // the spectrum is complete (the block's fifth barrier)
  for (int b = 0; b < 4; ++b) wg_barrier();  // the block's inverse exchanges
}

// Dedicated M = 8192 block kernel: radix sequence 2 x 16 x 16 x 16 over 512
// threads (one radix-16 butterfly per thread per pass) with the power-base
// twiddles of the 4096-point kernel for the passes Ls = 2, 32, 512
// (conjugated for the inverse).  The forward radix-2 pass (Ls = 1, no
// twiddles) is folded into the load: radix-16 butterfly j of the next pass
// takes X1[j + 512m] = x[j0] +- x[j0 + 4096] with j0 = j/2 + 256m and the
// sign given by j's parity, so every thread reads its 16 input pairs straight
// from HBM (its neighbour reads the same pairs for the other sign, from L1).
// The inverse radix-2 pass pairs bins k and k + 4096, which the forward's
// last pass leaves in one thread (slots q and q + 8), so it runs in
// registers before the first inverse exchange: five LDS exchanges per block.
struct Pc8192Args {
  PcArgs a;
  const float4 *bA, *bB, *bC;  // power-base rows for Ls = 2, 32, 512
};

__device__ __forceinline__ void ols_block8192(const OlsBlock ob, const float4 *bA, const float4 *bB,
                                              const float4 *bC, float2 *lds) {
  constexpr int NT = 512;
  const int j = threadIdx.x;
  const float2 *__restrict__ xr = ob.xr;
  const int lo = ob.lo, n32 = ob.n, seff32 = ob.seff;

  float2 v[16];
  // Stockham exchange in the padded layout (pad(i) = i + i / 16): outputs
  // base + q ls_out of butterfly j sit at pad(base) + 2 q + (q >= 8)
  // (ls_out = 2) or pad(base) + 34 q (ls_out = 32), input j + 512 m at
  // pad(j) + 544 m: one address per phase, the rest instruction offsets.
  const int rbase = lds_pad(j);
  auto xchg = [&](int ls_out) {
    const int k = j & (ls_out - 1), base = (j - k) * 16 + k;
    const int wbase = lds_pad(base);
#pragma unroll
    for (int q = 0; q < 16; ++q) lds[wbase + (ls_out == 2 ? 2 * q + (q >= 8) : 34 * q)] = v[q];
    wg_barrier();
#pragma unroll
    for (int m = 0; m < 16; ++m) v[m] = lds[rbase + 544 * m];
    wg_barrier();
  };
  auto ldx = [&](int i) {
    int gi = lo + i;
    if (gi >= n32) gi -= n32;
    return gi < seff32 ? xr[gi] : make_float2(0.f, 0.f);
  };
  const float4 *pA = bA + (j & 1) * 2;
  const float4 *pB = bB + (j & 31) * 2;
  const float4 *pC = bC + j * 2;

  // forward radix-2 pass folded into the load; a block that does not wrap
  // reads through a raw buffer resource ending at the row's last sample
  // (zeros past it from the hardware range check)
  {
    const int jh = j >> 1;
    const bool odd = (j & 1) != 0;
    float2 a[16], b[16];
    if (lo + 8192 <= n32) {
      int avail = seff32 - lo;
      avail = avail < 0 ? 0 : (avail > 8192 ? 8192 : avail);
      const buf_rsrc rx = make_rsrc(xr + lo, avail * 8);
#pragma unroll
      for (int m = 0; m < 16; ++m) {
        a[m] = buf_load2(rx, (jh + 256 * m) * 8);
        b[m] = buf_load2(rx, (jh + 256 * m + 4096) * 8);
      }
    } else {
#pragma unroll
      for (int m = 0; m < 16; ++m) {
        a[m] = ldx(jh + 256 * m);
        b[m] = ldx(jh + 256 * m + 4096);
      }
    }
#pragma unroll
    for (int m = 0; m < 16; ++m) v[m] = odd ? csub(a[m], b[m]) : cadd(a[m], b[m]);
  }
  float4 t0 = pA[0], t1 = pA[1];
  tw_from_base(v, t0, t1, false);  // forward pass Ls = 2
  dft16<false>(v);
  t0 = pB[0];
  t1 = pB[1];
  xchg(2);
  tw_from_base(v, t0, t1, false);  // forward pass Ls = 32
  dft16<false>(v);
  t0 = pC[0];
  t1 = pC[1];
  // the spectrum bins this thread multiplies, fetched before the exchange
  float2 sp[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) sp[q] = ob.wc[j + q * 512];
  xchg(32);
  tw_from_base(v, t0, t1, false);  // forward pass Ls = 512: bins j + 512q
  dft16<false>(v);
#pragma unroll
  for (int q = 0; q < 16; ++q) v[q] = cmul(v[q], sp[q]);
  // inverse radix-2 pass (Ls = 1) on bins j + 512q, j + 512q + 4096; its
  // outputs 2 (j + 512 q) and the next sit at pad(2 j) + 1088 q and + 1
  t0 = pA[0];
  t1 = pA[1];
  const int e0 = lds_pad(2 * j);
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    lds[e0 + 1088 * q] = cadd(v[q], v[q + 8]);
    lds[e0 + 1088 * q + 1] = csub(v[q], v[q + 8]);
  }
  wg_barrier();
#pragma unroll
  for (int m = 0; m < 16; ++m) v[m] = lds[rbase + 544 * m];
  wg_barrier();
  tw_from_base(v, t0, t1, true);  // inverse pass Ls = 2
  dft16<true>(v);
  t0 = pB[0];
  t1 = pB[1];
  xchg(2);
  tw_from_base(v, t0, t1, true);  // inverse pass Ls = 32
  dft16<true>(v);
  t0 = pC[0];
  t1 = pC[1];
  xchg(32);
  tw_from_base(v, t0, t1, true);  // inverse pass Ls = 512: lags j + 512q
  dft16<true>(v);
  // lags lo + j + 512 q below lo + lim are kept (stores past lim dropped by
  // the range check; the spectrum carries the 1/M)
  const int lim = (ob.nout - lo) < ob.V ? (ob.nout - lo) : ob.V;
  const buf_rsrc ry = make_rsrc(ob.yr + lo, lim * 8);
#pragma unroll
  for (int q = 0; q < 16; ++q) buf_store2(ry, (j + q * 512) * 8, v[q]);
}

__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4)))
void pc8192_kernel(Pc8192Args pa) {
  using LY = Layout<8192, 1, false>;
  __shared__ float2 lds[LY::LDS_ELEMS];
  const PcArgs &a = pa.a;
  const int64_t g = xcd_remap(blockIdx.x, a.groups);
  const int64_t row = g / a.nblk;
  int64_t ox, oy;
  batch_offsets(a.bi, row, &ox, &oy);
  const OlsBlock ob{a.x + ox, a.y + oy, (int)((g % a.nblk) * a.V), (int)a.n, (int)a.seff,
                    (int)a.nout, (int)a.V, a.wc};
  ols_block8192(ob, pa.bA, pa.bB, pa.bC, lds);
}

// One workgroup of M/16 threads.  Thread j loads samples j + m*M/16,
// m < 16 (all loads in flight at once), which are exactly the inputs of
// butterfly j of the first radix-16 Stockham pass, so the window, the
// energy and the scaling are applied in registers and pass 0 starts without
// an LDS round trip; the energy is reduced in double (wave shuffles, then
// one LDS slot per wave); the last pass stores the (conjugated) spectrum
// straight to memory.
template <int M>
__device__ __forceinline__ void wfspec_body(const WfArgs &a, float2 *lds, double *red) {
  constexpr int NT = PcCfg<M>::NT;
  constexpr int NWV = NT / 64;
  using LY = Layout<M, 1, false>;
  using SF = Seq<M, false>;
  constexpr int K = SF::n;
  static_assert(SF::template R<0> == 16 && M / 16 == NT, "first pass must be radix 16 over NT butterflies");
  const int tid = threadIdx.x;

  Pass<LY, 16, M, NT, 1> p0;
  double acc = 0.0;
#pragma unroll
  for (int m = 0; m < 16; ++m) {
    const int i = tid + m * NT;
    p0.v[0][m] = keep_if(a.w[(int64_t)(i < a.L ? i : 0) * a.wstride], i < a.L);
  }
#pragma unroll
  for (int m = 0; m < 16; ++m) {
    const int i = tid + m * NT;
    float2 v = p0.v[0][m];
    if (a.window && i < a.L) v = cscale(v, hamming_f32(i, (int)a.L));
    acc += (double)v.x * v.x + (double)v.y * v.y;
    p0.v[0][m] = v;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
  if ((tid & 63) == 0) red[tid >> 6] = acc;
  __syncthreads();
  double tot = 0.0;
#pragma unroll
  for (int q = 0; q < NWV; ++q) tot += red[q];
  if (a.normalize) {
    const float inv = (float)(1.0 / sqrt(tot));
#pragma unroll
    for (int m = 0; m < 16; ++m) p0.v[0][m] = cscale(p0.v[0][m], inv);
  }
  if (a.writeback) {
#pragma unroll
    for (int m = 0; m < 16; ++m) {
      const int i = tid + m * NT;
      if (i < a.L) a.w[(int64_t)i * a.wstride] = p0.v[0][m];
    }
  }
  p0.template compute<1, false>(nullptr);
  p0.template store<1>([&](int t, int i, float2 x) { lds[LY::addr(t, i)] = x; });
  wg_barrier();
  if constexpr (K > 2) lds_passes<LY, SF, M, NT, 1, false, 1, K - 1>(lds, a.tw);
  constexpr int RL = SF::template R<K - 1>;
  constexpr int LSL = SF::template LS<K - 1>;
  Pass<LY, RL, M, NT, 1> pl;
  pl.load([&](int t, int i) { return lds[LY::addr(t, i)]; });
  pl.template compute<LSL, false>(a.tw + SF::template TWOFF<K - 1>);
  const float os = a.out_scale;
  pl.template store<LSL>([&](int t, int i, float2 x) {
    if (a.conj) x.y = -x.y;
    a.wc[i] = cscale(x, os);
  });
}

template <int M>
__global__ __launch_bounds__(PcCfg<M>::NT) void wfspec_kernel(WfArgs a) {
  __shared__ float2 lds[Layout<M, 1, false>::LDS_ELEMS];
  __shared__ double red[PcCfg<M>::NT / 64];
  wfspec_body<M>(a, lds, red);
}

// Ragged batch: one workgroup per job's waveform.
template <int M>
__global__ __launch_bounds__(PcCfg<M>::NT) void wfspec_batch_kernel(const WfArgs *args) {
  __shared__ float2 lds[Layout<M, 1, false>::LDS_ELEMS];
  __shared__ double red[PcCfg<M>::NT / 64];
  wfspec_body<M>(args[blockIdx.x], lds, red);
}

// Ragged batch: the (job, row, block) items of every job of one block size
// (jobs[j].item0 are prefix sums of C*P*nblk).  item_job[it] is the job of
// item it, a table built with the plan (one load, where a search over the
// jobs costs a chain of dependent loads per workgroup); null means a batch
// of one job.  item_job_of is for a workgroup-uniform item.
__device__ __forceinline__ int item_job_of(const int *item_job, int64_t it) {
  return item_job ? __builtin_amdgcn_readfirstlane(item_job[it]) : 0;
}
__device__ __forceinline__ int item_job_lane(const int *item_job, int64_t it) {
  return item_job ? item_job[it] : 0;
}

__device__ __forceinline__ OlsBlock job_block(const PcJob &jb, int64_t it) {
  const int64_t local = it - jb.item0;
  const int64_t row = local / jb.nblk;
  const int b = (int)(local % jb.nblk);
  const int c = (int)(row / jb.P), p = (int)(row % jb.P);
  return OlsBlock{jb.x + c * jb.x_cs + p * jb.x_ps, jb.y + c * jb.y_cs + p * jb.y_ps, b * jb.V,
                  jb.n, jb.seff, jb.nout, jb.V, jb.wc};
}

// PERSIST: a fixed grid of resident workgroups loops over the items;
// otherwise one workgroup per item (the dispatcher does the balancing, and
// no per-pass state stays live across items).
template <int M, bool PERSIST>
__global__ __launch_bounds__(PcCfg<M>::NT) __attribute__((amdgpu_waves_per_eu(4)))
void pc_batch_kernel(const PcJob *jobs, const int *item_job, int64_t items, const float2 *twf,
                     const float2 *twi) {
  __shared__ float2 lds[Layout<M, 1, false>::LDS_ELEMS];
  if constexpr (!PERSIST) {
    const int64_t it = blockIdx.x;
    ols_block<M>(job_block(uniform_copy(&jobs[item_job_of(item_job, it)]), it), twf, twi, lds);
    return;
  }
  for (int64_t it = blockIdx.x; it < items; it += gridDim.x) {
    const PcJob jb = uniform_copy(&jobs[item_job_of(item_job, it)]);
    if (it != blockIdx.x) wg_barrier();  // previous item's LDS reads are done
    ols_block<M>(job_block(jb, it), opaque(twf), opaque(twi), lds);
  }
}

// Ragged batch, small blocks: 4096/M consecutive items per workgroup (items
// of different jobs may share one).
template <int M>
__global__ __launch_bounds__(256) void pc_batch_multi_kernel(const PcJob *jobs, const int *item_job,
                                                             int64_t items, const float2 *twf,
                                                             const float2 *twi) {
  constexpr int TB = 4096 / M;
  __shared__ float2 lds[Layout<M, TB, false>::LDS_ELEMS];
  __shared__ OlsBlock obs[TB];
  if (threadIdx.x < TB) {
    const int64_t it0 = (int64_t)blockIdx.x * TB;
    const int64_t it = it0 + threadIdx.x < items ? it0 + threadIdx.x : it0;
    OlsBlock ob = job_block(jobs[item_job_lane(item_job, it)], it);
    if (it0 + threadIdx.x >= items) {  // past the last item: idle transform
      ob.seff = 0;
      ob.nout = ob.lo;
    }
    obs[threadIdx.x] = ob;
  }
  __syncthreads();
  ols_multi<M>(obs, twf, twi, lds);
}

template <bool PERSIST>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4)))
void pc4096_batch_kernel(const PcJob *jobs, const int *item_job, int64_t items, const float4 *base1,
                         const float4 *base2) {
  __shared__ float2 lds[Layout<4096, 1, false>::LDS_ELEMS];
  if constexpr (!PERSIST) {
    const int64_t it = blockIdx.x;
    ols_block4096(job_block(uniform_copy(&jobs[item_job_of(item_job, it)]), it), base1, base2, lds);
    return;
  }
  for (int64_t it = blockIdx.x; it < items; it += gridDim.x) {
    const PcJob jb = uniform_copy(&jobs[item_job_of(item_job, it)]);
    if (it != blockIdx.x) wg_barrier();
    ols_block4096(job_block(jb, it), opaque(base1), opaque(base2), lds);
  }
}

template <bool PERSIST>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4)))
void pc8192_batch_kernel(const PcJob *jobs, const int *item_job, int64_t items, const float4 *bA,
                         const float4 *bB, const float4 *bC) {
  __shared__ float2 lds[Layout<8192, 1, false>::LDS_ELEMS];
  if constexpr (!PERSIST) {
    const int64_t it = blockIdx.x;
    ols_block8192(job_block(uniform_copy(&jobs[item_job_of(item_job, it)]), it), bA, bB, bC, lds);
    return;
  }
  for (int64_t it = blockIdx.x; it < items; it += gridDim.x) {
    const PcJob jb = uniform_copy(&jobs[item_job_of(item_job, it)]);
    if (it != blockIdx.x) wg_barrier();
    ols_block8192(job_block(jb, it), opaque(bA), opaque(bB), opaque(bC), lds);
  }
}

mxr_status launch_wfspec_batch(int M, const WfArgs *d_args, int njobs, hipStream_t s) {
  if (njobs == 0) return MXR_SUCCESS;
  ProfScope ps("wfspec_batch_kernel", s);
  switch (M) {
#define MXR_WB(MM) \
  case MM: wfspec_batch_kernel<MM><<<dim3(njobs), dim3(PcCfg<MM>::NT), 0, s>>>(d_args); break;
    MXR_WB(1024)
    MXR_WB(2048)
    MXR_WB(4096)
    MXR_WB(8192)
    MXR_WB(16384)
#undef MXR_WB
    default:
      set_error(MXR_NOT_SUPPORTED, "ragged batch: block size %d unsupported", M);
      return MXR_NOT_SUPPORTED;
  }
  MXR_HIP_TRY(hipGetLastError());
  return MXR_SUCCESS;
}

// Persistent grid: as many workgroups as stay resident at once.
int pc_batch_grid(int M) {
  int per_cu = 1;
  hipError_t e = hipSuccess;
  switch (M) {
    case 1024: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, pc_batch_kernel<1024, true>, PcCfg<1024>::NT, 0); break;
    case 2048: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, pc_batch_kernel<2048, true>, PcCfg<2048>::NT, 0); break;
    case 4096: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, pc4096_batch_kernel<true>, 256, 0); break;
    case 8192: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, pc8192_batch_kernel<true>, 512, 0); break;
    case 16384: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, pc_batch_kernel<16384, true>, PcCfg<16384>::NT, 0); break;
    default: break;
  }
  if (e != hipSuccess || per_cu < 1) per_cu = 1;
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device_id()) != hipSuccess || cus < 1)
    cus = 256;
  return per_cu * cus;
}

// grid == items: one workgroup per item; otherwise the persistent form.
mxr_status launch_pc_batch(int M, const PcJob *d_jobs, const int *d_item_job, int64_t items, int grid,
                           hipStream_t s) {
  if (items == 0) return MXR_SUCCESS;
  MXR_REQUIRE(grid >= 1 && grid <= items && items <= 0x7fffffff, MXR_INVALID_SIZE,
              "ragged batch: bad grid %d for %lld items", grid, (long long)items);
  const bool persist = grid < items;
  const float2 *twf = twiddles_device(M, false), *twi = twiddles_device(M, true);
  MXR_REQUIRE(twf && twi, MXR_OUT_OF_MEMORY, "ragged batch: twiddle table allocation failed");
  ProfScope ps("pc_batch_kernel", s);
  switch (M) {
    case 4096: {
      const float4 *b1 = pow_base_device(4096, 16), *b2 = pow_base_device(4096, 256);
      MXR_REQUIRE(b1 && b2, MXR_OUT_OF_MEMORY, "ragged batch: twiddle table allocation failed");
      if (persist) pc4096_batch_kernel<true><<<dim3(grid), dim3(256), 0, s>>>(d_jobs, d_item_job, items, b1, b2);
      else pc4096_batch_kernel<false><<<dim3(grid), dim3(256), 0, s>>>(d_jobs, d_item_job, items, b1, b2);
      break;
    }
#define MXR_PB(MM)                                                                               \
  case MM:                                                                                       \
    if (persist)                                                                                 \
      pc_batch_kernel<MM, true><<<dim3(grid), dim3(PcCfg<MM>::NT), 0, s>>>(d_jobs, d_item_job, items, twf, twi); \
    else                                                                                         \
      pc_batch_kernel<MM, false><<<dim3(grid), dim3(PcCfg<MM>::NT), 0, s>>>(d_jobs, d_item_job, items, twf, twi); \
    break;
#define MXR_PBM(MM)                                                                              \
  case MM:                                                                                       \
    if (persist)                                                                                 \
      pc_batch_kernel<MM, true><<<dim3(grid), dim3(PcCfg<MM>::NT), 0, s>>>(d_jobs, d_item_job, items, twf, twi); \
    else                                                                                         \
      pc_batch_multi_kernel<MM><<<dim3((unsigned)((items + 4096 / MM - 1) / (4096 / MM))), dim3(256), 0, s>>>( \
          d_jobs, d_item_job, items, twf, twi);                                                       \
    break;
    case 8192: {
      const float4 *bA = pow_base_device(8192, 2), *bB = pow_base_device(8192, 32),
                   *bC = pow_base_device(8192, 512);
      MXR_REQUIRE(bA && bB && bC, MXR_OUT_OF_MEMORY, "ragged batch: twiddle table allocation failed");
      if (persist) pc8192_batch_kernel<true><<<dim3(grid), dim3(512), 0, s>>>(d_jobs, d_item_job, items, bA, bB, bC);
      else pc8192_batch_kernel<false><<<dim3(grid), dim3(512), 0, s>>>(d_jobs, d_item_job, items, bA, bB, bC);
      break;
    }
    MXR_PBM(1024)
    MXR_PBM(2048)
    MXR_PB(16384)
#undef MXR_PB
#undef MXR_PBM
    default:
      set_error(MXR_NOT_SUPPORTED, "ragged batch: block size %d unsupported", M);
      return MXR_NOT_SUPPORTED;
  }
  MXR_HIP_TRY(hipGetLastError());
  return MXR_SUCCESS;
}

// ------------------------------------------------------------------ host
namespace {

struct PcPlan {
  int M;
  int64_t V, nblk;
};

int ilog2(int64_t v) {
  int l = 0;
  while ((int64_t(1) << l) < v) ++l;
  return l;
}

// Forced block size for experiments/tuning: MXR_PC_BLOCK=<M> (0 = auto).
int forced_block() {
  static int v = [] {
    const char *e = getenv("MXR_PC_BLOCK");
    return e ? atoi(e) : 0;
  }();
  return v;
}

// Chooses the overlap-save block size.  allow_ols is false for convolution
// (conj_w = 0), where only the single n-point block is exact.
mxr_status plan_pc(int64_t seff, int64_t L, int64_t n, int64_t nout, bool allow_ols, bool in_place,
                   PcPlan *out, int force_M = 0) {
  PcPlan best{0, 0, 0};
  double best_cost = 1e300;
  auto consider = [&](int M) {
    PcPlan p;
    p.M = M;
    if (M == n) {
      p.V = nout;
      p.nblk = 1;
    } else {
      if (M < L || M > n) return;
      p.V = M - L + 1;
      // keep whole 64-lag groups per block so every block's stores start on
      // a 512-byte boundary of an aligned row, unless that costs a block
      if (p.V >= 64 && (nout + p.V - 1) / p.V == (nout + (p.V / 64) * 64 - 1) / ((p.V / 64) * 64))
        p.V = (p.V / 64) * 64;
      p.nblk = (nout + p.V - 1) / p.V;
      if (in_place) {
        // a block that wraps past n would read lags already overwritten
        const int64_t last_lo = (p.nblk - 1) * p.V;
        if (last_lo + M > n && (last_lo + M - n) > 0 && seff > 0) return;
      }
    }
    // relative cost per point: 1024..8192 have dedicated or multi-block
    // kernels with several workgroups per CU; the 16384-point kernel holds
    // one workgroup per CU and runs about half as fast per point (measured)
    const double eff = M <= 8192 ? 1.0 : (M == 16384 ? 2.0 : 1e9);
    const double cost = (double)p.nblk * M * ilog2(M) * eff;
    if (cost < best_cost || (cost == best_cost && M > best.M)) {
      best_cost = cost;
      best = p;
    }
  };
  const int fb = force_M > 0 ? force_M : forced_block();
  if (fb > 0) {
    consider(fb);
  } else {
    if (n <= 16384) consider((int)n);
    if (allow_ols)
      for (int M = 1024; M <= 16384 && M < n; M *= 2) consider(M);
  }
  MXR_REQUIRE(best.M > 0 && best.M <= 16384, MXR_NOT_SUPPORTED,
              "pulse compression: no block size fits (S=%lld, L=%lld, n=%lld, nout=%lld%s)",
              (long long)seff, (long long)L, (long long)n, (long long)nout,
              in_place ? ", in place" : "");
  *out = best;
  return MXR_SUCCESS;
}

// Per (device, stream, M) waveform-spectrum buffer: calls on different
// streams never share one (cf. per-stream workspaces, cache.h:261-291).
std::mutex g_wc_mu;
std::map<std::tuple<int, uintptr_t, int>, float2 *> g_wc;

float2 *wc_buffer(int M, hipStream_t s) {
  const auto key = std::make_tuple(device_id(), (uintptr_t)s, M);
  std::lock_guard<std::mutex> g(g_wc_mu);
  auto it = g_wc.find(key);
  if (it != g_wc.end()) return it->second;
  float2 *p = nullptr;
  if (hipMalloc(&p, sizeof(float2) * M) != hipSuccess) return nullptr;
  g_wc[key] = p;
  return p;
}

void clear_wc() {
  std::lock_guard<std::mutex> g(g_wc_mu);
  for (auto &kv : g_wc) {
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(std::get<0>(kv.first));
    (void)hipFree(kv.second);
    (void)hipSetDevice(cur);
  }
  g_wc.clear();
}
CacheClearRegistrar g_reg_wc(clear_wc);

// Either part may be skipped: wa == nullptr keeps the spectrum already in
// the buffer (a plan's), pa == nullptr only (re)computes the spectrum.
// Grids up to this many workgroups take pc4096_wf_kernel when the spectrum
// is computed in the same call (two resident workgroups per CU).
constexpr int64_t PC_WF_FUSE_MAX_GROUPS = 512;

template <int M>
mxr_status launch_pc(const WfArgs *wa, const PcArgs *pa, hipStream_t s) {
  if constexpr (M == 4096) {
    if (wa && pa && pa->groups > 0 && pa->bpw == 1 && pa->groups <= PC_WF_FUSE_MAX_GROUPS &&
        !wa->window && !wa->normalize && !wa->writeback && wa->L >= 1 && wa->L <= 4096) {
      Pc4096WfArgs q{{*pa, pow_base_device(4096, 16), pow_base_device(4096, 256)}, wa->w, wa->wstride,
                     (int)wa->L, wa->conj, wa->out_scale};
      MXR_REQUIRE(q.p.base1 && q.p.base2, MXR_OUT_OF_MEMORY,
                  "pulse compression: twiddle table allocation failed");
      ProfScope ps("pc4096_wf_kernel", s);
      pc4096_wf_kernel<<<dim3((unsigned)pa->groups), dim3(512), 0, s>>>(q);
      MXR_HIP_TRY(hipGetLastError());
      return MXR_SUCCESS;
    }
  }
  if (wa) {
    ProfScope ps("wfspec_kernel", s);
    wfspec_kernel<M><<<dim3(1), dim3(PcCfg<M>::NT), 0, s>>>(*wa);
  }
  MXR_HIP_TRY(hipGetLastError());
  if (!pa || pa->groups <= 0) return MXR_SUCCESS;
  if constexpr (M == 8192) {
    if (pa->bpw == 1) {
      Pc8192Args p8{*pa, pow_base_device(8192, 2), pow_base_device(8192, 32),
                    pow_base_device(8192, 512)};
      MXR_REQUIRE(p8.bA && p8.bB && p8.bC, MXR_OUT_OF_MEMORY,
                  "pulse compression: twiddle table allocation failed");
      ProfScope ps("pc8192_kernel", s);
      pc8192_kernel<<<dim3((unsigned)pa->groups), dim3(512), 0, s>>>(p8);
      MXR_HIP_TRY(hipGetLastError());
      return MXR_SUCCESS;
    }
  }
  if constexpr (M == 4096) {
    if (pa->bpw == 1) {
      Pc4096Args p4{*pa, pow_base_device(4096, 16), pow_base_device(4096, 256)};
      MXR_REQUIRE(p4.base1 && p4.base2, MXR_OUT_OF_MEMORY, "pulse compression: twiddle table allocation failed");
      ProfScope ps("pc4096_kernel", s);
      pc4096_kernel<<<dim3((unsigned)pa->groups), dim3(256), 0, s>>>(p4);
      MXR_HIP_TRY(hipGetLastError());
      return MXR_SUCCESS;
    }
  }
  {
    ProfScope ps("pc_ols_kernel", s);
    if (pa->bpw > 1) {
      pc_ols_kernel<M, true><<<dim3((unsigned)pa->groups), dim3(PcCfg<M>::NT), 0, s>>>(*pa);
    } else if constexpr (M <= 2048) {
      constexpr int TB = 4096 / M;
      pc_ols_multi_kernel<M><<<dim3((unsigned)((pa->groups + TB - 1) / TB)), dim3(256), 0, s>>>(*pa);
    } else {
      pc_ols_kernel<M, false><<<dim3((unsigned)pa->groups), dim3(PcCfg<M>::NT), 0, s>>>(*pa);
    }
  }
  MXR_HIP_TRY(hipGetLastError());
  return MXR_SUCCESS;
}

mxr_status dispatch_pc(int M, const WfArgs *wa, const PcArgs *pa, hipStream_t s) {
  switch (M) {
    case 1024: return launch_pc<1024>(wa, pa, s);
    case 2048: return launch_pc<2048>(wa, pa, s);
    case 4096: return launch_pc<4096>(wa, pa, s);
    case 8192: return launch_pc<8192>(wa, pa, s);
    case 16384: return launch_pc<16384>(wa, pa, s);
    default: break;
  }
  set_error(MXR_NOT_SUPPORTED, "pulse compression: block size %d unsupported (1024..16384)", M);
  return MXR_NOT_SUPPORTED;
}

}  // namespace

// Builds every table the block kernels of size M read (forward and reversed
// pass twiddles, the power-base rows of the dedicated 4096 and 8192 kernels),
// so that later launches allocate nothing (and can be graph-captured).
mxr_status pc_tables_prepare(int M, const char *fn) {
  bool ok = twiddles_device(M, false) && twiddles_device(M, true);
  if (M == 4096) ok = ok && pow_base_device(4096, 16) && pow_base_device(4096, 256);
  if (M == 8192)
    ok = ok && pow_base_device(8192, 2) && pow_base_device(8192, 32) && pow_base_device(8192, 512);
  MXR_REQUIRE(ok, MXR_OUT_OF_MEMORY, "%s: twiddle table allocation failed", fn);
  return MXR_SUCCESS;
}

// Overlap-save block plan of an out-of-place correlation (the planner's
// choice, or the given block size M_req when it fits).
mxr_status pc_plan_blocks(int64_t seff, int64_t L, int64_t n, int64_t nout, int M_req, int *M,
                          int64_t *V, int64_t *nblk) {
  PcPlan p;
  mxr_status st = plan_pc(seff, L, n, nout, true, false, &p, M_req);
  if (st) return st;
  *M = p.M;
  *V = p.V;
  *nblk = p.nblk;
  return MXR_SUCCESS;
}

// Validates the signal/output pair of a pulse-compression call and fills
// the row fields of PcArgs (x, y, rows, batch strides, seff, n, nout).
static mxr_status pc_rows(const char *fn, const mxr_tensor *y, const mxr_tensor *x, int64_t n,
                          PcArgs *pa, bool *in_place) {
  mxr_status st;
  if ((st = check_tensor(y, "output", MXR_C64, 1, MXR_MAX_RANK))) return st;
  if ((st = check_tensor(x, "input", MXR_C64, 1, MXR_MAX_RANK))) return st;
  MXR_REQUIRE(x->rank == y->rank, MXR_INVALID_DIM, "%s: input rank %d != output rank %d", fn,
              x->rank, y->rank);
  const int rank = x->rank;
  for (int d = 0; d < rank - 1; ++d)
    MXR_REQUIRE(x->shape[d] == y->shape[d], MXR_INVALID_SIZE, "%s: batch dim %d differs", fn, d);
  MXR_REQUIRE(x->stride[rank - 1] == 1 && y->stride[rank - 1] == 1, MXR_INVALID_PARAMETER,
              "%s: the sample axis must have unit stride", fn);
  MXR_REQUIRE(n > 0 && (n & (n - 1)) == 0, MXR_NOT_SUPPORTED,
              "%s: transform size %lld must be a power of two", fn, (long long)n);
  const int64_t S = x->shape[rank - 1];
  const int64_t nout = y->shape[rank - 1];
  MXR_REQUIRE(nout <= n, MXR_INVALID_SIZE, "%s: %lld output lags exceed the transform size %lld", fn,
              (long long)nout, (long long)n);
  *in_place = false;
  if (views_overlap(*x, *y)) {
    bool same_rows = x->data == y->data;
    for (int d = 0; d < rank - 1; ++d) same_rows = same_rows && x->stride[d] == y->stride[d];
    MXR_REQUIRE(same_rows, MXR_ALIAS_ERROR,
                "%s: output overlaps input without being the same rows (in place)", fn);
    *in_place = true;
  }
  pa->x = (const float2 *)x->data;
  pa->y = (float2 *)y->data;
  pa->rows = 1;
  pa->bi.nd = rank - 1;
  for (int d = 0; d < rank - 1; ++d) {
    pa->bi.shape[d] = x->shape[d];
    pa->bi.sa[d] = x->stride[d];
    pa->bi.sb[d] = y->stride[d];
    pa->rows *= x->shape[d];
  }
  pa->seff = S < n ? S : n;
  pa->n = n;
  pa->nout = nout;
  return MXR_SUCCESS;
}

// Block fields of PcArgs from a plan.
static mxr_status pc_blocks(const char *fn, const PcPlan &plan, bool in_place, const float2 *wc,
                            const float2 *twf, const float2 *twi, PcArgs *pa) {
  pa->V = plan.V;
  pa->nblk = plan.nblk;
  pa->bpw = in_place ? plan.nblk : 1;
  const int64_t gpr = (pa->nblk + pa->bpw - 1) / pa->bpw;
  pa->groups = pa->nout > 0 ? pa->rows * gpr : 0;
  MXR_REQUIRE(pa->groups <= 0x7fffffff, MXR_INVALID_SIZE, "%s: too many rows", fn);
  pa->wc = wc;
  pa->twf = twf;
  pa->twi = twi;
  return MXR_SUCCESS;
}

mxr_status fft_conv_rows(const mxr_tensor *out, const mxr_tensor *x, const mxr_tensor *w,
                         int64_t n, int conj_w, hipStream_t s);

// Shared body of mxr_pulse_compress and mxr_fft_conv.
mxr_status pc_common(const char *fn, const mxr_tensor *y, const mxr_tensor *x, const mxr_tensor *w,
                     int64_t n, int window, int normalize, int writeback, int conj_w,
                     hipStream_t stream) {
  mxr_status st;
  if ((st = check_tensor(w, "waveform", MXR_C64, 1, 1))) return st;
  PcArgs pa{};
  bool in_place = false;
  if ((st = pc_rows(fn, y, x, n, &pa, &in_place))) return st;
  const int64_t L = w->shape[0];
  MXR_REQUIRE(L >= 1 && L <= n, MXR_INVALID_SIZE, "%s: waveform length %lld not in [1, n=%lld]", fn,
              (long long)L, (long long)n);
  MXR_REQUIRE(!views_overlap(*w, *x) && !views_overlap(*w, *y), MXR_ALIAS_ERROR,
              "%s: waveform aliases the signal", fn);
  PcPlan plan;
  if ((st = plan_pc(pa.seff, L, n, pa.nout, conj_w != 0, in_place, &plan))) return st;
  MXR_REQUIRE(L <= plan.M, MXR_NOT_SUPPORTED, "%s: waveform longer than the block", fn);

  float2 *wc = wc_buffer(plan.M, stream);
  MXR_REQUIRE(wc != nullptr, MXR_OUT_OF_MEMORY, "%s: spectrum buffer allocation failed", fn);
  const float2 *twf = twiddles_device(plan.M, false);
  const float2 *twi = twiddles_device(plan.M, true);
  MXR_REQUIRE(twf && twi, MXR_OUT_OF_MEMORY, "%s: twiddle table allocation failed", fn);

  WfArgs wa{};
  wa.w = (float2 *)w->data;
  wa.wstride = w->stride[0];
  wa.L = L;
  wa.window = window;
  wa.normalize = normalize;
  wa.writeback = writeback;
  wa.conj = conj_w;
  wa.wc = wc;
  wa.tw = twf;
  wa.out_scale = (float)(1.0 / (double)plan.M);  // exact: a power of two
  if ((st = pc_blocks(fn, plan, in_place, wc, twf, twi, &pa))) return st;
  return dispatch_pc(plan.M, &wa, &pa, stream);
}

// --------------------------------------------------------------- plans
// A plan fixes (S, L, n, nout, flags), picks the block size once and owns
// the M-point spectrum of one waveform, so a stream of cubes filtered by the
// same waveform runs only the block kernel per cube.
struct PcPlanObj {
  int dev;
  int64_t S, L, n, nout;
  int flags;
  PcPlan plan;
  float2 *wc;
  const float2 *twf, *twi;
};

static mxr_status pc_plan_create(mxr_pc_plan *out, int64_t S, int64_t L, int64_t n, int64_t nout,
                                 int flags) {
  MXR_REQUIRE(out != nullptr, MXR_INVALID_PARAMETER, "pc_plan_create: null plan pointer");
  *out = nullptr;
  MXR_REQUIRE(n > 0 && (n & (n - 1)) == 0, MXR_NOT_SUPPORTED,
              "pc_plan_create: transform size %lld must be a power of two", (long long)n);
  MXR_REQUIRE(S >= 1 && L >= 1 && L <= n && nout >= 0 && nout <= n, MXR_INVALID_SIZE,
              "pc_plan_create: need S >= 1, 1 <= L <= n and 0 <= nout <= n (S=%lld L=%lld "
              "n=%lld nout=%lld)", (long long)S, (long long)L, (long long)n, (long long)nout);
  MXR_REQUIRE((flags & ~(MXR_PC_WINDOW | MXR_PC_NORMALIZE)) == 0, MXR_INVALID_PARAMETER,
              "pc_plan_create: flags may hold only MXR_PC_WINDOW | MXR_PC_NORMALIZE");
  PcPlan plan;
  mxr_status st;
  if ((st = plan_pc(S < n ? S : n, L, n, nout, true, false, &plan))) return st;
  MXR_REQUIRE(L <= plan.M, MXR_NOT_SUPPORTED, "pc_plan_create: waveform longer than the block");
  if ((st = pc_tables_prepare(plan.M, "pc_plan_create"))) return st;
  const float2 *twf = twiddles_device(plan.M, false);
  const float2 *twi = twiddles_device(plan.M, true);
  float2 *wc = nullptr;
  MXR_HIP_TRY(hipMalloc(&wc, sizeof(float2) * plan.M));
  plan_tables_retain();
  *out = reinterpret_cast<mxr_pc_plan>(new PcPlanObj{device_id(), S, L, n, nout, flags, plan, wc, twf, twi});
  log_msg(LOG_DEBUG, "pc plan: S=%lld L=%lld n=%lld -> %d-point blocks, %lld per row", (long long)S,
          (long long)L, (long long)n, plan.M, (long long)plan.nblk);
  return MXR_SUCCESS;
}

static mxr_status plan_on_device(const PcPlanObj *p, const char *fn) {
  MXR_REQUIRE(p != nullptr, MXR_INVALID_PARAMETER, "%s: null plan", fn);
  MXR_REQUIRE(device_id() == p->dev, MXR_INVALID_PARAMETER,
              "%s: plan was made on device %d, current device is %d", fn, p->dev, device_id());
  return MXR_SUCCESS;
}

static mxr_status pc_plan_set_waveform(mxr_pc_plan h, const mxr_tensor *w, hipStream_t s) {
  const PcPlanObj *p = reinterpret_cast<const PcPlanObj *>(h);
  mxr_status st;
  if ((st = plan_on_device(p, "pc_plan_set_waveform"))) return st;
  if ((st = check_tensor(w, "waveform", MXR_C64, 1, 1))) return st;
  MXR_REQUIRE(w->shape[0] == p->L, MXR_INVALID_SIZE,
              "pc_plan_set_waveform: waveform has %lld samples, the plan %lld",
              (long long)w->shape[0], (long long)p->L);
  WfArgs wa{};
  wa.w = (float2 *)w->data;
  wa.wstride = w->stride[0];
  wa.L = p->L;
  wa.window = (p->flags & MXR_PC_WINDOW) != 0;
  wa.normalize = (p->flags & MXR_PC_NORMALIZE) != 0;
  wa.writeback = 0;
  wa.conj = 1;
  wa.wc = p->wc;
  wa.tw = p->twf;
  wa.out_scale = (float)(1.0 / (double)p->plan.M);
  return dispatch_pc(p->plan.M, &wa, nullptr, s);
}

static mxr_status pc_plan_execute(mxr_pc_plan h, const mxr_tensor *y, const mxr_tensor *x,
                                  hipStream_t s) {
  const PcPlanObj *p = reinterpret_cast<const PcPlanObj *>(h);
  mxr_status st;
  if ((st = plan_on_device(p, "pc_plan_execute"))) return st;
  PcArgs pa{};
  bool in_place = false;
  if ((st = pc_rows("pc_plan_execute", y, x, p->n, &pa, &in_place))) return st;
  MXR_REQUIRE(!in_place, MXR_ALIAS_ERROR,
              "pc_plan_execute: output overlaps the input (use mxr_pulse_compress in place)");
  MXR_REQUIRE(x->shape[x->rank - 1] == p->S && pa.nout == p->nout, MXR_INVALID_SIZE,
              "pc_plan_execute: rows of %lld samples -> %lld lags, the plan has %lld -> %lld",
              (long long)x->shape[x->rank - 1], (long long)pa.nout, (long long)p->S,
              (long long)p->nout);
  if ((st = pc_blocks("pc_plan_execute", p->plan, false, p->wc, p->twf, p->twi, &pa))) return st;
  return dispatch_pc(p->plan.M, nullptr, &pa, s);
}

static mxr_status pc_plan_waveform_grad(mxr_pc_plan h, const mxr_tensor *gw, const mxr_tensor *gy,
                                        const mxr_tensor *x, const mxr_tensor *w, hipStream_t s) {
  const char *fn = "pc_plan_waveform_grad";
  const PcPlanObj *p = reinterpret_cast<const PcPlanObj *>(h);
  mxr_status st;
  if ((st = plan_on_device(p, fn))) return st;
  PcArgs pa{};
  bool in_place = false;
  if ((st = pc_rows(fn, gy, x, p->n, &pa, &in_place))) return st;
  MXR_REQUIRE(!in_place, MXR_ALIAS_ERROR, "%s: gy overlaps x", fn);
  MXR_REQUIRE(x->shape[x->rank - 1] == p->S && pa.nout == p->nout, MXR_INVALID_SIZE,
              "%s: rows of %lld samples -> %lld lags, the plan has %lld -> %lld", fn,
              (long long)x->shape[x->rank - 1], (long long)pa.nout, (long long)p->S,
              (long long)p->nout);
  if ((st = check_tensor(w, "waveform", MXR_C64, 1, 1))) return st;
  if ((st = check_tensor(gw, "waveform gradient", MXR_C64, 1, 1))) return st;
  MXR_REQUIRE(w->shape[0] == p->L && gw->shape[0] == p->L, MXR_INVALID_SIZE,
              "%s: waveform and its gradient need %lld samples", fn, (long long)p->L);
  for (const mxr_tensor *i : {x, gy, w})
    MXR_REQUIRE(!views_overlap(*gw, *i), MXR_ALIAS_ERROR, "%s: gw overlaps an input", fn);
  MXR_REQUIRE(pa.rows <= 0x7fffffffll, MXR_INVALID_SIZE, "%s: too many rows", fn);
  WgradJob j{};
  j.x = pa.x;
  j.gy = pa.y;
  j.w = (const float2 *)w->data;
  j.gw = (float2 *)gw->data;
  j.wstride = w->stride[0];
  j.gwstride = gw->stride[0];
  j.bi = pa.bi;
  j.rows = pa.nout > 0 ? pa.rows : 0;
  j.V = p->plan.V;
  j.nblk = p->plan.nblk;
  j.M = p->plan.M;
  if (j.M > 8192 && p->L <= 8192) {
    // two spectra and their running product stay in registers: at 16384
    // points (1024 threads) they no longer fit, so use 8192-point blocks
    int M2 = 0;
    if ((st = pc_plan_blocks(pa.seff, p->L, p->n, pa.nout, 8192, &M2, &j.V, &j.nblk))) return st;
    j.M = M2;
  }
  j.n = (int)p->n;
  j.seff = (int)pa.seff;
  j.nout = (int)pa.nout;
  j.L = (int)p->L;
  j.window = (p->flags & MXR_PC_WINDOW) != 0;
  j.normalize = (p->flags & MXR_PC_NORMALIZE) != 0;
  return pc_waveform_grad_run(j, s);
}

static mxr_status pc_plan_destroy(mxr_pc_plan h) {
  PcPlanObj *p = reinterpret_cast<PcPlanObj *>(h);
  if (!p) return MXR_SUCCESS;
  int cur = 0;
  (void)hipGetDevice(&cur);
  (void)hipSetDevice(p->dev);
  const hipError_t e = hipFree(p->wc);
  (void)hipSetDevice(cur);
  delete p;
  plan_tables_release();
  MXR_HIP_TRY(e);
  return MXR_SUCCESS;
}

}  // namespace mxr

extern "C" {

mxr_status mxr_pulse_compress(const mxr_tensor *y, const mxr_tensor *x, const mxr_tensor *w,
                              int64_t n, int flags, hipStream_t stream) {
  return mxr::guarded("mxr_pulse_compress", [&] {
    return mxr::pc_common("pulse_compress", y, x, w, n, (flags & MXR_PC_WINDOW) != 0,
                          (flags & MXR_PC_NORMALIZE) != 0, (flags & MXR_PC_WRITEBACK) != 0, 1,
                          stream);
  });
}

mxr_status mxr_fft_conv(const mxr_tensor *out, const mxr_tensor *x, const mxr_tensor *w, int64_t n,
                        int conj_w, hipStream_t stream) {
  return mxr::guarded("mxr_fft_conv", [&] {
    if (w && w->rank > 1) return mxr::fft_conv_rows(out, x, w, n, conj_w ? 1 : 0, stream);
    return mxr::pc_common("fft_conv", out, x, w, n, 0, 0, 0, conj_w ? 1 : 0, stream);
  });
}

mxr_status mxr_pc_plan_create(mxr_pc_plan *plan, int64_t S, int64_t L, int64_t n, int64_t nout,
                              int flags) {
  return mxr::guarded("mxr_pc_plan_create",
                      [&] { return mxr::pc_plan_create(plan, S, L, n, nout, flags); });
}

mxr_status mxr_pc_plan_set_waveform(mxr_pc_plan plan, const mxr_tensor *w, hipStream_t stream) {
  return mxr::guarded("mxr_pc_plan_set_waveform",
                      [&] { return mxr::pc_plan_set_waveform(plan, w, stream); });
}

mxr_status mxr_pc_plan_execute(mxr_pc_plan plan, const mxr_tensor *y, const mxr_tensor *x,
                               hipStream_t stream) {
  return mxr::guarded("mxr_pc_plan_execute",
                      [&] { return mxr::pc_plan_execute(plan, y, x, stream); });
}

mxr_status mxr_pc_plan_waveform_grad(mxr_pc_plan plan, const mxr_tensor *gw, const mxr_tensor *gy,
                                     const mxr_tensor *x, const mxr_tensor *w, hipStream_t stream) {
  return mxr::guarded("mxr_pc_plan_waveform_grad",
                      [&] { return mxr::pc_plan_waveform_grad(plan, gw, gy, x, w, stream); });
}

mxr_status mxr_pc_plan_destroy(mxr_pc_plan plan) {
  return mxr::guarded("mxr_pc_plan_destroy", [&] { return mxr::pc_plan_destroy(plan); });
}

}  // extern "C"
