// Pulse-axis pass: three-pulse canceller [1,-2,1] -> Doppler window ->
// nd-point Doppler FFT along pulses -> |.|^2, fused in one kernel
// (replaces ThreePulseCanceller + DopplerProcessing + the abs2 of
// CFARDetections, examples/simple_radar_pipeline.h:321-357,398).
//
// A workgroup owns a tile of TR consecutive range bins of one channel across
// all P pulses.  It loads the tile with coalesced TR-wide row reads into LDS,
// evaluates the canceller and the window while feeding the first Stockham
// pass, runs the remaining passes down the LDS columns (element (pulse i,
// range t) at i*TR + t, so consecutive lanes touch consecutive words), and
// writes the range-Doppler map and/or its power with TR-wide coalesced rows.
// HBM traffic per tile: P*TR*8 bytes in, nd*TR*(4 and/or 8) bytes out.
#include "fft_engine.hpp"
#include "mxr_cache.hpp"
#include "ragged.hpp"

namespace mxr {

template <int ND>
struct DopCfg {
  // 512 threads from 512 Doppler bins up: one radix-16 butterfly per thread
  // per pass instead of two (12 % faster forward at 512 x 16 tiles)
  static constexpr int NT = ND >= 512 ? 512 : 256;
  // range bins per tile (measured on MI355X): 32 (256-byte rows, 32 KiB of
  // LDS at 128 Doppler bins) up to 128; 16 (one 128-byte line per pulse row)
  // at 256 and 512, where 32 would hold a 64 KiB tile and halve the resident
  // workgroups (40% slower at 256); 8 at 1024.  Narrower tiles waste part of
  // every 128-byte line they touch.
  static constexpr int TR = ND <= 128 ? 32 : (ND <= 512 ? 16 : 8);
};

// Coalesced load of a pulse tile: rows p < P of TR range bins from r0 into
// lds[p * TR + t] (zero past R), 16 loads in flight per lane.
template <int ND, int TR = DopCfg<ND>::TR, int NT = DopCfg<ND>::NT>
__device__ __forceinline__ void load_pulse_tile(const float2 *yc, int64_t ys_p, int64_t ys_r,
                                                int P, int R, int r0, float2 *lds) {
  constexpr int RPI = NT / TR;  // pulses covered by one sweep of the workgroup
  constexpr int BATCH = 16;
  const int t = threadIdx.x % TR, p0 = threadIdx.x / TR;
  const int r = r0 + t;
  const bool rok = r < R;
  const float2 *col = yc + (int64_t)r * ys_r;
  for (int pb = 0; pb < P; pb += RPI * BATCH) {
    float2 v[BATCH];
#pragma unroll
    for (int u = 0; u < BATCH; ++u) {
      const int p = pb + p0 + u * RPI;
      v[u] = (rok && p < P) ? col[(int64_t)p * ys_p] : make_float2(0.f, 0.f);
    }
#pragma unroll
    for (int u = 0; u < BATCH; ++u) {
      const int p = pb + p0 + u * RPI;
      if (p < P) lds[p * TR + t] = v[u];
    }
  }
}

// One tile: TR range bins from r0 of channel c.  lds holds rows_buf * TR
// elements, win ND floats.  Pre: the workgroup's previous use of both is
// finished.
template <int ND, int TR = DopCfg<ND>::TR, int NT = DopCfg<ND>::NT>
__device__ __forceinline__ void doppler_tile(const DopArgs &a, int c, int r0, int rows_buf,
                                             float2 *lds, float *win) {
  using LY = Layout<ND, TR, true>;
  using SQ = Seq<ND, false>;
  constexpr int K = SQ::n;
  for (int i = threadIdx.x; i < ND; i += NT) {
    float h = 0.f;
    if (i < a.rows_t)
      h = a.taps ? a.taps[(int64_t)i * a.taps_s] : (i < a.win_rows ? hamming_f32(i, a.win_rows) : 1.0f);
    win[i] = h;
  }
  load_pulse_tile<ND, TR, NT>(a.y + (int64_t)c * a.ys_c, a.ys_p, a.ys_r, a.P, a.R, r0, lds);
  wg_barrier();

  auto tpc = [&](int t, int i) -> float2 {
    if (i >= a.rows_t) return make_float2(0.f, 0.f);
    const int p = i + a.p_off;
    float2 m = lds[p * TR + t];
    if (a.no_tpc) return cscale(m, win[i]);
    float2 s = p >= 1 ? lds[(p - 1) * TR + t] : make_float2(0.f, 0.f);
    float2 u = p + 1 < a.P ? lds[(p + 1) * TR + t] : make_float2(0.f, 0.f);
    float2 v = make_float2(s.x + u.x - 2.0f * m.x, s.y + u.y - 2.0f * m.y);
    return cscale(v, win[i]);
  };
  auto store_out = [&](int t, int k, float2 v) {
    const int r = r0 + t;
    if (r >= a.R) return;
    if (a.rd) a.rd[(int64_t)c * a.rs_c + (int64_t)k * a.rs_k + (int64_t)r * a.rs_r] = v;
    if (a.pw) a.pw[(int64_t)c * a.ps_c + (int64_t)k * a.ps_k + (int64_t)r * a.ps_r] = cabs2(v);
  };

  constexpr int R0 = SQ::template R<0>;
  Pass<LY, R0, ND, NT, TR> p0;
  p0.load(tpc);
  if constexpr (K == 1) {
    p0.template compute<1, false>(nullptr);
    p0.template store<1>(store_out);
  } else {
    wg_barrier();  // every canceller read is done before the tile is overwritten
    p0.template compute<1, false>(nullptr);
    p0.template store<1>([&](int t, int i, float2 v) { lds[LY::addr(t, i)] = v; });
    wg_barrier();
    lds_passes<LY, SQ, ND, NT, TR, false, 1, K - 1>(lds, a.tw);
    constexpr int RL = SQ::template R<K - 1>;
    constexpr int LSL = SQ::template LS<K - 1>;
    Pass<LY, RL, ND, NT, TR> pl;
    pl.load([&](int t, int i) { return lds[LY::addr(t, i)]; });
    pl.template compute<LSL, false>(a.tw + SQ::template TWOFF<K - 1>);
    pl.template store<LSL>(store_out);
  }
}

template <int ND>
__global__ __launch_bounds__(DopCfg<ND>::NT) void doppler_kernel(DopArgs a, int rows_buf) {
  extern __shared__ float2 lds[];  // rows_buf * TR
  __shared__ float win[ND];
  doppler_tile<ND>(a, blockIdx.y, blockIdx.x * DopCfg<ND>::TR, rows_buf, lds, win);
}

// Ragged batch: the (job, channel, tile) items of every job of one Doppler
// size (jobs[j].item0 are prefix sums of C*tiles); item_job[it] is the job
// of item it (a table built with the plan).  PERSIST: a resident grid loops
// over the items; otherwise one workgroup per item.
// Batch tiles hold 32 KiB of pulses where that fits: 128 range bins at 16
// and 32 Doppler bins, 64 at 64, 32 at 128 and 256 (the per-item decode, a
// job lookup and a record copy, is paid per tile, and wider rows read whole
// 512- and 1024-byte runs per pulse).  At 256 the 32-wide tile measured
// faster in the batch although the single call prefers 16; 64 at 128 (a
// 64 KiB tile) measured slower.
template <int ND>
struct DopBatchCfg {
  static constexpr int TR = ND <= 32 ? 128 : (ND <= 64 ? 64 : (ND <= 256 ? 32 : DopCfg<ND>::TR));
  // 512 threads once the first radix-16 pass has 512 butterflies (256 x 32)
  static constexpr int NT = ND >= 256 ? 512 : 256;
};

template <int ND, bool PERSIST>
__global__ __launch_bounds__(DopBatchCfg<ND>::NT) void doppler_batch_kernel(const DopJob *jobs,
                                                                          const int *item_job,
                                                                          int64_t items) {
  constexpr int TR = DopBatchCfg<ND>::TR;
  __shared__ float2 lds[ND * TR];
  __shared__ float win[ND];
  auto item = [&](int64_t it) {
    const int ji = item_job ? __builtin_amdgcn_readfirstlane(item_job[it]) : 0;
    const DopJob jb = uniform_copy(&jobs[ji]);
    const int64_t local = it - jb.item0;
    doppler_tile<ND, TR, DopBatchCfg<ND>::NT>(jb.a, (int)(local / jb.tiles), (int)(local % jb.tiles) * TR,
                                             ND, lds, win);
  };
  if constexpr (!PERSIST) {
    item(blockIdx.x);
    return;
  }
  for (int64_t it = blockIdx.x; it < items; it += gridDim.x) {
    if (it != blockIdx.x) __syncthreads();  // previous item's LDS use is done
    item(it);
  }
}

mxr_status launch_doppler_batch(int nd, const DopJob *jobs, const int *item_job, int64_t items, int grid,
                                hipStream_t s) {
  if (items == 0) return MXR_SUCCESS;
  MXR_REQUIRE(grid >= 1 && grid <= items && items <= 0x7fffffff, MXR_INVALID_SIZE,
              "ragged batch: bad grid %d for %lld items", grid, (long long)items);
  const bool persist = grid < items;
  ProfScope ps("doppler_batch_kernel", s);
  switch (nd) {
#define MXR_DB(N)                                                                          \
  case N:                                                                                  \
    if (persist) doppler_batch_kernel<N, true><<<dim3(grid), dim3(DopBatchCfg<N>::NT), 0, s>>>(jobs, item_job, items); \
    else doppler_batch_kernel<N, false><<<dim3(grid), dim3(DopBatchCfg<N>::NT), 0, s>>>(jobs, item_job, items); \
    break;
    MXR_DB(2)
    MXR_DB(4)
    MXR_DB(8)
    MXR_DB(16)
    MXR_DB(32)
    MXR_DB(64)
    MXR_DB(128)
    MXR_DB(256)
    MXR_DB(512)
    MXR_DB(1024)
#undef MXR_DB
    default:
      set_error(MXR_NOT_SUPPORTED, "ragged batch: Doppler size %d unsupported", nd);
      return MXR_NOT_SUPPORTED;
  }
  MXR_HIP_TRY(hipGetLastError());
  return MXR_SUCCESS;
}

int doppler_batch_tile(int nd) {
  switch (nd) {
    case 2: return DopBatchCfg<2>::TR;
    case 4: return DopBatchCfg<4>::TR;
    case 8: return DopBatchCfg<8>::TR;
    case 16: return DopBatchCfg<16>::TR;
    case 32: return DopBatchCfg<32>::TR;
    case 64: return DopBatchCfg<64>::TR;
    case 128: return DopBatchCfg<128>::TR;
    case 256: return DopBatchCfg<256>::TR;
    case 512: return DopBatchCfg<512>::TR;
    case 1024: return DopBatchCfg<1024>::TR;
  }
  return 0;
}

template <int ND>
static mxr_status launch_doppler(const DopArgs &a, hipStream_t s) {
  constexpr int TR = DopCfg<ND>::TR;
  const int rows_buf = a.P > ND ? a.P : ND;
  const size_t lds = (size_t)rows_buf * TR * sizeof(float2);
  MXR_REQUIRE(lds + ND * sizeof(float) <= (size_t)max_lds_per_block(), MXR_NOT_SUPPORTED,
              "doppler: %d pulses x %d range bins do not fit in LDS", rows_buf, TR);
  if (lds + ND * sizeof(float) > 65536)
    MXR_HIP_TRY(hipFuncSetAttribute((const void *)doppler_kernel<ND>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  dim3 grid((a.R + TR - 1) / TR, a.C);
  {
    ProfScope ps("doppler_kernel", s);
    doppler_kernel<ND><<<grid, dim3(DopCfg<ND>::NT), lds, s>>>(a, rows_buf);
  }
  MXR_HIP_TRY(hipGetLastError());
  return MXR_SUCCESS;
}

static mxr_status doppler_run(const mxr_tensor *pow_out, const mxr_tensor *rd_out,
                              const mxr_tensor *y, int tpc_mode, int64_t win_rows,
                              const mxr_tensor *taps, int64_t nd, hipStream_t stream) {
  mxr_status st;
  if ((st = check_tensor(y, "doppler input", MXR_C64, 3, 3))) return st;
  MXR_REQUIRE(pow_out || rd_out, MXR_INVALID_PARAMETER, "doppler: no output requested");
  if (rd_out && (st = check_tensor(rd_out, "doppler rd output", MXR_C64, 3, 3))) return st;
  if (pow_out && (st = check_tensor(pow_out, "doppler power output", MXR_F32, 3, 3))) return st;
  MXR_REQUIRE(tpc_mode == MXR_CONV_SAME || tpc_mode == MXR_CONV_VALID || tpc_mode == MXR_TPC_NONE,
              MXR_INVALID_PARAMETER, "doppler: canceller mode must be SAME, VALID or NONE");
  const int64_t C = y->shape[0], P = y->shape[1], R = y->shape[2];
  MXR_REQUIRE(P >= 3 || tpc_mode == MXR_TPC_NONE, MXR_INVALID_SIZE,
              "doppler: need at least 3 pulses, got %lld", (long long)P);
  const int64_t rows_t = tpc_mode == MXR_CONV_VALID ? P - 2 : P;
  MXR_REQUIRE(nd >= rows_t, MXR_INVALID_SIZE,
              "doppler: %lld-point FFT is shorter than the %lld canceller rows", (long long)nd,
              (long long)rows_t);
  MXR_REQUIRE(win_rows >= 0 && win_rows <= rows_t, MXR_INVALID_SIZE, "doppler: bad window length");
  for (const mxr_tensor *o : {pow_out, rd_out}) {
    if (!o) continue;
    MXR_REQUIRE(o->shape[0] == C && o->shape[1] == nd && o->shape[2] == R, MXR_INVALID_SIZE,
                "doppler: output must be [%lld, %lld, %lld]", (long long)C, (long long)nd,
                (long long)R);
    const bool exact = o == rd_out && o->data == y->data && o->shape[1] == P &&
                       o->stride[0] == y->stride[0] && o->stride[1] == y->stride[1] &&
                       o->stride[2] == y->stride[2];
    MXR_REQUIRE(exact || !views_overlap(*o, *y), MXR_ALIAS_ERROR,
                "doppler: output partially aliases the input");
  }
  if (pow_out && rd_out)
    MXR_REQUIRE(!views_overlap(*pow_out, *rd_out), MXR_ALIAS_ERROR, "doppler: outputs overlap");
  if (C == 0 || R == 0) return MXR_SUCCESS;
  MXR_REQUIRE(C <= 65535 && R < (1ll << 30) && P < (1 << 20), MXR_INVALID_SIZE,
              "doppler: shape too large");
  DopArgs a{};
  a.y = (const float2 *)y->data;
  a.ys_c = y->stride[0];
  a.ys_p = y->stride[1];
  a.ys_r = y->stride[2];
  if (rd_out) {
    a.rd = (float2 *)rd_out->data;
    a.rs_c = rd_out->stride[0];
    a.rs_k = rd_out->stride[1];
    a.rs_r = rd_out->stride[2];
  }
  if (pow_out) {
    a.pw = (float *)pow_out->data;
    a.ps_c = pow_out->stride[0];
    a.ps_k = pow_out->stride[1];
    a.ps_r = pow_out->stride[2];
  }
  a.C = (int)C;
  a.P = (int)P;
  a.R = (int)R;
  a.rows_t = (int)rows_t;
  a.p_off = tpc_mode == MXR_CONV_VALID ? 1 : 0;
  a.no_tpc = tpc_mode == MXR_TPC_NONE;
  a.win_rows = (int)win_rows;
  if (taps) {
    if ((st = check_tensor(taps, "doppler taps", MXR_F32, 1, 1))) return st;
    MXR_REQUIRE(taps->shape[0] == rows_t, MXR_INVALID_SIZE,
                "doppler: %lld taps for %lld rows", (long long)taps->shape[0], (long long)rows_t);
    a.taps = (const float *)taps->data;
    a.taps_s = taps->stride[0];
  }
  MXR_REQUIRE(nd >= 2 && nd <= 1024 && (nd & (nd - 1)) == 0, MXR_NOT_SUPPORTED,
              "doppler: FFT size %lld must be a power of two in [2, 1024]", (long long)nd);
  a.tw = twiddles_device((int)nd, false);
  MXR_REQUIRE(a.tw != nullptr, MXR_OUT_OF_MEMORY, "doppler: twiddle table allocation failed");
  switch (nd) {
    case 2: return launch_doppler<2>(a, stream);
    case 4: return launch_doppler<4>(a, stream);
    case 8: return launch_doppler<8>(a, stream);
    case 16: return launch_doppler<16>(a, stream);
    case 32: return launch_doppler<32>(a, stream);
    case 64: return launch_doppler<64>(a, stream);
    case 128: return launch_doppler<128>(a, stream);
    case 256: return launch_doppler<256>(a, stream);
    case 512: return launch_doppler<512>(a, stream);
    case 1024: return launch_doppler<1024>(a, stream);
  }
  return MXR_NOT_SUPPORTED;
}

mxr_status doppler_impl(const mxr_tensor *pow_out, const mxr_tensor *rd_out, const mxr_tensor *y,
                        int tpc_mode, int64_t win_rows, int64_t nd, hipStream_t stream) {
  return doppler_run(pow_out, rd_out, y, tpc_mode, win_rows, nullptr, nd, stream);
}

// gh[p] = sum over parts of partial[part][p]: one workgroup per tap, a
// fixed-order tree in double (deterministic).
__global__ __launch_bounds__(256) void taps_reduce_kernel(const float *partial, int nparts, int P,
                                                          float *gh, int64_t ghs) {
  __shared__ double acc[256];
  const int p = blockIdx.x;
  double s = 0.0;
  for (int part = threadIdx.x; part < nparts; part += 256) s += (double)partial[(int64_t)part * P + p];
  acc[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) acc[threadIdx.x] += acc[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) gh[(int64_t)p * ghs] = (float)acc[0];
}

// ------------------------------------------------ pulse-axis adjoints
// Backward of mxr_doppler's power output with respect to its input y:
//   t_i = h_i (y[i+o-1] - 2 y[i+o] + y[i+o+1])   i < rows_t   (o = p_off)
//   D = FFT_nd(t),  pow = |D|^2
//   gD = 2 gpow D;  gu = F^H gD (unnormalised inverse DFT);  s_i = h_i gu_i
//   gy[p] = s[p-o+1] - 2 s[p-o] + s[p-o-1]   (s = 0 outside [0, rows_t))
// (no canceller: t_i = h_i y[i], gy[p] = s_p).  The canceller coefficients
// are real, so its adjoint is the same symmetric stencil.  With free taps h
// (mxr_doppler_taps, no canceller) the same kernel also gives the tap
// gradient gh[p] = sum_{c,r} Re(conj(gu[c,p,r]) y[c,p,r]): each tile reduces
// its range bins into one partial row, and taps_reduce_kernel sums the rows
// in a fixed order.  D is recomputed from y inside the tile (checkpointing:
// no nd x R complex spectrum is kept between the passes), so a tile reads y
// and gpow once and writes gy once.  For free taps (a design loss on a few
// cells leaves most of gpow zero) a tile whose gpow is all zero writes zeros
// and skips the transforms; the test costs 13 % on a dense gpow, so the
// plain adjoint (a loss over the whole map) does not make it.
struct DopGradArgs {
  const float2 *y;
  int64_t ys_c, ys_p, ys_r;
  const float *gp;
  int64_t gs_c, gs_k, gs_r;
  float2 *gy;  // may be null (tap gradient only)
  int64_t gys_c, gys_p, gys_r;
  const float *taps;  // null: Hamming(win_rows) then 1
  int64_t taps_s;
  float *partial;  // [C * tiles][P] tap-gradient rows, or null
  int skip_zero;   // test each tile's gpow for all zeros first
  int C, P, R, tiles;
  int rows_t, p_off, no_tpc, win_rows;
  const float2 *tw;
};

// Adjoint tiles: as the forward's, except at 512 Doppler bins, where two
// transforms per tile favour occupancy: 8 range bins (32 KiB) measured 17 %
// faster than 16 and 2x faster than 4.
template <int ND>
struct DopAdjCfg {
  static constexpr int TR = ND == 512 ? 8 : DopCfg<ND>::TR;
};

template <int ND>
__global__ __launch_bounds__(DopCfg<ND>::NT) void doppler_adj_kernel(DopGradArgs a) {
  constexpr int NT = DopCfg<ND>::NT;
  constexpr int TR = DopAdjCfg<ND>::TR;
  static_assert(TR <= 64 && 64 % TR == 0, "range tile must divide a wave");
  using LY = Layout<ND, TR, true>;
  using SQ = Seq<ND, false>;
  constexpr int K = SQ::n;
  extern __shared__ float2 lds[];  // max(P, ND) * TR: y tile, then D, then gu
  __shared__ float win[ND];
  const int c = blockIdx.y;
  const int r0 = blockIdx.x * TR;
  float *part = a.partial ? a.partial + ((int64_t)c * a.tiles + blockIdx.x) * a.P : nullptr;
  float2 *gyc = a.gy ? a.gy + (int64_t)c * a.gys_c : nullptr;
  const float *gpc = a.gp + (int64_t)c * a.gs_c;
  if (a.skip_zero) {
    int nz = 0;
#pragma unroll 8
    for (int idx = threadIdx.x; idx < ND * TR; idx += NT) {
      const int k = idx / TR, r = r0 + idx % TR;
      // unconditional loads (clamped column) so they issue back to back
      const float g = gpc[(int64_t)k * a.gs_k + (int64_t)min(r, a.R - 1) * a.gs_r];
      nz |= (r < a.R) & (g != 0.f);
    }
    if (!__syncthreads_or(nz)) {
      if (gyc)
        for (int idx = threadIdx.x; idx < a.P * TR; idx += NT) {
          const int p = idx / TR, r = r0 + idx % TR;
          if (r < a.R) gyc[(int64_t)p * a.gys_p + (int64_t)r * a.gys_r] = make_float2(0.f, 0.f);
        }
      if (part)
        for (int p = threadIdx.x; p < a.P; p += NT) part[p] = 0.f;
      return;
    }
  }
  for (int i = threadIdx.x; i < ND; i += NT) {
    float h = 0.f;
    if (i < a.rows_t)
      h = a.taps ? a.taps[(int64_t)i * a.taps_s] : (i < a.win_rows ? hamming_f32(i, a.win_rows) : 1.0f);
    win[i] = h;
  }
  load_pulse_tile<ND, TR>(a.y + (int64_t)c * a.ys_c, a.ys_p, a.ys_r, a.P, a.R, r0, lds);
  wg_barrier();

  // forward: first pass fed by the windowed canceller, the rest in LDS; the
  // last forward pass stores gD = 2 gpow D.  The gpow values of the last
  // pass's outputs are fetched before its LDS reads (column clamped, zeroed
  // past R by a mask, so the loads issue back to back).
  constexpr int RL = SQ::template R<K - 1>;
  constexpr int LSL = SQ::template LS<K - 1>;
  using PL = Pass<LY, RL, ND, NT, TR>;
  float gpv[PL::BPT][RL];
  auto fetch_gp = [&]() {
    PL::template outputs<LSL>([&](int u, int q, int t, int k) {
      const int r = r0 + t;
      const float g = gpc[(int64_t)k * a.gs_k + (int64_t)min(r, a.R - 1) * a.gs_r];
      gpv[u][q] = __int_as_float(__float_as_int(g) & (r < a.R ? -1 : 0));
    });
  };
  auto store_gd = [&](PL &pl) {
    PL::template outputs<LSL>([&](int u, int q, int t, int k) {
      lds[LY::addr(t, k)] = cscale(pl.v[u][q], 2.0f * gpv[u][q]);
    });
  };
  constexpr int R0 = SQ::template R<0>;
  Pass<LY, R0, ND, NT, TR> p0;
  if constexpr (K == 1) fetch_gp();
  p0.load([&](int t, int i) -> float2 {
    if (i >= a.rows_t) return make_float2(0.f, 0.f);
    const int p = i + a.p_off;
    const float2 m = lds[p * TR + t];
    if (a.no_tpc) return cscale(m, win[i]);
    const float2 s = p >= 1 ? lds[(p - 1) * TR + t] : make_float2(0.f, 0.f);
    const float2 u = p + 1 < a.P ? lds[(p + 1) * TR + t] : make_float2(0.f, 0.f);
    return cscale(make_float2(s.x + u.x - 2.0f * m.x, s.y + u.y - 2.0f * m.y), win[i]);
  });
  wg_barrier();  // every canceller read is done before the tile is overwritten
  p0.template compute<1, false>(nullptr);
  if constexpr (K == 1) {
    store_gd(p0);
  } else {
    p0.template store<1>([&](int t, int i, float2 v) { lds[LY::addr(t, i)] = v; });
    wg_barrier();
    lds_passes<LY, SQ, ND, NT, TR, false, 1, K - 1>(lds, a.tw);
    PL pl;
    fetch_gp();
    pl.load([&](int t, int i) { return lds[LY::addr(t, i)]; });
    wg_barrier();
    pl.template compute<LSL, false>(a.tw + SQ::template TWOFF<K - 1>);
    store_gd(pl);
  }
  wg_barrier();
  // gu = F^H gD: the same passes with conjugate twiddles, unscaled
  lds_passes<LY, SQ, ND, NT, TR, true, 0, K>(lds, a.tw);

  // window, canceller transpose, coalesced store; tap-gradient partials
  auto s_at = [&](int t, int i) -> float2 {
    return (i >= 0 && i < a.rows_t) ? cscale(lds[LY::addr(t, i)], win[i]) : make_float2(0.f, 0.f);
  };
  const float2 *yc = a.y + (int64_t)c * a.ys_c;
  for (int base = 0; base < a.P * TR; base += NT) {  // uniform trip count (shuffles)
    const int idx = base + threadIdx.x;
    const int p = idx / TR, t = idx % TR, r = r0 + t;
    const bool ok = p < a.P && r < a.R;
    float contrib = 0.f;
    if (ok && gyc) {
      const int i = p - a.p_off;
      float2 g;
      if (a.no_tpc) {
        g = s_at(t, i);
      } else {
        const float2 lo = s_at(t, i - 1), mid = s_at(t, i), hi = s_at(t, i + 1);
        g = make_float2(lo.x + hi.x - 2.0f * mid.x, lo.y + hi.y - 2.0f * mid.y);
      }
      gyc[(int64_t)p * a.gys_p + (int64_t)r * a.gys_r] = g;
    }
    if (part) {  // no canceller here: row p of gu pairs with pulse p
      if (ok) {
        const float2 gu = lds[LY::addr(t, p)];
        const float2 yv = yc[(int64_t)p * a.ys_p + (int64_t)r * a.ys_r];
        contrib = gu.x * yv.x + gu.y * yv.y;
      }
#pragma unroll
      for (int o = TR / 2; o > 0; o >>= 1) contrib += __shfl_xor(contrib, o);
      if (t == 0 && p < a.P) part[p] = contrib;
    }
  }
}

template <int ND>
static mxr_status launch_doppler_adj(DopGradArgs a, const mxr_tensor *gtaps, hipStream_t s) {
  constexpr int TR = DopAdjCfg<ND>::TR;
  const int rows_buf = a.P > ND ? a.P : ND;
  const size_t lds = (size_t)rows_buf * TR * sizeof(float2);
  MXR_REQUIRE(lds + ND * sizeof(float) <= (size_t)max_lds_per_block(), MXR_NOT_SUPPORTED,
              "doppler adjoint: %d pulses x %d range bins do not fit in LDS", rows_buf, TR);
  if (lds + ND * sizeof(float) > 65536)
    MXR_HIP_TRY(hipFuncSetAttribute((const void *)doppler_adj_kernel<ND>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  a.tiles = (a.R + TR - 1) / TR;
  a.partial = nullptr;
  if (gtaps) {
    a.partial = (float *)stream_workspace("doppler_taps_partial",
                                          sizeof(float) * (size_t)a.C * a.tiles * a.P, s);
    if (a.partial == nullptr) return workspace_failure("doppler adjoint");
  }
  {
    ProfScope ps("doppler_adj_kernel", s);
    doppler_adj_kernel<ND><<<dim3(a.tiles, a.C), dim3(DopCfg<ND>::NT), lds, s>>>(a);
  }
  MXR_HIP_TRY(hipGetLastError());
  if (gtaps) {
    ProfScope ps("taps_reduce_kernel", s);
    taps_reduce_kernel<<<dim3(a.P), dim3(256), 0, s>>>(a.partial, a.C * a.tiles, a.P,
                                                       (float *)gtaps->data, gtaps->stride[0]);
    MXR_HIP_TRY(hipGetLastError());
  }
  return MXR_SUCCESS;
}

static mxr_status dispatch_doppler_adj(int64_t nd, const DopGradArgs &a, const mxr_tensor *gtaps,
                                       hipStream_t s) {
  switch (nd) {
    case 2: return launch_doppler_adj<2>(a, gtaps, s);
    case 4: return launch_doppler_adj<4>(a, gtaps, s);
    case 8: return launch_doppler_adj<8>(a, gtaps, s);
    case 16: return launch_doppler_adj<16>(a, gtaps, s);
    case 32: return launch_doppler_adj<32>(a, gtaps, s);
    case 64: return launch_doppler_adj<64>(a, gtaps, s);
    case 128: return launch_doppler_adj<128>(a, gtaps, s);
    case 256: return launch_doppler_adj<256>(a, gtaps, s);
    case 512: return launch_doppler_adj<512>(a, gtaps, s);
    case 1024: return launch_doppler_adj<1024>(a, gtaps, s);
  }
  set_error(MXR_NOT_SUPPORTED, "doppler adjoint: FFT size %lld unsupported", (long long)nd);
  return MXR_NOT_SUPPORTED;
}

static mxr_status doppler_adjoint_impl(const mxr_tensor *gy, const mxr_tensor *gpow,
                                       const mxr_tensor *y, int tpc_mode, int64_t win_rows,
                                       int64_t nd, hipStream_t stream) {
  const char *fn = "doppler_adjoint";
  mxr_status st;
  if ((st = check_tensor(y, "doppler_adjoint y", MXR_C64, 3, 3))) return st;
  if ((st = check_tensor(gpow, "doppler_adjoint gpow", MXR_F32, 3, 3))) return st;
  if ((st = check_tensor(gy, "doppler_adjoint gy", MXR_C64, 3, 3))) return st;
  MXR_REQUIRE(tpc_mode == MXR_CONV_SAME || tpc_mode == MXR_CONV_VALID || tpc_mode == MXR_TPC_NONE,
              MXR_INVALID_PARAMETER, "%s: canceller mode must be SAME, VALID or NONE", fn);
  const int64_t C = y->shape[0], P = y->shape[1], R = y->shape[2];
  MXR_REQUIRE(P >= 3 || tpc_mode == MXR_TPC_NONE, MXR_INVALID_SIZE,
              "%s: need at least 3 pulses, got %lld", fn, (long long)P);
  const int64_t rows_t = tpc_mode == MXR_CONV_VALID ? P - 2 : P;
  MXR_REQUIRE(nd >= rows_t, MXR_INVALID_SIZE,
              "%s: %lld-point FFT is shorter than the %lld canceller rows", fn, (long long)nd,
              (long long)rows_t);
  MXR_REQUIRE(win_rows >= 0 && win_rows <= rows_t, MXR_INVALID_SIZE, "%s: bad window length", fn);
  MXR_REQUIRE(gpow->shape[0] == C && gpow->shape[1] == nd && gpow->shape[2] == R, MXR_INVALID_SIZE,
              "%s: gpow must be [%lld, %lld, %lld]", fn, (long long)C, (long long)nd, (long long)R);
  MXR_REQUIRE(gy->shape[0] == C && gy->shape[1] == P && gy->shape[2] == R, MXR_INVALID_SIZE,
              "%s: gy must have y's shape [%lld, %lld, %lld]", fn, (long long)C, (long long)P,
              (long long)R);
  MXR_REQUIRE(!views_overlap(*gy, *y) && !views_overlap(*gy, *gpow), MXR_ALIAS_ERROR,
              "%s: gy overlaps an input", fn);
  MXR_REQUIRE(nd >= 2 && nd <= 1024 && (nd & (nd - 1)) == 0, MXR_NOT_SUPPORTED,
              "%s: FFT size %lld must be a power of two in [2, 1024]", fn, (long long)nd);
  if (C == 0 || R == 0 || P == 0) return MXR_SUCCESS;
  MXR_REQUIRE(C <= 65535 && R < (1ll << 30) && P < (1 << 20), MXR_INVALID_SIZE,
              "%s: shape too large", fn);
  DopGradArgs a{};
  a.y = (const float2 *)y->data;
  a.ys_c = y->stride[0];
  a.ys_p = y->stride[1];
  a.ys_r = y->stride[2];
  a.gp = (const float *)gpow->data;
  a.gs_c = gpow->stride[0];
  a.gs_k = gpow->stride[1];
  a.gs_r = gpow->stride[2];
  a.gy = (float2 *)gy->data;
  a.gys_c = gy->stride[0];
  a.gys_p = gy->stride[1];
  a.gys_r = gy->stride[2];
  a.C = (int)C;
  a.P = (int)P;
  a.R = (int)R;
  a.rows_t = (int)rows_t;
  a.p_off = tpc_mode == MXR_CONV_VALID ? 1 : 0;
  a.no_tpc = tpc_mode == MXR_TPC_NONE;
  a.win_rows = (int)win_rows;
  a.tw = twiddles_device((int)nd, false);
  MXR_REQUIRE(a.tw != nullptr, MXR_OUT_OF_MEMORY, "%s: twiddle table allocation failed", fn);
  return dispatch_doppler_adj(nd, a, nullptr, stream);
}

static mxr_status doppler_taps_adjoint_impl(const mxr_tensor *gtaps, const mxr_tensor *gx,
                                            const mxr_tensor *gpow, const mxr_tensor *x,
                                            const mxr_tensor *taps, int64_t nd, hipStream_t s) {
  const char *fn = "doppler_taps_adjoint";
  mxr_status st;
  if ((st = check_tensor(x, "x", MXR_C64, 3, 3))) return st;
  if ((st = check_tensor(taps, "taps", MXR_F32, 1, 1))) return st;
  if ((st = check_tensor(gpow, "gpow", MXR_F32, 3, 3))) return st;
  MXR_REQUIRE(gtaps || gx, MXR_INVALID_PARAMETER, "%s: no gradient requested", fn);
  const int64_t C = x->shape[0], P = x->shape[1], R = x->shape[2];
  MXR_REQUIRE(taps->shape[0] == P, MXR_INVALID_SIZE, "%s: %lld taps for %lld pulses", fn,
              (long long)taps->shape[0], (long long)P);
  MXR_REQUIRE(nd >= 2 && nd <= 1024 && (nd & (nd - 1)) == 0 && nd >= P, MXR_NOT_SUPPORTED,
              "%s: FFT size %lld must be a power of two in [max(2, P), 1024]", fn, (long long)nd);
  MXR_REQUIRE(gpow->shape[0] == C && gpow->shape[1] == nd && gpow->shape[2] == R,
              MXR_INVALID_SIZE, "%s: gpow must be [%lld, %lld, %lld]", fn, (long long)C,
              (long long)nd, (long long)R);
  if (gtaps) {
    if ((st = check_tensor(gtaps, "gtaps", MXR_F32, 1, 1))) return st;
    MXR_REQUIRE(gtaps->shape[0] == P, MXR_INVALID_SIZE, "%s: gtaps must have %lld entries", fn,
                (long long)P);
  }
  if (gx) {
    if ((st = check_tensor(gx, "gx", MXR_C64, 3, 3))) return st;
    MXR_REQUIRE(gx->shape[0] == C && gx->shape[1] == P && gx->shape[2] == R, MXR_INVALID_SIZE,
                "%s: gx must have x's shape", fn);
  }
  for (const mxr_tensor *o : {gtaps, gx}) {
    if (!o) continue;
    for (const mxr_tensor *i : {x, taps, gpow})
      MXR_REQUIRE(!views_overlap(*o, *i), MXR_ALIAS_ERROR, "%s: an output overlaps an input", fn);
  }
  if (gtaps && gx) MXR_REQUIRE(!views_overlap(*gtaps, *gx), MXR_ALIAS_ERROR, "%s: outputs overlap", fn);
  if (C == 0 || R == 0 || P == 0) return MXR_SUCCESS;
  MXR_REQUIRE(C <= 65535 && R < (1ll << 30), MXR_INVALID_SIZE, "%s: shape too large", fn);
  DopGradArgs a{};
  a.y = (const float2 *)x->data;
  a.ys_c = x->stride[0];
  a.ys_p = x->stride[1];
  a.ys_r = x->stride[2];
  a.gp = (const float *)gpow->data;
  a.gs_c = gpow->stride[0];
  a.gs_k = gpow->stride[1];
  a.gs_r = gpow->stride[2];
  if (gx) {
    a.gy = (float2 *)gx->data;
    a.gys_c = gx->stride[0];
    a.gys_p = gx->stride[1];
    a.gys_r = gx->stride[2];
  }
  a.taps = (const float *)taps->data;
  a.taps_s = taps->stride[0];
  a.C = (int)C;
  a.P = (int)P;
  a.R = (int)R;
  a.rows_t = (int)P;
  a.p_off = 0;
  a.no_tpc = 1;
  a.win_rows = 0;
  a.skip_zero = 1;
  a.tw = twiddles_device((int)nd, false);
  MXR_REQUIRE(a.tw != nullptr, MXR_OUT_OF_MEMORY, "%s: twiddle table allocation failed", fn);
  return dispatch_doppler_adj(nd, a, gtaps, s);
}

}  // namespace mxr

extern "C" mxr_status mxr_doppler_adjoint(const mxr_tensor *gy, const mxr_tensor *gpow,
                                          const mxr_tensor *y, int tpc_mode, int64_t win_rows,
                                          int64_t nd, hipStream_t stream) {
  return mxr::guarded("mxr_doppler_adjoint", [&] {
    return mxr::doppler_adjoint_impl(gy, gpow, y, tpc_mode, win_rows, nd, stream);
  });
}

extern "C" mxr_status mxr_doppler(const mxr_tensor *pow_out, const mxr_tensor *rd_out,
                                  const mxr_tensor *y, int tpc_mode, int64_t win_rows, int64_t nd,
                                  hipStream_t stream) {
  return mxr::guarded("mxr_doppler", [&] {
    return mxr::doppler_impl(pow_out, rd_out, y, tpc_mode, win_rows, nd, stream);
  });
}

extern "C" mxr_status mxr_doppler_taps(const mxr_tensor *pow_out, const mxr_tensor *rd_out,
                                       const mxr_tensor *x, const mxr_tensor *taps, int64_t nd,
                                       hipStream_t stream) {
  return mxr::guarded("mxr_doppler_taps", [&] {
    MXR_REQUIRE(taps != nullptr, MXR_INVALID_PARAMETER, "doppler_taps: null taps");
    return mxr::doppler_run(pow_out, rd_out, x, MXR_TPC_NONE, 0, taps, nd, stream);
  });
}

extern "C" mxr_status mxr_doppler_taps_adjoint(const mxr_tensor *gtaps, const mxr_tensor *gx,
                                               const mxr_tensor *gpow, const mxr_tensor *x,
                                               const mxr_tensor *taps, int64_t nd,
                                               hipStream_t stream) {
  return mxr::guarded("mxr_doppler_taps_adjoint", [&] {
    return mxr::doppler_taps_adjoint_impl(gtaps, gx, gpow, x, taps, nd, stream);
  });
}
