// Ragged batch: many independent pulse-compression + Doppler jobs of uneven
// shapes run as a handful of persistent launches (one per block-size class
// and stage), which a caller captures once into a hipGraph and replays.
//
// Per job (the stage semantics of examples/simple_radar_pipeline.h:273-301,
// 344-357, with the canceller left out):
//   y[c, p, r]  = pulse compression of x[c, p, :S] with window-normalised w
//                 (n = next_pow2(S), lags r < R = S - L + 1), as
//                 mxr_pulse_compress computes it
//   pw[c, k, r] = |sum_p hamming_P(p) y[c, p, r] e^{-2 pi i k p / nd}|^2,
//                 nd = next_pow2(P), as mxr_doppler(..., MXR_TPC_NONE, P, nd)
// The plan owns the compressed cubes (scratch) and the waveform spectra.
//
// Block size per job: the overlap-save planner's cost model, or (with
// MXR_RAGGED_AUTOTUNE) the fastest of the candidate block sizes measured on
// the device for the job's size class (waveform length rounded up to a power
// of two, transform size), cached per device for the process.
#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "fft_engine.hpp"
#include "mxr_cache.hpp"
#include "mxr_trace.hpp"
#include "ragged.hpp"

namespace mxr {

namespace {

int64_t next_pow2(int64_t v) {
  int64_t p = 1;
  while (p < v) p <<= 1;
  return p;
}

struct PcClass {
  int M = 0;
  int nd = 0;  // Doppler size of its jobs (the chain it belongs to)
  int njobs = 0;
  int64_t items = 0;
  int grid = 0;
  PcJob *d_jobs = nullptr;
  WfArgs *d_wf = nullptr;
  const int *d_item_job = nullptr;  // job (index within the class) of each item
};

struct DopClass {
  int nd = 0;
  int njobs = 0;
  int64_t items = 0;
  int grid = 0;
  DopJob *d_jobs = nullptr;
  const int *d_item_job = nullptr;
};

// The jobs of one Doppler size form a chain: the spectrum and pulse-
// compression launches of their block-size classes, then their Doppler
// launch.  Chains are independent, so they run concurrently on helper
// streams: the caller's stream records a fork event, each helper stream
// waits on it, runs its chains in turn and records a join event, and the
// caller's stream waits on every join (the same fork/join, captured into a
// graph, becomes parallel branches).  One chain's Doppler pass thus overlaps
// another chain's pulse compression.  All at default priority.
constexpr int RAGGED_BRANCHES = 4;

struct RaggedPlanObj {
  int dev = 0;
  std::vector<mxr_ragged_job> jobs;
  std::vector<int> M;
  float2 *y = nullptr;     // compressed cubes of every job
  float2 *spec = nullptr;  // waveform spectra of every job
  void *tables = nullptr;  // device copies of the job tables
  std::vector<PcClass> pc;
  std::vector<DopClass> dop;
  std::vector<int> dop_order;   // chains by decreasing work
  std::vector<int> dop_branch;  // helper stream of each chain
  hipStream_t branch[RAGGED_BRANCHES] = {};
  hipEvent_t fork = {}, join[RAGGED_BRANCHES] = {};
};

std::mutex g_tune_mu;
std::map<std::tuple<int, int64_t, int64_t>, int> g_tuned;  // (device, L class, n) -> M

// Times the pulse-compression batch kernel for candidate block sizes on a
// synthetic class representative (256 rows of n samples, waveform length
// Lc) and returns the size with the least time per output lag.
mxr_status autotune_block(int64_t Lc, int64_t n, hipStream_t s, int *bestM) {
  const auto key = std::make_tuple(device_id(), Lc, n);
  {
    std::lock_guard<std::mutex> g(g_tune_mu);
    auto it = g_tuned.find(key);
    if (it != g_tuned.end()) {
      *bestM = it->second;
      return MXR_SUCCESS;
    }
  }
  const int rows = 256;
  const int64_t nout = n - Lc + 1;
  float2 *x = nullptr, *y = nullptr, *wc = nullptr;
  PcJob *dj = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  mxr_status st = MXR_SUCCESS;
  double best = 1e300;
  int bm = 0;
  auto fail = [&](hipError_t e) {
    set_error(MXR_HIP_ERROR, "ragged autotune: %s", hipGetErrorString(e));
    st = MXR_HIP_ERROR;
  };
  hipError_t e;
  if ((e = hipMalloc(&x, sizeof(float2) * rows * n)) != hipSuccess ||
      (e = hipMalloc(&y, sizeof(float2) * rows * n)) != hipSuccess ||
      (e = hipMalloc(&wc, sizeof(float2) * 16384)) != hipSuccess ||
      (e = hipMalloc(&dj, sizeof(PcJob))) != hipSuccess ||
      (e = hipMemsetAsync(x, 0, sizeof(float2) * rows * n, s)) != hipSuccess ||
      (e = hipMemsetAsync(wc, 0, sizeof(float2) * 16384, s)) != hipSuccess ||
      (e = hipEventCreate(&e0)) != hipSuccess || (e = hipEventCreate(&e1)) != hipSuccess) {
    fail(e);
  }
  for (int M = 1024; st == MXR_SUCCESS && M <= 16384 && M <= n; M *= 2) {
    int Mp;
    int64_t V, nblk;
    if (pc_plan_blocks(n, Lc, n, nout, M, &Mp, &V, &nblk) != MXR_SUCCESS) continue;
    PcJob j{};
    j.x = x;
    j.x_cs = 0;
    j.x_ps = n;
    j.y = y;
    j.y_cs = 0;
    j.y_ps = n;
    j.C = 1;
    j.P = rows;
    j.seff = (int)n;
    j.n = (int)n;
    j.nout = (int)nout;
    j.V = (int)V;
    j.nblk = (int)nblk;
    j.wc = wc;
    j.item0 = 0;
    if ((e = hipMemcpyAsync(dj, &j, sizeof(PcJob), hipMemcpyHostToDevice, s)) != hipSuccess) {
      fail(e);
      break;
    }
    const int64_t items = (int64_t)rows * nblk;
    const int grid = (int)items;
    if ((st = launch_pc_batch(M, dj, nullptr, items, grid, s))) break;  // warm-up
    if ((e = hipEventRecord(e0, s)) != hipSuccess) {
      fail(e);
      break;
    }
    for (int r = 0; r < 3 && st == MXR_SUCCESS; ++r) st = launch_pc_batch(M, dj, nullptr, items, grid, s);
    if (st) break;
    float ms = 0.f;
    if ((e = hipEventRecord(e1, s)) != hipSuccess || (e = hipEventSynchronize(e1)) != hipSuccess ||
        (e = hipEventElapsedTime(&ms, e0, e1)) != hipSuccess) {
      fail(e);
      break;
    }
    const double per_lag = ms / ((double)rows * nout);
    if (per_lag < best) {
      best = per_lag;
      bm = M;
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipStreamSynchronize(s);
  (void)hipFree(x);
  (void)hipFree(y);
  (void)hipFree(wc);
  (void)hipFree(dj);
  if (st) return st;
  MXR_REQUIRE(bm > 0, MXR_NOT_SUPPORTED, "ragged autotune: no block size fits L=%lld n=%lld",
              (long long)Lc, (long long)n);
  std::lock_guard<std::mutex> g(g_tune_mu);
  g_tuned[key] = bm;
  *bestM = bm;
  return MXR_SUCCESS;
}

void free_plan(RaggedPlanObj *p) {
  if (!p) return;
  int cur = 0;
  (void)hipGetDevice(&cur);
  (void)hipSetDevice(p->dev);
  (void)hipFree(p->y);
  (void)hipFree(p->spec);
  (void)hipFree(p->tables);
  for (int b = 0; b < RAGGED_BRANCHES; ++b) {
    if (p->branch[b]) (void)hipStreamDestroy(p->branch[b]);
    if (p->join[b]) (void)hipEventDestroy(p->join[b]);
  }
  if (p->fork) (void)hipEventDestroy(p->fork);
  (void)hipSetDevice(cur);
  delete p;
  plan_tables_release();
}

}  // namespace

static mxr_status ragged_create(mxr_ragged_plan *out, const mxr_ragged_job *jobs, int njobs,
                                int flags, hipStream_t s) {
  const char *fn = "ragged_plan_create";
  MXR_REQUIRE(out != nullptr, MXR_INVALID_PARAMETER, "%s: null plan pointer", fn);
  *out = nullptr;
  MXR_REQUIRE(njobs >= 1 && jobs != nullptr, MXR_INVALID_PARAMETER, "%s: no jobs", fn);
  MXR_REQUIRE((flags & ~(MXR_RAGGED_AUTOTUNE | MXR_RAGGED_PERSISTENT)) == 0, MXR_INVALID_PARAMETER,
              "%s: unknown flags", fn);
  const bool persist = (flags & MXR_RAGGED_PERSISTENT) != 0;
  for (int i = 0; i < njobs; ++i) {
    const mxr_ragged_job &j = jobs[i];
    MXR_REQUIRE(j.x && j.w && j.pw, MXR_INVALID_PARAMETER, "%s: job %d has a null pointer", fn, i);
    MXR_REQUIRE(j.C >= 1 && j.P >= 1 && j.S >= 1 && j.L >= 1 && j.L <= j.S, MXR_INVALID_SIZE,
                "%s: job %d needs C, P, S >= 1 and 1 <= L <= S", fn, i);
    MXR_REQUIRE(next_pow2(j.S) <= 16384 && next_pow2(j.P) <= 1024 && j.C <= 65535,
                MXR_NOT_SUPPORTED, "%s: job %d exceeds S <= 16384, P <= 1024", fn, i);
    MXR_REQUIRE(j.P >= 2, MXR_NOT_SUPPORTED,
                "%s: job %d has one pulse; the Doppler transform needs P >= 2", fn, i);
    MXR_REQUIRE(next_pow2(j.S) >= 1024, MXR_NOT_SUPPORTED,
                "%s: job %d has S = %lld; the batch's block sizes start at 1024, so S >= 513", fn, i,
                (long long)j.S);
    MXR_REQUIRE(j.x_ps >= j.S && j.x_cs >= j.x_ps * j.P && j.pw_ks >= j.S - j.L + 1 &&
                    j.pw_cs >= j.pw_ks * next_pow2(j.P),
                MXR_INVALID_PARAMETER, "%s: job %d strides overlap rows", fn, i);
  }
  auto *p = new RaggedPlanObj();
  plan_tables_retain();  // released by free_plan
  p->dev = device_id();
  p->jobs.assign(jobs, jobs + njobs);
  p->M.resize(njobs);
  mxr_status st = MXR_SUCCESS;
  // block size per job
  for (int i = 0; i < njobs && !st; ++i) {
    const mxr_ragged_job &j = jobs[i];
    const int64_t n = next_pow2(j.S), R = j.S - j.L + 1;
    int M = 0, Mreq = 0;
    int64_t V, nblk;
    if (flags & MXR_RAGGED_AUTOTUNE) st = autotune_block(next_pow2(j.L), n, s, &Mreq);
    if (!st && (pc_plan_blocks(j.S, j.L, n, R, Mreq, &M, &V, &nblk) != MXR_SUCCESS))
      st = pc_plan_blocks(j.S, j.L, n, R, 0, &M, &V, &nblk);  // tuned size does not fit this job
    // the batch kernels take blocks of 1024 .. 16384 (n >= 1024 here)
    if (!st && M < 1024) st = pc_plan_blocks(j.S, j.L, n, R, 1024, &M, &V, &nblk);
    if (!st && (M < 1024 || M > 16384)) {
      set_error(MXR_NOT_SUPPORTED, "%s: job %d: block size %d outside [1024, 16384]", fn, i, M);
      st = MXR_NOT_SUPPORTED;
    }
    p->M[i] = M;
  }
  // scratch: compressed cubes (rows padded to 64 lags) and spectra
  std::vector<int64_t> yoff(njobs), soff(njobs), yps(njobs);
  int64_t ytot = 0, stot = 0;
  for (int i = 0; i < njobs && !st; ++i) {
    const mxr_ragged_job &j = jobs[i];
    yps[i] = (j.S - j.L + 1 + 63) / 64 * 64;
    yoff[i] = ytot;
    ytot += j.C * j.P * yps[i];
    soff[i] = stot;
    stot += p->M[i];
  }
  hipError_t e = hipSuccess;
  if (!st && ((e = hipMalloc(&p->y, sizeof(float2) * ytot)) != hipSuccess ||
              (e = hipMalloc(&p->spec, sizeof(float2) * stot)) != hipSuccess)) {
    set_error(MXR_OUT_OF_MEMORY, "%s: scratch allocation failed: %s", fn, hipGetErrorString(e));
    st = MXR_OUT_OF_MEMORY;
  }
  // classes
  std::vector<PcJob> hpc;
  std::vector<WfArgs> hwf;
  std::vector<DopJob> hdop;
  std::vector<std::pair<size_t, size_t>> pc_ranges, dop_ranges;
  std::vector<int> Ms, nds;
  for (int i = 0; i < njobs; ++i) {
    if (std::find(Ms.begin(), Ms.end(), p->M[i]) == Ms.end()) Ms.push_back(p->M[i]);
    const int nd = (int)next_pow2(jobs[i].P);
    if (std::find(nds.begin(), nds.end(), nd) == nds.end()) nds.push_back(nd);
  }
  for (int nd : nds)
  for (int M : Ms) {
    if (st) break;
    bool any = false;
    for (int i = 0; i < njobs && !any; ++i) any = p->M[i] == M && (int)next_pow2(jobs[i].P) == nd;
    if (!any) continue;
    PcClass cl;
    cl.M = M;
    cl.nd = nd;
    const size_t first = hpc.size();
    if ((st = pc_tables_prepare(M, fn))) break;
    const float2 *tw = twiddles_device(M, false);
    for (int i = 0; i < njobs; ++i) {
      if (p->M[i] != M || (int)next_pow2(jobs[i].P) != nd) continue;
      const mxr_ragged_job &j = jobs[i];
      const int64_t n = next_pow2(j.S), R = j.S - j.L + 1;
      int Mp;
      int64_t V, nblk;
      if ((st = pc_plan_blocks(j.S, j.L, n, R, M, &Mp, &V, &nblk))) break;
      PcJob pj{};
      pj.x = (const float2 *)j.x;
      pj.x_cs = j.x_cs;
      pj.x_ps = j.x_ps;
      pj.y = p->y + yoff[i];
      pj.y_cs = j.P * yps[i];
      pj.y_ps = yps[i];
      pj.C = (int)j.C;
      pj.P = (int)j.P;
      pj.seff = (int)j.S;
      pj.n = (int)n;
      pj.nout = (int)R;
      pj.V = (int)V;
      pj.nblk = (int)nblk;
      pj.wc = p->spec + soff[i];
      pj.item0 = cl.items;
      cl.items += j.C * j.P * nblk;
      hpc.push_back(pj);
      WfArgs wa{};
      wa.w = (float2 *)j.w;
      wa.wstride = 1;
      wa.L = j.L;
      wa.window = 1;
      wa.normalize = 1;
      wa.writeback = 0;
      wa.conj = 1;
      wa.wc = p->spec + soff[i];
      wa.tw = tw;
      wa.out_scale = (float)(1.0 / (double)M);
      hwf.push_back(wa);
    }
    cl.njobs = (int)(hpc.size() - first);
    cl.grid = persist ? (int)std::min<int64_t>(cl.items, pc_batch_grid(M)) : (int)cl.items;
    pc_ranges.push_back({first, hpc.size()});
    p->pc.push_back(cl);
  }
  for (int nd : nds) {
    if (st) break;
    DopClass cl;
    cl.nd = nd;
    const size_t first = hdop.size();
    const float2 *tw = twiddles_device(nd, false);
    const int TR = doppler_batch_tile(nd);
    if (!tw || TR <= 0) {
      set_error(MXR_OUT_OF_MEMORY, "%s: twiddle table allocation failed", fn);
      st = MXR_OUT_OF_MEMORY;
      break;
    }
    for (int i = 0; i < njobs; ++i) {
      const mxr_ragged_job &j = jobs[i];
      if (next_pow2(j.P) != nd) continue;
      const int64_t R = j.S - j.L + 1;
      DopJob dj{};
      dj.a.y = p->y + yoff[i];
      dj.a.ys_c = j.P * yps[i];
      dj.a.ys_p = yps[i];
      dj.a.ys_r = 1;
      dj.a.pw = (float *)j.pw;
      dj.a.ps_c = j.pw_cs;
      dj.a.ps_k = j.pw_ks;
      dj.a.ps_r = 1;
      dj.a.C = (int)j.C;
      dj.a.P = (int)j.P;
      dj.a.R = (int)R;
      dj.a.rows_t = (int)j.P;
      dj.a.p_off = 0;
      dj.a.win_rows = (int)j.P;
      dj.a.no_tpc = 1;
      dj.a.tw = tw;
      dj.tiles = (int)((R + TR - 1) / TR);
      dj.item0 = cl.items;
      cl.items += j.C * dj.tiles;
      hdop.push_back(dj);
    }
    cl.njobs = (int)(hdop.size() - first);
    int cus = 256;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, p->dev);
    cl.grid = persist ? (int)std::min<int64_t>(cl.items, (int64_t)cus * 4) : (int)cl.items;
    dop_ranges.push_back({first, hdop.size()});
    p->dop.push_back(cl);
  }
  // chains to helper streams, largest first onto the least loaded stream.
  // A chain's cost is the time its jobs' kernels would take at the rates
  // they reach on their own: pulse compression bound by the larger of its
  // HBM bytes (x in, y out) at 5 TB/s and its FFT flops (forward + inverse,
  // 5 M log2 M each per block) at 35 TFLOP/s; the Doppler pass by its bytes
  // (y in, power out) at 4 TB/s.
  if (!st) {
    std::vector<double> cost(p->dop.size(), 0.0);
    for (int i = 0; i < njobs; ++i) {
      const mxr_ragged_job &j = jobs[i];
      const int nd = (int)next_pow2(j.P);
      const int64_t n = next_pow2(j.S), R = j.S - j.L + 1;
      int Mp = 0;
      int64_t V = 0, nblk = 0;
      if (pc_plan_blocks(j.S, j.L, n, R, p->M[i], &Mp, &V, &nblk) != MXR_SUCCESS) nblk = (R + p->M[i] - 1) / p->M[i];
      const double CP = (double)j.C * (double)j.P, M = (double)p->M[i];
      const double t_pc = std::max(8.0 * CP * (double)(j.S + R) / 5.0e12,
                                   10.0 * CP * (double)nblk * M * std::log2(M) / 35.0e12);
      const double t_dop = (8.0 * CP * (double)R + 4.0 * (double)j.C * nd * (double)R) / 4.0e12;
      for (size_t d = 0; d < p->dop.size(); ++d)
        if (p->dop[d].nd == nd) cost[d] += t_pc + t_dop;
    }
    p->dop_order.resize(p->dop.size());
    for (size_t d = 0; d < p->dop.size(); ++d) p->dop_order[d] = (int)d;
    std::stable_sort(p->dop_order.begin(), p->dop_order.end(),
                     [&](int a, int b) { return cost[a] > cost[b]; });
    p->dop_branch.assign(p->dop.size(), 0);
    double load[RAGGED_BRANCHES] = {};
    for (int d : p->dop_order) {
      int b = 0;
      for (int k = 1; k < RAGGED_BRANCHES; ++k)
        if (load[k] < load[b]) b = k;
      p->dop_branch[d] = b;
      load[b] += cost[d];
    }
  }
  // item -> job tables (job index within its class), classes back to back
  std::vector<int> hitem;
  std::vector<size_t> pc_item0, dop_item0;
  for (size_t c = 0; c < p->pc.size() && !st; ++c) {
    pc_item0.push_back(hitem.size());
    for (size_t k = pc_ranges[c].first; k < pc_ranges[c].second; ++k) {
      const PcJob &pj = hpc[k];
      hitem.insert(hitem.end(), (size_t)pj.C * pj.P * pj.nblk, (int)(k - pc_ranges[c].first));
    }
  }
  for (size_t c = 0; c < p->dop.size() && !st; ++c) {
    dop_item0.push_back(hitem.size());
    for (size_t k = dop_ranges[c].first; k < dop_ranges[c].second; ++k) {
      const DopJob &dj = hdop[k];
      hitem.insert(hitem.end(), (size_t)dj.a.C * dj.tiles, (int)(k - dop_ranges[c].first));
    }
  }
  // one device allocation for every table
  const size_t bpc = sizeof(PcJob) * hpc.size(), bwf = sizeof(WfArgs) * hwf.size(),
               bdop = sizeof(DopJob) * hdop.size(), bitem = sizeof(int) * hitem.size();
  if (!st) {
    if ((e = hipMalloc(&p->tables, bpc + bwf + bdop + bitem + 64)) != hipSuccess) {
      set_error(MXR_OUT_OF_MEMORY, "%s: table allocation failed", fn);
      st = MXR_OUT_OF_MEMORY;
    }
  }
  if (!st) {
    char *base = (char *)p->tables;
    PcJob *dpc = (PcJob *)base;
    WfArgs *dwf = (WfArgs *)(base + bpc);
    DopJob *ddop = (DopJob *)(base + bpc + bwf);
    int *ditem = (int *)(base + bpc + bwf + bdop);
    if ((e = hipMemcpy(dpc, hpc.data(), bpc, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(dwf, hwf.data(), bwf, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(ddop, hdop.data(), bdop, hipMemcpyHostToDevice)) != hipSuccess ||
        (bitem && (e = hipMemcpy(ditem, hitem.data(), bitem, hipMemcpyHostToDevice)) != hipSuccess)) {
      set_error(MXR_HIP_ERROR, "%s: table upload failed: %s", fn, hipGetErrorString(e));
      st = MXR_HIP_ERROR;
    }
    for (size_t c = 0; c < p->pc.size(); ++c) {
      p->pc[c].d_jobs = dpc + pc_ranges[c].first;
      p->pc[c].d_wf = dwf + pc_ranges[c].first;
      p->pc[c].d_item_job = ditem + pc_item0[c];
    }
    for (size_t c = 0; c < p->dop.size(); ++c) {
      p->dop[c].d_jobs = ddop + dop_ranges[c].first;
      p->dop[c].d_item_job = ditem + dop_item0[c];
    }
  }
  for (int b = 0; b < RAGGED_BRANCHES && !st; ++b) {
    if (hipStreamCreateWithFlags(&p->branch[b], hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&p->join[b], hipEventDisableTiming) != hipSuccess) {
      set_error(MXR_HIP_ERROR, "%s: branch stream creation failed", fn);
      st = MXR_HIP_ERROR;
    }
  }
  if (!st && hipEventCreateWithFlags(&p->fork, hipEventDisableTiming) != hipSuccess) {
    set_error(MXR_HIP_ERROR, "%s: event creation failed", fn);
    st = MXR_HIP_ERROR;
  }
  if (st) {
    free_plan(p);
    return st;
  }
  *out = reinterpret_cast<mxr_ragged_plan>(p);
  log_msg(LOG_DEBUG, "ragged plan: %d jobs, %zu pulse-compression classes, %zu Doppler chains", njobs,
          p->pc.size(), p->dop.size());
  return MXR_SUCCESS;
}

static mxr_status ragged_execute(mxr_ragged_plan h, hipStream_t s) {
  auto *p = reinterpret_cast<RaggedPlanObj *>(h);
  MXR_REQUIRE(p != nullptr, MXR_INVALID_PARAMETER, "ragged_execute: null plan");
  MXR_REQUIRE(device_id() == p->dev, MXR_INVALID_PARAMETER,
              "ragged_execute: plan was made on device %d, current device is %d", p->dev,
              device_id());
  // fork into the helper streams that hold a chain, run every chain on its
  // stream (its pulse-compression classes, then its Doppler pass), join
  int nb = 0;
  for (int b : p->dop_branch) nb = std::max(nb, b + 1);
  MXR_HIP_TRY(hipEventRecord(p->fork, s));
  for (int b = 0; b < nb; ++b) MXR_HIP_TRY(hipStreamWaitEvent(p->branch[b], p->fork, 0));
  mxr_status st = MXR_SUCCESS;
  for (size_t k = 0; k < p->dop_order.size() && !st; ++k) {
    const int d = p->dop_order[k];
    const DopClass &dc = p->dop[d];
    hipStream_t bs = p->branch[p->dop_branch[d]];
    for (const PcClass &c : p->pc) {
      if (c.nd != dc.nd || st) continue;
      st = launch_wfspec_batch(c.M, c.d_wf, c.njobs, bs);
      if (!st) st = launch_pc_batch(c.M, c.d_jobs, c.d_item_job, c.items, c.grid, bs);
    }
    if (!st) st = launch_doppler_batch(dc.nd, dc.d_jobs, dc.d_item_job, dc.items, dc.grid, bs);
  }
  for (int b = 0; b < nb; ++b) {
    MXR_HIP_TRY(hipEventRecord(p->join[b], p->branch[b]));
    MXR_HIP_TRY(hipStreamWaitEvent(s, p->join[b], 0));
  }
  return st;
}

static mxr_status ragged_block_size(mxr_ragged_plan h, int job, int *M) {
  auto *p = reinterpret_cast<RaggedPlanObj *>(h);
  MXR_REQUIRE(p != nullptr && M != nullptr, MXR_INVALID_PARAMETER, "ragged_plan_block_size: null");
  MXR_REQUIRE(job >= 0 && job < (int)p->M.size(), MXR_INVALID_PARAMETER,
              "ragged_plan_block_size: job %d out of range", job);
  *M = p->M[job];
  return MXR_SUCCESS;
}

}  // namespace mxr

extern "C" {

mxr_status mxr_ragged_plan_create(mxr_ragged_plan *plan, const mxr_ragged_job *jobs, int njobs,
                                  int flags, hipStream_t stream) {
  return mxr::guarded("mxr_ragged_plan_create",
                      [&] { return mxr::ragged_create(plan, jobs, njobs, flags, stream); });
}

mxr_status mxr_ragged_execute(mxr_ragged_plan plan, hipStream_t stream) {
  return mxr::guarded("mxr_ragged_execute", [&] { return mxr::ragged_execute(plan, stream); });
}

mxr_status mxr_ragged_plan_block_size(mxr_ragged_plan plan, int job, int *M) {
  return mxr::guarded("mxr_ragged_plan_block_size",
                      [&] { return mxr::ragged_block_size(plan, job, M); });
}

mxr_status mxr_ragged_plan_destroy(mxr_ragged_plan plan) {
  return mxr::guarded("mxr_ragged_plan_destroy", [&] {
    mxr::free_plan(reinterpret_cast<mxr::RaggedPlanObj *>(plan));
    return MXR_SUCCESS;
  });
}

}  // extern "C"
