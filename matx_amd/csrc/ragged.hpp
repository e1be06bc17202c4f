// Argument blocks shared by the single-call kernels and the ragged-batch
// (persistent, many-job) kernels of the pulse-compression and pulse-axis
// passes, and the host launchers of the batched kernels.
#pragma once

#include "mxr_internal.hpp"

namespace mxr {

// Pulse-axis pass (canceller, window, Doppler FFT, |.|^2) of one cube.
struct DopArgs {
  const float2 *y;
  int64_t ys_c, ys_p, ys_r;
  float2 *rd;
  int64_t rs_c, rs_k, rs_r;
  float *pw;
  int64_t ps_c, ps_k, ps_r;
  int C, P, R;
  int rows_t;    // canceller output rows fed to the FFT (P or P-2)
  int p_off;     // canceller row i is centred on pulse i + p_off
  int win_rows;  // rows < win_rows get hamming(win_rows)
  int no_tpc;    // input is already the canceller output
  const float *taps;  // optional real window replacing hamming(win_rows)
  int64_t taps_s;
  const float2 *tw;
};

// Waveform spectrum: Wc = conj?(fft_M(pad_M(w2))), w2 = w1/||w1|| (optional),
// w1 = w .* hamming(L) (optional); optionally writes w2 back to w.
struct WfArgs {
  float2 *w;
  int64_t wstride;
  int64_t L;
  int window, normalize, writeback, conj;
  float2 *wc;
  const float2 *tw;
  float out_scale;  // applied to the stored spectrum (the block kernels' 1/M)
};

// One job of a ragged pulse-compression batch: rows (c, p) of a [C, P, S]
// cube, overlap-save blocks of the class's size M.
struct PcJob {
  const float2 *x;
  int64_t x_cs, x_ps;
  float2 *y;
  int64_t y_cs, y_ps;
  int C, P;
  int seff, n, nout, V, nblk;
  const float2 *wc;  // the job's spectrum, scaled by 1/M
  int64_t item0;  // first batch item (row-block) of this job
};

// One job of a ragged pulse-axis batch: items are (channel, range tile).
struct DopJob {
  DopArgs a;
  int tiles;
  int64_t item0;
};

// A wave-uniform copy of a job record: every 32-bit word goes through
// readfirstlane, so the fields live in scalar registers for the item.
template <class T>
__device__ __forceinline__ T uniform_copy(const T *src) {
  static_assert(sizeof(T) % 4 == 0, "record size");
  T out;
  const uint32_t *s = reinterpret_cast<const uint32_t *>(src);
  uint32_t *d = reinterpret_cast<uint32_t *>(&out);
#pragma unroll
  for (int i = 0; i < (int)(sizeof(T) / 4); ++i) d[i] = __builtin_amdgcn_readfirstlane(s[i]);
  return out;
}

// Hides a pointer's value from the optimiser for one loop trip, so loads
// through it are not hoisted out of a persistent loop (and kept live in
// registers across every item).
template <class T>
__device__ __forceinline__ T *opaque(T *p) {
  asm volatile("" : "+s"(p));
  return p;
}

// Builds the tables block kernels of size M read (no allocation at launch).
mxr_status pc_tables_prepare(int M, const char *fn);
mxr_status pc_plan_blocks(int64_t seff, int64_t L, int64_t n, int64_t nout, int M_req, int *M,
                          int64_t *V, int64_t *nblk);
mxr_status launch_wfspec_batch(int M, const WfArgs *d_args, int njobs, hipStream_t s);
mxr_status launch_pc_batch(int M, const PcJob *d_jobs, const int *d_item_job, int64_t items, int grid,
                           hipStream_t s);
int pc_batch_grid(int M);
mxr_status launch_doppler_batch(int nd, const DopJob *d_jobs, const int *d_item_job, int64_t items,
                                int grid,
                                hipStream_t s);
int doppler_batch_tile(int nd);

}  // namespace mxr
