// Pulse-axis pass fused with CA-CFAR: three-pulse canceller [1,-2,1] ->
// Doppler window -> nd-point Doppler FFT -> |.|^2 -> masked window sums ->
// threshold, one kernel, one read of the compressed cube and one write of
// the detection mask (replaces ThreePulseCanceller + DopplerProcessing +
// CFARDetections, examples/simple_radar_pipeline.h:321-357,394-436; the
// power map never leaves LDS).
//
// Tile: a workgroup owns TB = 64 consecutive range bins (one wave wide) of
// one channel across all pulses.  The CFAR window needs HX range bins on
// either side of a cell, so a tile computes the Doppler spectrum of all 64
// columns and emits detections for the OUT = 64 - 2*HXA inner columns (HXA =
// HX rounded up to even, so OUT is a multiple of 4 and every output quad
// starts on a 16-byte boundary of the detection rows); neighbouring tiles
// overlap by 2*HXA columns, which the XCD-local L2 serves.  The narrowest
// halo matters: at 16 x 8001 range bins and two resident workgroups per CU,
// 52 outputs per tile fill 4.8 rounds of the machine where 48 need 5.2.
//
//  1. load: lane = range column, wave = a strip of Doppler-input rows; each
//     lane reads its strip's pulses plus one on either side (512-byte
//     coalesced rows), forms the canceller output and the window in
//     registers and writes one LDS element per row (element (row i, column
//     t) at i*TB + t).
//  2. Stockham passes down the LDS columns (lane = column: unit-stride LDS);
//     the last pass writes |X|^2 over the first half of the tile.
//  3. CFAR: a thread owns 4 adjacent output columns on every 32nd Doppler
//     row (16 lanes per row).  It reads each of its power rows with 16-byte LDS loads and forms
//     the positive partial sums L (left of the hole), M (hole columns), Rt
//     (right of the hole) of its 4 columns once per row; lr = L + Rt and
//     full = lr + M go back to LDS, and each cell's noise sum adds full over
//     the window rows outside the hole's Doppler extent and lr over the rows
//     inside it.  One 16-byte detection quad per row.  Sums never subtract,
//     so a strong cell in the hole cannot cancel (the same decomposition as
//     the stand-alone CFAR kernel).
#include <cmath>
#include <cstdlib>

#include "fft_engine.hpp"
#include "mxr_cache.hpp"

namespace mxr {

constexpr int DC_MAXCELL = 255;  // largest window area with a fused instance

struct DcArgs {
  const float2 *y;  // [C, P, R], unit range stride
  int64_t ys_c;
  int ys_p;
  int *dets;  // [C, nd, R], unit range stride
  int64_t ds_c;
  int ds_k;
  float *ba;  // optional, unit range stride
  int64_t bs_c;
  int bs_k;
  float *pw;  // optional power output, unit range stride
  int64_t ps_c;
  int ps_k;
  int C, P, R;
  int rows_t;    // canceller rows fed to the FFT
  int p_off;     // canceller row i is centred on pulse i + p_off
  int win_rows;  // rows < win_rows get hamming(win_rows)
  int no_tpc;    // y already holds the canceller output
  int tiles;     // range tiles per channel
  int vec_out;   // every output row is 16-byte aligned
  const float2 *tw;
  // threshold factor per in-bounds cell count n: alpha(n) (exact form) or
  // alpha(n)/n, alpha(n) = n*(pfa^(-1/n) - 1) evaluated in float on the host
  float thr[DC_MAXCELL + 1];
};

// Element access at a 32-bit byte offset from a wave-uniform base (lets the
// compiler use the scalar-base + 32-bit vector-offset addressing form).
template <class T>
__device__ __forceinline__ T ld32(const T *base, uint32_t byte_off) {
  return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + byte_off);
}
template <class T>
__device__ __forceinline__ void st32(T *base, uint32_t byte_off, T v) {
  *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + byte_off) = v;
}

constexpr int DC_TB = 64;   // computed range columns per tile
constexpr int DC_NT = 512;  // threads per workgroup
constexpr int DC_PS = 64;   // power-tile row stride (floats)

template <int HX>
struct DcGeom {
  static constexpr int HXA = (HX + 1) & ~1;
  static constexpr int OUT = DC_TB - 2 * HXA;
};

template <int W>
__device__ __forceinline__ void dc_win4(const float *u, float (&o)[4]) {
  if constexpr (W <= 0) {
    o[0] = o[1] = o[2] = o[3] = 0.f;
  } else if constexpr (W == 1) {
#pragma unroll
    for (int j = 0; j < 4; ++j) o[j] = u[j];
  } else if constexpr (W == 2) {
#pragma unroll
    for (int j = 0; j < 4; ++j) o[j] = u[j] + u[j + 1];
  } else if constexpr (W == 3) {
    const float p1 = u[1] + u[2], p3 = u[3] + u[4];
    o[0] = u[0] + p1;
    o[1] = p1 + u[3];
    o[2] = u[2] + p3;
    o[3] = p3 + u[5];
  } else {
    float core = u[3];
#pragma unroll
    for (int i = 4; i < W; ++i) core += u[i];
    const float d = u[1] + u[2], e = u[W] + u[W + 1];
    o[0] = (u[0] + d) + core;
    o[1] = (d + core) + u[W];
    o[2] = (u[2] + core) + e;
    o[3] = (core + e) + u[W + 2];
  }
}

// Two 512-thread workgroups per CU (LDS bound): 4 waves per SIMD, so at most
// 128 VGPRs.
template <int ND, int HY, int HX, int EY, int EX, bool EXACT>
__global__ __launch_bounds__(DC_NT) __attribute__((amdgpu_waves_per_eu(4)))
void dopcfar_kernel(DcArgs a) {
  constexpr int TB = DC_TB, NT = DC_NT, NW = NT / 64;
  constexpr int HXA = DcGeom<HX>::HXA, OUT = DcGeom<HX>::OUT;
  static_assert(OUT >= 16, "window too wide for a 64-column tile");
  static_assert(ND >= NW, "Doppler size below one row per wave");
  constexpr int PS = DC_PS;
  constexpr int SR = ND / NW;  // Doppler-input rows per wave in the load phase
  using LY = Layout<ND, TB, true>;
  using SQ = Seq<ND, false>;
  constexpr int K = SQ::n;
  constexpr int NCELL = (2 * HY + 1) * (2 * HX + 1);
  static_assert(NCELL <= DC_MAXCELL, "window area");
  // the complex tile; after the transform it holds the power rows (stride
  // PS, HY zero rows either side) followed by the CFAR's full-sum rows
  // (stride OUT)
  constexpr int LDS2 = ND * TB > ((ND + 2 * HY) * (PS + OUT) + 1) / 2 ? ND * TB
                                                                      : ((ND + 2 * HY) * (PS + OUT) + 1) / 2;
  __shared__ __align__(16) float2 lds[LDS2];
  static_assert((ND + 2 * HY) * (PS + OUT) <= 2 * LDS2, "power and full-sum rows must fit the tile");
  __shared__ float2 coef[ND];  // Doppler-input row i = coef.x*(y[p-1]+y[p+1]) + coef.y*y[p]
  __shared__ float thr_tab[NCELL + 1];

  const int G = a.tiles * a.C;
  const int g = (int)xcd_remap(blockIdx.x, G);
  const int c = g / a.tiles, tile = g - c * a.tiles;
  const int r0 = tile * OUT - HXA;  // range bin of tile column 0
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);

  // canceller [1,-2,1] and window folded into two coefficients per row:
  // rows past the canceller output are zero, NONE passes y[p] windowed
  for (int i = threadIdx.x; i < ND; i += NT) {
    float h = 0.f;
    if (i < a.rows_t) h = i < a.win_rows ? hamming_f32(i, a.win_rows) : 1.0f;
    coef[i] = a.no_tpc ? make_float2(0.f, h) : make_float2(h, -2.0f * h);
  }
  for (int n = threadIdx.x; n <= NCELL; n += NT) thr_tab[n] = a.thr[n];
  __syncthreads();

  // ---- 1. canceller + window in registers, one LDS write per element.
  // Indices are clamped into the cube so every load is unconditional; the
  // values outside it are replaced by zeros afterwards.
  {
    const int r = r0 + lane;
    const bool rok = r >= 0 && r < a.R;
    const int rc = min(max(r, 0), a.R - 1);
    const float2 *chan = a.y + (int64_t)c * a.ys_c;
    const int i0 = wv * SR;
    const int pb = i0 + a.p_off - 1;  // pulse held in v[0]
    float2 v[SR + 2];
#pragma unroll
    for (int u = 0; u < SR + 2; ++u) {
      const int p = pb + u;
      const int pc = min(max(p, 0), a.P - 1);
      const float2 t = ld32(chan, (uint32_t)(rc + pc * a.ys_p) * 8u);
      const bool in = rok && p >= 0 && p < a.P;
      v[u] = make_float2(in ? t.x : 0.f, in ? t.y : 0.f);
    }
#pragma unroll
    for (int u = 0; u < SR; ++u) {
      const int i = i0 + u;
      const float2 cf = coef[i];
      const float2 x = make_float2(cf.x * (v[u].x + v[u + 2].x) + cf.y * v[u + 1].x,
                                   cf.x * (v[u].y + v[u + 2].y) + cf.y * v[u + 1].y);
      lds[LY::addr(lane, i)] = x;
    }
  }
  wg_barrier();

  // ---- 2. Doppler FFT down the columns; the last pass leaves |X|^2 in LDS
  // as rows 0..ND-1 of a power tile with HY zero rows above and below (the
  // CFAR's zero padding along Doppler)
  float *const pwt = reinterpret_cast<float *>(lds) + HY * PS;  // row 0
  if constexpr (K > 1) lds_passes<LY, SQ, ND, NT, TB, false, 0, K - 1>(lds, a.tw);
  {
    constexpr int RL = SQ::template R<K - 1>;
    constexpr int LSL = SQ::template LS<K - 1>;
    Pass<LY, RL, ND, NT, TB> pl;
    pl.load([&](int t, int i) { return lds[LY::addr(t, i)]; });
    wg_barrier();  // the power map overwrites the tile
    for (int i = threadIdx.x; i < 2 * HY * TB; i += NT) {
      const int row = i / TB < HY ? i / TB - HY : ND + i / TB - HY;
      pwt[row * PS + (i % TB)] = 0.f;
    }
    pl.template compute<LSL, false>(a.tw + SQ::template TWOFF<K - 1>);
    pl.template store<LSL>([&](int t, int k, float2 x) { pwt[k * PS + t] = cabs2(x); });
  }
  wg_barrier();

  // ---- 3. CFAR over the inner columns, in two steps over the same cells.
  // A thread owns one output column quad (tile columns tc0..tc0+3) on the
  // rows k = kb + KS*it.
  //  3a. per row: the positive partial sums of the row segment, lr = left +
  //      right of the hole and full = lr + hole columns (each row's sums are
  //      formed once; full goes to the free upper part of the tile, lr
  //      replaces the power row once every thread has read its row);
  //  3b. per cell: noise = sum over the window rows of full (rows outside the
  //      hole's Doppler extent) or lr (rows inside it), then the threshold.
  // The cell under test and the row's own lr stay in registers in between.
  constexpr int NQO = OUT / 4;                // output column quads per row
  // 16 lanes per row (quads q < NQO active): the lanes of one 16-lane LDS
  // access group read one row, so their 16-byte reads never collide in a
  // bank (rows of different lanes 256 bytes apart would)
  constexpr int GQ = 16;
  static_assert(NQO <= GQ, "output quads per row");
  constexpr int KS = NT / GQ;                 // row step between a thread's rows
  constexpr int QIT = (ND + KS - 1) / KS;     // rows per thread (at most)
  constexpr int NQ = 1 + 2 * HXA / 4;         // power quads read per row segment
  constexpr int NV = 4 * NQ;
  constexpr int FS = OUT;                     // full-sum row stride (floats)
  float *const fsum = reinterpret_cast<float *>(lds) + (ND + 2 * HY) * PS + HY * FS;  // row 0
  const int q = threadIdx.x % GQ, kb = threadIdx.x / GQ;
  const bool act = q < NQO;
  const int tc0 = HXA + 4 * q;                // first tile column of the quad
  float cut[QIT][4], lrk[QIT][4];
  // zero rows of the full sums (the CFAR's zero padding along Doppler)
  for (int i = threadIdx.x; i < 2 * HY * NQO; i += NT) {
    const int row = i / NQO < HY ? i / NQO - HY : ND + i / NQO - HY;
    *reinterpret_cast<float4 *>(fsum + row * FS + 4 * (i % NQO)) = make_float4(0.f, 0.f, 0.f, 0.f);
  }
#pragma unroll
  for (int it = 0; it < QIT; ++it) {
    const int k = kb + KS * it;
    if (!act || k >= ND) continue;
    float u[NV];
    const float *row = pwt + k * PS + tc0 - HXA;
#pragma unroll
    for (int m = 0; m < NQ; ++m) {
      const float4 z = *reinterpret_cast<const float4 *>(row + 4 * m);
      u[4 * m] = z.x;
      u[4 * m + 1] = z.y;
      u[4 * m + 2] = z.z;
      u[4 * m + 3] = z.w;
    }
    // u[HXA + j + d] is tile column tc0 + j + d
    float L[4], Mh[4], Rt[4], fl[4];
    dc_win4<HX - EX>(u + HXA - HX, L);
    dc_win4<2 * EX + 1>(u + HXA - EX, Mh);
    dc_win4<HX - EX>(u + HXA + EX + 1, Rt);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      lrk[it][j] = L[j] + Rt[j];
      fl[j] = lrk[it][j] + Mh[j];
      cut[it][j] = u[HXA + j];
    }
    *reinterpret_cast<float4 *>(fsum + k * FS + 4 * q) = make_float4(fl[0], fl[1], fl[2], fl[3]);
  }
  wg_barrier();  // every power row has been read
  // lr in place of the power rows (columns 0..OUT-1; the zero rows stay zero)
#pragma unroll
  for (int it = 0; it < QIT; ++it) {
    const int k = kb + KS * it;
    if (!act || k >= ND) continue;
    *reinterpret_cast<float4 *>(pwt + k * PS + 4 * q) =
        make_float4(lrk[it][0], lrk[it][1], lrk[it][2], lrk[it][3]);
  }
  wg_barrier();

  const int gr0 = r0 + tc0;  // range bin of the quad (a multiple of 4, >= 0)
  if (!act || gr0 >= a.R) return;  // no barrier follows
  const int R = a.R;
  // in-bounds cell count = ro*co - rh*chh; away from the first and last HY
  // Doppler rows ro = 2HY+1 and rh = 2EY+1, so it depends on the column only
  int co[4], chh[4], nrm_int[4];
  float thr_int[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int gr = gr0 + j;
    co[j] = min(gr + HX, R - 1) - max(gr - HX, 0) + 1;
    chh[j] = min(gr + EX, R - 1) - max(gr - EX, 0) + 1;
    nrm_int[j] = (2 * HY + 1) * co[j] - (2 * EY + 1) * chh[j];
    thr_int[j] = thr_tab[min(max(nrm_int[j], 0), NCELL)];
  }
  const bool vec = a.vec_out && gr0 + 3 < R;
  int *const dets_c = a.dets + (int64_t)c * a.ds_c;
  float *const ba_c = EXACT ? a.ba + (int64_t)c * a.bs_c : nullptr;
  float *const pw_c = a.pw ? a.pw + (int64_t)c * a.ps_c : nullptr;

#pragma unroll
  for (int it = 0; it < QIT; ++it) {
    const int kc = kb + KS * it;  // Doppler row of the cells under test
    if (kc >= ND) continue;
    float sm[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) sm[j] = 0.f;
#pragma unroll
    for (int dy = -HY; dy <= HY; ++dy) {
      float t[4];
      if (dy == 0) {
#pragma unroll
        for (int j = 0; j < 4; ++j) t[j] = lrk[it][j];
      } else {
        const bool inner = dy >= -EY && dy <= EY;
        const float *src = inner ? pwt + (kc + dy) * PS + 4 * q : fsum + (kc + dy) * FS + 4 * q;
        const float4 z = *reinterpret_cast<const float4 *>(src);
        t[0] = z.x;
        t[1] = z.y;
        t[2] = z.z;
        t[3] = z.w;
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) sm[j] = dy == -HY ? t[j] : sm[j] + t[j];
    }
    float thr[4];
    int nrm[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      thr[j] = thr_int[j];
      nrm[j] = nrm_int[j];
    }
    if (kc < HY || kc + HY >= ND) {  // rows whose window leaves the map
      const int ro = min(kc + HY, ND - 1) - max(kc - HY, 0) + 1;
      const int rh = min(kc + EY, ND - 1) - max(kc - EY, 0) + 1;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        nrm[j] = ro * co[j] - rh * chh[j];
        thr[j] = thr_tab[min(max(nrm[j], 0), NCELL)];
      }
    }
    const float *cv = cut[it];
    int det[4];
    float bav[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if constexpr (EXACT) {
        bav[j] = sm[j] / (float)nrm[j];
        det[j] = cv[j] > thr[j] * bav[j] ? 1 : 0;
      } else {
        det[j] = cv[j] > thr[j] * sm[j] ? 1 : 0;
      }
    }
    const uint32_t od = (uint32_t)(kc * a.ds_k + gr0) * 4u;
    const uint32_t ob = EXACT ? (uint32_t)(kc * a.bs_k + gr0) * 4u : 0u;
    const uint32_t op = pw_c ? (uint32_t)(kc * a.ps_k + gr0) * 4u : 0u;
    if (vec) {
      *reinterpret_cast<int4 *>(reinterpret_cast<char *>(dets_c) + od) =
          make_int4(det[0], det[1], det[2], det[3]);
      if constexpr (EXACT)
        *reinterpret_cast<float4 *>(reinterpret_cast<char *>(ba_c) + ob) =
            make_float4(bav[0], bav[1], bav[2], bav[3]);
      if (pw_c)
        *reinterpret_cast<float4 *>(reinterpret_cast<char *>(pw_c) + op) =
            make_float4(cv[0], cv[1], cv[2], cv[3]);
    } else {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        if (gr0 + j >= R) break;
        st32(dets_c, od + 4u * j, det[j]);
        if constexpr (EXACT) st32(ba_c, ob + 4u * j, bav[j]);
        if (pw_c) st32(pw_c, op + 4u * j, cv[j]);
      }
    }
  }
}

template <int ND, int HY, int HX, int EY, int EX>
static mxr_status launch_dopcfar(DcArgs a, hipStream_t s) {
  a.tiles = (a.R + DcGeom<HX>::OUT - 1) / DcGeom<HX>::OUT;
  const int64_t G = (int64_t)a.tiles * a.C;
  MXR_REQUIRE(G < (1ll << 31), MXR_INVALID_SIZE, "doppler_cfar: too many tiles");
  {
    ProfScope ps("dopcfar_kernel", s);
    if (a.ba)
      dopcfar_kernel<ND, HY, HX, EY, EX, true><<<dim3((unsigned)G), dim3(DC_NT), 0, s>>>(a);
    else
      dopcfar_kernel<ND, HY, HX, EY, EX, false><<<dim3((unsigned)G), dim3(DC_NT), 0, s>>>(a);
  }
  MXR_HIP_TRY(hipGetLastError());
  return MXR_SUCCESS;
}

template <int HY, int HX, int EY, int EX>
static mxr_status dopcfar_nd(int nd, const DcArgs &a, hipStream_t s) {
  switch (nd) {
    case 8: return launch_dopcfar<8, HY, HX, EY, EX>(a, s);
    case 16: return launch_dopcfar<16, HY, HX, EY, EX>(a, s);
    case 32: return launch_dopcfar<32, HY, HX, EY, EX>(a, s);
    case 64: return launch_dopcfar<64, HY, HX, EY, EX>(a, s);
    case 128: return launch_dopcfar<128, HY, HX, EY, EX>(a, s);
    default: return MXR_NOT_SUPPORTED;
  }
}

mxr_status doppler_impl(const mxr_tensor *pow_out, const mxr_tensor *rd_out, const mxr_tensor *y,
                        int tpc_mode, int64_t win_rows, int64_t nd, hipStream_t stream);
mxr_status cfar_impl(const mxr_tensor *dets, const mxr_tensor *ba_out, const mxr_tensor *pow_out,
                     const mxr_tensor *in, const mxr_cfar_params *pp, hipStream_t stream);

// Unfused fallback: pulse-axis pass into the power map (the caller's, or a
// stream-ordered temporary), then the stand-alone CFAR kernel.
static mxr_status dopcfar_unfused(const mxr_tensor *dets, const mxr_tensor *ba_out,
                                  const mxr_tensor *pow_out, const mxr_tensor *y, int tpc_mode,
                                  int64_t win_rows, int64_t nd, const mxr_cfar_params *pp,
                                  hipStream_t s) {
  mxr_status st;
  if (pow_out) {
    if ((st = doppler_impl(pow_out, nullptr, y, tpc_mode, win_rows, nd, s))) return st;
    return cfar_impl(dets, ba_out, nullptr, pow_out, pp, s);
  }
  const int64_t C = y->shape[0], R = y->shape[2];
  void *tmp = nullptr;
  MXR_HIP_TRY(hipMallocAsync(&tmp, (size_t)(C * nd * R) * sizeof(float), s));
  mxr_tensor pw{};
  pw.data = tmp;
  pw.dtype = MXR_F32;
  pw.rank = 3;
  pw.shape[0] = C;
  pw.shape[1] = nd;
  pw.shape[2] = R;
  pw.stride[0] = nd * R;
  pw.stride[1] = R;
  pw.stride[2] = 1;
  st = doppler_impl(&pw, nullptr, y, tpc_mode, win_rows, nd, s);
  if (!st) st = cfar_impl(dets, ba_out, nullptr, &pw, pp, s);
  const hipError_t e = hipFreeAsync(tmp, s);
  if (!st && e != hipSuccess) {
    set_error(MXR_HIP_ERROR, "doppler_cfar: hipFreeAsync failed: %s", hipGetErrorString(e));
    return MXR_HIP_ERROR;
  }
  return st;
}

mxr_status doppler_cfar_impl(const mxr_tensor *dets, const mxr_tensor *ba_out,
                             const mxr_tensor *pow_out, const mxr_tensor *y, int tpc_mode,
                             int64_t win_rows, int64_t nd, const mxr_cfar_params *pp,
                             hipStream_t stream) {
  mxr_status st;
  MXR_REQUIRE(pp != nullptr, MXR_INVALID_PARAMETER, "doppler_cfar: null params");
  if ((st = check_tensor(y, "doppler_cfar input", MXR_C64, 3, 3))) return st;
  if ((st = check_tensor(dets, "doppler_cfar detections", MXR_I32, 3, 3))) return st;
  if (ba_out && (st = check_tensor(ba_out, "doppler_cfar ba", MXR_F32, 3, 3))) return st;
  if (pow_out && (st = check_tensor(pow_out, "doppler_cfar power", MXR_F32, 3, 3))) return st;
  MXR_REQUIRE(tpc_mode == MXR_CONV_SAME || tpc_mode == MXR_CONV_VALID || tpc_mode == MXR_TPC_NONE,
              MXR_INVALID_PARAMETER, "doppler_cfar: canceller mode must be SAME, VALID or NONE");
  const int64_t C = y->shape[0], P = y->shape[1], R = y->shape[2];
  MXR_REQUIRE(P >= 3 || tpc_mode == MXR_TPC_NONE, MXR_INVALID_SIZE,
              "doppler_cfar: need at least 3 pulses, got %lld", (long long)P);
  const int64_t rows_t = tpc_mode == MXR_CONV_VALID ? P - 2 : P;
  MXR_REQUIRE(nd >= rows_t, MXR_INVALID_SIZE,
              "doppler_cfar: %lld-point FFT is shorter than the %lld canceller rows",
              (long long)nd, (long long)rows_t);
  MXR_REQUIRE(nd >= 2 && nd <= 1024 && (nd & (nd - 1)) == 0, MXR_NOT_SUPPORTED,
              "doppler_cfar: FFT size %lld must be a power of two in [2, 1024]", (long long)nd);
  MXR_REQUIRE(win_rows >= 0 && win_rows <= rows_t, MXR_INVALID_SIZE,
              "doppler_cfar: bad window length");
  for (const mxr_tensor *o : {dets, ba_out, pow_out}) {
    if (!o) continue;
    MXR_REQUIRE(o->shape[0] == C && o->shape[1] == nd && o->shape[2] == R, MXR_INVALID_SIZE,
                "doppler_cfar: outputs must be [%lld, %lld, %lld]", (long long)C, (long long)nd,
                (long long)R);
    MXR_REQUIRE(!views_overlap(*o, *y), MXR_ALIAS_ERROR, "doppler_cfar: an output aliases the input");
  }
  const mxr_tensor *outs[3] = {dets, ba_out, pow_out};
  for (int i = 0; i < 3; ++i)
    for (int j = i + 1; j < 3; ++j)
      if (outs[i] && outs[j])
        MXR_REQUIRE(!views_overlap(*outs[i], *outs[j]), MXR_ALIAS_ERROR,
                    "doppler_cfar: outputs overlap");
  const int wy = pp->win_y, wx = pp->win_x, gy = pp->guard_y, gx = pp->guard_x;
  MXR_REQUIRE(wy >= 1 && wx >= 1 && (wy & 1) && (wx & 1), MXR_INVALID_PARAMETER,
              "doppler_cfar: window %dx%d must have odd positive sides", wy, wx);
  MXR_REQUIRE(gy >= 1 && gx >= 1 && (gy & 1) && (gx & 1) && gy <= wy && gx <= wx,
              MXR_INVALID_PARAMETER, "doppler_cfar: guard %dx%d must be odd and inside the window",
              gy, gx);
  MXR_REQUIRE(pp->pfa > 0.f && pp->pfa < 1.f, MXR_INVALID_PARAMETER,
              "doppler_cfar: pfa must be in (0,1)");
  if (C == 0 || R == 0) return MXR_SUCCESS;
  MXR_REQUIRE(C <= 65535 && R < (1ll << 30) && P < (1 << 20), MXR_INVALID_SIZE,
              "doppler_cfar: shape too large");

  auto unit_rows = [](const mxr_tensor *t) { return !t || t->stride[2] == 1; };
  // the kernel addresses one channel of each tensor with 32-bit byte offsets
  auto span_ok = [&](const mxr_tensor *t, int64_t rows, int64_t esz) {
    return !t || (t->stride[1] >= 0 && ((rows - 1) * t->stride[1] + R) * esz < (1ll << 32));
  };
  const bool fusable = nd <= 128 && nd >= 8 && unit_rows(y) && unit_rows(dets) &&
                       unit_rows(ba_out) && unit_rows(pow_out) && span_ok(y, P, 8) &&
                       span_ok(dets, nd, 4) && span_ok(ba_out, nd, 4) && span_ok(pow_out, nd, 4);
  DcArgs a{};
  a.y = (const float2 *)y->data;
  a.ys_c = y->stride[0];
  a.ys_p = (int)y->stride[1];
  a.dets = (int *)dets->data;
  a.ds_c = dets->stride[0];
  a.ds_k = (int)dets->stride[1];
  if (ba_out) {
    a.ba = (float *)ba_out->data;
    a.bs_c = ba_out->stride[0];
    a.bs_k = (int)ba_out->stride[1];
  }
  if (pow_out) {
    a.pw = (float *)pow_out->data;
    a.ps_c = pow_out->stride[0];
    a.ps_k = (int)pow_out->stride[1];
  }
  a.C = (int)C;
  a.P = (int)P;
  a.R = (int)R;
  a.rows_t = (int)rows_t;
  a.p_off = tpc_mode == MXR_CONV_VALID ? 1 : 0;
  a.no_tpc = tpc_mode == MXR_TPC_NONE;
  a.win_rows = (int)win_rows;
  auto al16 = [](const mxr_tensor *t) {
    return !t || ((uintptr_t)t->data % 16 == 0 && t->stride[1] % 4 == 0 && t->stride[0] % 4 == 0);
  };
  a.vec_out = al16(dets) && al16(ba_out) && al16(pow_out);
  if (fusable) {
    const int ncell = wy * wx;
    a.thr[0] = 0.f;
    for (int n = 1; n <= DC_MAXCELL && n <= ncell; ++n) {
      const float nf = (float)n;
      const float alpha = nf * (std::pow(pp->pfa, -1.0f / nf) - 1.0f);
      a.thr[n] = ba_out ? alpha : alpha / nf;
    }
    a.tw = twiddles_device((int)nd, false);
    MXR_REQUIRE(a.tw != nullptr, MXR_OUT_OF_MEMORY, "doppler_cfar: twiddle table allocation failed");
    const int hy = wy / 2, hx = wx / 2, ey = gy / 2, ex = gx / 2;
    st = MXR_NOT_SUPPORTED;
#define MXR_DC_CASE(HY_, HX_, EY_, EX_) \
  else if (hy == HY_ && hx == HX_ && ey == EY_ && ex == EX_) st = dopcfar_nd<HY_, HX_, EY_, EX_>((int)nd, a, stream);
    if (false) {
    }
    MXR_DC_CASE(2, 6, 1, 1)  // 5 x 13 window, 3 x 3 hole (MatX mask)
    MXR_DC_CASE(2, 6, 1, 2)  // 5 x 13 window, 3 x 5 hole
    MXR_DC_CASE(1, 2, 0, 1)
    MXR_DC_CASE(2, 4, 1, 1)
    MXR_DC_CASE(3, 8, 1, 2)
#undef MXR_DC_CASE
    if (st != MXR_NOT_SUPPORTED) return st;
  }
  return dopcfar_unfused(dets, ba_out, pow_out, y, tpc_mode, win_rows, nd, pp, stream);
}

}  // namespace mxr

extern "C" mxr_status mxr_doppler_cfar(const mxr_tensor *dets, const mxr_tensor *ba_out,
                                       const mxr_tensor *pow_out, const mxr_tensor *y,
                                       int tpc_mode, int64_t win_rows, int64_t nd,
                                       const mxr_cfar_params *params, hipStream_t stream) {
  return mxr::guarded("mxr_doppler_cfar", [&] {
    return mxr::doppler_cfar_impl(dets, ba_out, pow_out, y, tpc_mode, win_rows, nd, params, stream);
  });
}
