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
// columns and emits detections for the OUT = 64 - 2*HXA inner columns
// (HXA = HX rounded up to a multiple of 4, so output quads stay 16-byte
// aligned); neighbouring tiles overlap by 2*HXA columns, which the XCD-local
// L2 serves.
//
//  1. load: lane = range column, wave = a strip of Doppler-input rows; each
//     lane reads its strip's pulses plus one on either side (512-byte
//     coalesced rows), forms the canceller output and the window in
//     registers and writes one LDS element per row (element (row i, column
//     t) at i*TB + t).
//  2. Stockham passes down the LDS columns (lane = column: unit-stride LDS);
//     the last pass writes |X|^2 over the first half of the tile.
//  3. CFAR: a thread owns 4 adjacent columns and a band of Doppler rows,
//     reads each power row with 16-byte LDS loads, forms the positive
//     partial sums L (left of the hole), M (hole columns), Rt (right of the
//     hole) of its 4 columns, keeps 2HY+1 rows of them in registers and
//     writes one 16-byte detection quad per row.  Sums never subtract, so a
//     strong cell in the hole cannot cancel (the same decomposition as the
//     stand-alone CFAR kernel).
#include <cstdlib>

#include "fft_engine.hpp"
#include "mxr_cache.hpp"

namespace mxr {

struct DcArgs {
  const float2 *y;  // [C, P, R], unit range stride
  int64_t ys_c, ys_p;
  int *dets;  // [C, nd, R], unit range stride
  int64_t ds_c, ds_k;
  float *ba;  // optional, unit range stride
  int64_t bs_c, bs_k;
  float *pw;  // optional power output, unit range stride
  int64_t ps_c, ps_k;
  int C, P, R;
  int rows_t;    // canceller rows fed to the FFT
  int p_off;     // canceller row i is centred on pulse i + p_off
  int win_rows;  // rows < win_rows get hamming(win_rows)
  int no_tpc;    // y already holds the canceller output
  int tiles;     // range tiles per channel
  int vec_out;   // every output row is 16-byte aligned
  float pfa;
  const float2 *tw;
};

constexpr int DC_TB = 64;   // computed range columns per tile
constexpr int DC_NT = 512;  // threads per workgroup

template <int HX>
struct DcGeom {
  static constexpr int HXA = (HX + 3) & ~3;
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

template <int ND, int HY, int HX, int EY, int EX, bool EXACT>
__global__ __launch_bounds__(DC_NT) void dopcfar_kernel(DcArgs a) {
  constexpr int TB = DC_TB, NT = DC_NT, NW = NT / 64;
  constexpr int HXA = DcGeom<HX>::HXA, OUT = DcGeom<HX>::OUT;
  static_assert(OUT >= 16, "window too wide for a 64-column tile");
  static_assert(ND >= NW, "Doppler size below one row per wave");
  constexpr int SR = ND / NW;  // Doppler-input rows per wave in the load phase
  using LY = Layout<ND, TB, true>;
  using SQ = Seq<ND, false>;
  constexpr int K = SQ::n;
  constexpr int NCELL = (2 * HY + 1) * (2 * HX + 1);
  __shared__ __align__(16) float2 lds[ND * TB];
  __shared__ float win[ND];
  __shared__ float thr_tab[NCELL + 1];

  const int G = a.tiles * a.C;
  const int g = (int)xcd_remap(blockIdx.x, G);
  const int c = g / a.tiles, tile = g - c * a.tiles;
  const int r0 = tile * OUT - HXA;  // range bin of tile column 0
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;

  for (int i = threadIdx.x; i < ND; i += NT) {
    float h = 0.f;
    if (i < a.rows_t) h = i < a.win_rows ? hamming_f32(i, a.win_rows) : 1.0f;
    win[i] = h;
  }
  for (int n = threadIdx.x; n <= NCELL; n += NT) {
    const float nf = (float)n;
    const float alpha = nf * (powf(a.pfa, -1.0f / nf) - 1.0f);
    thr_tab[n] = EXACT ? alpha : alpha / nf;
  }
  __syncthreads();

  // ---- 1. canceller + window in registers, one LDS write per element
  {
    const int r = r0 + lane;
    const bool rok = r >= 0 && r < a.R;
    const float2 *col = a.y + (int64_t)c * a.ys_c + (rok ? r : 0);
    const int i0 = wv * SR;
    const int pb = i0 + a.p_off - 1;  // pulse held in v[0]
    float2 v[SR + 2];
#pragma unroll
    for (int u = 0; u < SR + 2; ++u) {
      const int p = pb + u;
      v[u] = (rok && p >= 0 && p < a.P) ? col[(int64_t)p * a.ys_p] : make_float2(0.f, 0.f);
    }
#pragma unroll
    for (int u = 0; u < SR; ++u) {
      const int i = i0 + u;
      float2 x;
      if (i >= a.rows_t) {
        x = make_float2(0.f, 0.f);
      } else if (a.no_tpc) {
        x = cscale(v[u + 1], win[i]);
      } else {
        x = make_float2(v[u].x + v[u + 2].x - 2.0f * v[u + 1].x,
                        v[u].y + v[u + 2].y - 2.0f * v[u + 1].y);
        x = cscale(x, win[i]);
      }
      lds[LY::addr(lane, i)] = x;
    }
  }
  wg_barrier();

  // ---- 2. Doppler FFT down the columns; the last pass leaves |X|^2 in LDS
  if constexpr (K > 1) lds_passes<LY, SQ, ND, NT, TB, false, 0, K - 1>(lds, a.tw);
  {
    constexpr int RL = SQ::template R<K - 1>;
    constexpr int LSL = SQ::template LS<K - 1>;
    Pass<LY, RL, ND, NT, TB> pl;
    pl.load([&](int t, int i) { return lds[LY::addr(t, i)]; });
    wg_barrier();  // the power map overwrites the tile
    pl.template compute<LSL, false>(a.tw + SQ::template TWOFF<K - 1>);
    float *pwl = reinterpret_cast<float *>(lds);
    pl.template store<LSL>([&](int t, int k, float2 x) { pwl[k * TB + t] = cabs2(x); });
  }
  wg_barrier();

  // ---- 3. CFAR over the inner columns
  constexpr int GR = TB / 4;                  // column quads per row
  constexpr int NB = NT / GR;                 // row bands
  constexpr int RB = ND >= NB ? ND / NB : 1;  // rows per band
  constexpr int RING = 2 * HY + 1;
  constexpr int NQ = 1 + 2 * HXA / 4;         // quads read per row
  constexpr int NV = 4 * NQ;
  const float *pwl = reinterpret_cast<const float *>(lds);
  const int grp = threadIdx.x % GR, band = threadIdx.x / GR;
  const int tc0 = 4 * grp;  // first tile column of the quad
  const int k0 = band * RB;
  if (k0 >= ND || tc0 < HXA || tc0 >= HXA + OUT) return;  // no barrier follows
  const int gr0 = r0 + tc0;  // its range bin (a multiple of 4, >= 0)
  if (gr0 >= a.R) return;
  const int R = a.R;

  int co[4], chh[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int gr = gr0 + j;
    co[j] = min(gr + HX, R - 1) - max(gr - HX, 0) + 1;
    chh[j] = min(gr + EX, R - 1) - max(gr - EX, 0) + 1;
  }
  const bool quad_in = gr0 + 3 < R;
  int *const dets_c = a.dets + (int64_t)c * a.ds_c + gr0;

  float full[RING][4], lr[RING][4], ctr[RING][4];
#pragma unroll
  for (int ly = 0; ly < RB + 2 * HY; ++ly) {
    const int k = k0 - HY + ly;
    float u[NV];
    if (k >= 0 && k < ND) {
      const float *row = pwl + k * TB + tc0 - HXA;
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        const float4 z = *reinterpret_cast<const float4 *>(row + 4 * q);
        u[4 * q] = z.x;
        u[4 * q + 1] = z.y;
        u[4 * q + 2] = z.z;
        u[4 * q + 3] = z.w;
      }
    } else {
#pragma unroll
      for (int i = 0; i < NV; ++i) u[i] = 0.f;
    }
    // u[HXA + j + d] is tile column tc0 + j + d
    float L[4], Mh[4], Rt[4];
    dc_win4<HX - EX>(u + HXA - HX, L);
    dc_win4<2 * EX + 1>(u + HXA - EX, Mh);
    dc_win4<HX - EX>(u + HXA + EX + 1, Rt);
    const int s = ly % RING;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      lr[s][j] = L[j] + Rt[j];
      full[s][j] = lr[s][j] + Mh[j];
      ctr[s][j] = u[HXA + j];
    }
    if (ly < 2 * HY) continue;
    const int oy = ly - HY;  // ring coordinate of the cells under test
    const int kc = k0 + ly - 2 * HY;
    if (kc >= ND) break;
    const int ro = min(kc + HY, ND - 1) - max(kc - HY, 0) + 1;
    const int rh = min(kc + EY, ND - 1) - max(kc - EY, 0) + 1;
    int det[4];
    float bav[4], pv[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int nrm = ro * co[j] - rh * chh[j];
      const float thr = thr_tab[min(max(nrm, 0), NCELL)];
      float sm = 0.f;
#pragma unroll
      for (int dy = -HY; dy <= HY; ++dy) {
        const int kk = (oy + dy) % RING;
        const float t = (dy >= -EY && dy <= EY) ? lr[kk][j] : full[kk][j];
        sm = dy == -HY ? t : sm + t;
      }
      const float p = ctr[oy % RING][j];
      pv[j] = p;
      if constexpr (EXACT) {
        bav[j] = sm / (float)nrm;
        det[j] = p > thr * bav[j] ? 1 : 0;
      } else {
        det[j] = p > thr * sm ? 1 : 0;
      }
    }
    int *drow = dets_c + (int64_t)kc * a.ds_k;
    float *brow = EXACT ? a.ba + (int64_t)c * a.bs_c + (int64_t)kc * a.bs_k + gr0 : nullptr;
    float *prow = a.pw ? a.pw + (int64_t)c * a.ps_c + (int64_t)kc * a.ps_k + gr0 : nullptr;
    if (a.vec_out && quad_in) {
      *reinterpret_cast<int4 *>(drow) = make_int4(det[0], det[1], det[2], det[3]);
      if constexpr (EXACT) *reinterpret_cast<float4 *>(brow) = make_float4(bav[0], bav[1], bav[2], bav[3]);
      if (prow) *reinterpret_cast<float4 *>(prow) = make_float4(pv[0], pv[1], pv[2], pv[3]);
    } else {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        if (gr0 + j >= R) break;
        drow[j] = det[j];
        if constexpr (EXACT) brow[j] = bav[j];
        if (prow) prow[j] = pv[j];
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
  const bool fusable = nd <= 128 && nd >= 8 && unit_rows(y) && unit_rows(dets) &&
                       unit_rows(ba_out) && unit_rows(pow_out);
  DcArgs a{};
  a.y = (const float2 *)y->data;
  a.ys_c = y->stride[0];
  a.ys_p = y->stride[1];
  a.dets = (int *)dets->data;
  a.ds_c = dets->stride[0];
  a.ds_k = dets->stride[1];
  if (ba_out) {
    a.ba = (float *)ba_out->data;
    a.bs_c = ba_out->stride[0];
    a.bs_k = ba_out->stride[1];
  }
  if (pow_out) {
    a.pw = (float *)pow_out->data;
    a.ps_c = pow_out->stride[0];
    a.ps_k = pow_out->stride[1];
  }
  a.C = (int)C;
  a.P = (int)P;
  a.R = (int)R;
  a.rows_t = (int)rows_t;
  a.p_off = tpc_mode == MXR_CONV_VALID ? 1 : 0;
  a.no_tpc = tpc_mode == MXR_TPC_NONE;
  a.win_rows = (int)win_rows;
  a.pfa = pp->pfa;
  auto al16 = [](const mxr_tensor *t) {
    return !t || ((uintptr_t)t->data % 16 == 0 && t->stride[1] % 4 == 0 && t->stride[0] % 4 == 0);
  };
  a.vec_out = al16(dets) && al16(ba_out) && al16(pow_out);
  if (fusable) {
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
