// Cell-averaging CFAR: rectangular training window minus a centred guard
// hole, threshold from Pfa (replaces conv2d(xPow, mask, FULL) + trim +
// baTrim/normTrim + calcDets, examples/simple_radar_pipeline.h:394-436,
// 485-492, op body :86-98).
//
// The masked sum is built from positive partial sums only, never as
// "outer box minus hole": a strong target inside the hole would otherwise
// cancel catastrophically in float.  Per input row the kernel forms
//   L = sum of the HX-EX cells left of the hole columns,
//   M = sum over the 2EX+1 hole columns,  Rt = the cells right of them,
// and per output cell
//   sum = sum_{|dy| <= HY, |dy| > EY} (L+M+Rt)(y+dy) + sum_{|dy| <= EY} (L+Rt)(y+dy).
// norm (the count of in-bounds training cells) is exact integer arithmetic,
// ba = sum/norm (IEEE division, as MatX divides), alpha = norm*(pfa^(-1/norm)-1)
// in float as calcDets evaluates it, det = pow > alpha*ba.
//
// Tiling: a one-wave workgroup owns 256 consecutive columns x 16 rows of one
// channel; each lane owns four adjacent columns.  Walking down the rows, a
// lane reads its columns plus the window halo straight from memory (16-byte
// accesses; the overlap between neighbouring lanes is served by the cache,
// so HBM sees each cell about once), forms the horizontal window sums of its
// four columns with shared partial sums, keeps the last 2HY+1 rows in
// registers, and writes each detection row as one 16-byte store.  No LDS
// staging and no barriers: residency is bounded by registers only.
#include <cmath>
#include <type_traits>

#include "mxr_internal.hpp"
#include "mxr_cache.hpp"

namespace mxr {

struct CfarArgs {
  const void *in;
  int in_c64;
  int64_t is_c, is_y, is_x;
  int *dets;
  int64_t ds_c, ds_y, ds_x;
  float *ba;
  int64_t bs_c, bs_y, bs_x;
  float *pwo;
  int64_t os_c, os_y, os_x;
  int C, Ny, Nx;
  float pfa;
  int vec_in;    // f32 input, unit column stride, 16-byte aligned rows
  int vec_dets;  // detections: unit column stride, 16-byte aligned rows
};

constexpr int CF_NT = 64;             // one wave per workgroup
constexpr int CF_CPL = 4;             // adjacent columns per lane
constexpr int CF_TX = CF_NT * CF_CPL;  // 256 columns per tile
constexpr int CF_TY = 16;             // output rows per tile

// Sums of the four windows u[j .. j+W-1], j = 0..3, sharing partial sums
// (positive terms only, any association is exact up to rounding).
template <int W>
__device__ __forceinline__ void win4(const float *u, float (&o)[4]) {
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
    float core = u[3];  // u[3 .. W-1] is inside all four windows
#pragma unroll
    for (int i = 4; i < W; ++i) core += u[i];
    const float d = u[1] + u[2], e = u[W] + u[W + 1];
    o[0] = (u[0] + d) + core;
    o[1] = (d + core) + u[W];
    o[2] = (u[2] + core) + e;
    o[3] = (core + e) + u[W + 2];
  }
}

// HY/HX: half window (rows/cols); EY/EX: half guard hole.  CIN: input is
// complex (|.|^2 on load).  EXACT: ba = s/norm (IEEE division, written) and
// det = pow > alpha*ba, MatX's operation order; otherwise det = pow >
// (alpha/norm)*s with alpha/norm tabulated per norm count (the same
// threshold up to a few ulp of float rounding).  WPOW: also write pow.
template <int HY, int HX, int EY, int EX, bool CIN, bool EXACT, bool WPOW>
__global__ __launch_bounds__(CF_NT) void cfar_kernel(CfarArgs a) {
  constexpr int HALO = (HX + 3) & ~3;  // whole quads of halo on each side
  constexpr int NQ = 1 + 2 * HALO / 4; // quads a lane reads per row
  constexpr int NV = 4 * NQ;
  constexpr int LH = CF_TY + 2 * HY;
  constexpr int NCELL = (2 * HY + 1) * (2 * HX + 1);
  constexpr int RING = 2 * HY + 1;
  __shared__ float thr_tab[NCELL + 1];

  const int c = blockIdx.z;
  const int y0 = blockIdx.y * CF_TY;
  const int lane = threadIdx.x;
  const int gx0 = blockIdx.x * CF_TX + CF_CPL * lane;  // first column of this lane
  const int Nx = a.Nx, Ny = a.Ny;

  for (int n = lane; n <= NCELL; n += CF_NT) {
    const float nf = (float)n;
    const float alpha = nf * (powf(a.pfa, -1.0f / nf) - 1.0f);
    thr_tab[n] = EXACT ? alpha : alpha / nf;
  }
  __syncthreads();

  using T = typename std::conditional<CIN, float2, float>::type;
  const T *const in_c = reinterpret_cast<const T *>(a.in) + (int64_t)c * a.is_c;
  const int isy = (int)a.is_y, isx = (int)a.is_x;
  // Row ly of the tile -> the NV values at columns gx0-HALO .. gx0+3+HALO,
  // zero outside the map.  Quads wholly inside a row with unit stride load
  // as one 16-byte access; the few quads crossing the map edge go per cell.
  // Neighbouring lanes' quads overlap and are merged by the cache.
  auto load_row = [&](int ly, float (&u)[NV]) {
    const int gy = y0 - HY + ly;
    if (gy < 0 || gy >= Ny) {
#pragma unroll
      for (int i = 0; i < NV; ++i) u[i] = 0.f;
      return;
    }
    const T *rp = in_c + gy * isy;
#pragma unroll
    for (int qd = 0; qd < NQ; ++qd) {
      const int x = gx0 - HALO + 4 * qd;
      if (!CIN && a.vec_in && x >= 0 && x + 3 < Nx) {
        const float4 z = *reinterpret_cast<const float4 *>(reinterpret_cast<const float *>(rp) + x);
        u[4 * qd] = z.x;
        u[4 * qd + 1] = z.y;
        u[4 * qd + 2] = z.z;
        u[4 * qd + 3] = z.w;
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int xe = x + e;
          float val = 0.f;
          if (xe >= 0 && xe < Nx) {
            const T z = rp[xe * isx];
            if constexpr (CIN) val = z.x * z.x + z.y * z.y;
            else val = z;
          }
          u[4 * qd + e] = val;
        }
      }
    }
  };

  // in-bounds cell count = ro*co - rh*chh: co/chh per column, ro/rh per row
  int co[CF_CPL], chh[CF_CPL];
  float thr_int[CF_CPL];
#pragma unroll
  for (int j = 0; j < CF_CPL; ++j) {
    const int gx = gx0 + j;
    co[j] = min(gx + HX, Nx - 1) - max(gx - HX, 0) + 1;
    chh[j] = min(gx + EX, Nx - 1) - max(gx - EX, 0) + 1;
    thr_int[j] = thr_tab[min(max((2 * HY + 1) * co[j] - (2 * EY + 1) * chh[j], 0), NCELL)];
  }
  const bool lane_live = gx0 < Nx;
  const bool vec_out = a.vec_dets && gx0 + 3 < Nx;
  const int dsy = (int)a.ds_y, dsx = (int)a.ds_x;
  int *const dets_c = a.dets + (int64_t)c * a.ds_c;

  float full[RING][CF_CPL], lr[RING][CF_CPL], ctr[RING][CF_CPL];
  float ua[NV], ub[NV];  // two rows in flight
  load_row(0, ua);
  if (LH > 1) load_row(1, ub);
#pragma unroll
  for (int ly = 0; ly < LH; ++ly) {
    float u[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) u[i] = (ly & 1) ? ub[i] : ua[i];
    if (ly + 2 < LH) {
      if (ly & 1) load_row(ly + 2, ub);
      else load_row(ly + 2, ua);
    }
    // u[HALO + j + d] is column gx0 + j + d
    float L[4], Mh[4], Rt[4];
    win4<HX - EX>(u + HALO - HX, L);
    win4<2 * EX + 1>(u + HALO - EX, Mh);
    win4<HX - EX>(u + HALO + EX + 1, Rt);
    const int k = ly % RING;
#pragma unroll
    for (int j = 0; j < CF_CPL; ++j) {
      lr[k][j] = L[j] + Rt[j];
      full[k][j] = lr[k][j] + Mh[j];
      ctr[k][j] = u[HALO + j];
    }
    if (ly < 2 * HY) continue;
    const int oy = ly - HY;  // tile row of the cells under test
    const int gy = y0 + oy - HY;
    if (gy >= Ny) break;
    if (!lane_live) continue;
    float thr[CF_CPL];
    int nrm[CF_CPL];
    const bool edge = gy < HY || gy + HY >= Ny;
    const int ro = min(gy + HY, Ny - 1) - max(gy - HY, 0) + 1;
    const int rh = min(gy + EY, Ny - 1) - max(gy - EY, 0) + 1;
#pragma unroll
    for (int j = 0; j < CF_CPL; ++j) {
      nrm[j] = ro * co[j] - rh * chh[j];
      // columns past Nx (never stored) can give a meaningless count: clamp
      // the table index
      thr[j] = edge ? thr_tab[min(max(nrm[j], 0), NCELL)] : thr_int[j];
    }
    int det[CF_CPL];
#pragma unroll
    for (int j = 0; j < CF_CPL; ++j) {
      float s = 0.f;
#pragma unroll
      for (int dy = -HY; dy <= HY; ++dy) {
        const int kk = (oy + dy) % RING;
        const float t = (dy >= -EY && dy <= EY) ? lr[kk][j] : full[kk][j];
        s = dy == -HY ? t : s + t;
      }
      const float p = ctr[oy % RING][j];
      if constexpr (EXACT) {
        const float b = s / (float)nrm[j];
        det[j] = p > thr[j] * b ? 1 : 0;
        if (gx0 + j < Nx)
          a.ba[(int64_t)c * a.bs_c + (int64_t)gy * a.bs_y + (int64_t)(gx0 + j) * a.bs_x] = b;
      } else {
        det[j] = p > thr[j] * s ? 1 : 0;
      }
      if constexpr (WPOW) {
        if (gx0 + j < Nx)
          a.pwo[(int64_t)c * a.os_c + (int64_t)gy * a.os_y + (int64_t)(gx0 + j) * a.os_x] = p;
      }
    }
    int *drow = dets_c + gy * dsy;
    if (vec_out) {
      *reinterpret_cast<int4 *>(drow + gx0) = make_int4(det[0], det[1], det[2], det[3]);
    } else {
#pragma unroll
      for (int j = 0; j < CF_CPL; ++j)
        if (gx0 + j < Nx) drow[(gx0 + j) * dsx] = det[j];
    }
  }
}

// LDS-staged form for an f32 power map with unit column stride and 16-byte
// aligned rows (the pulse-axis pass's output).  A 256-thread workgroup
// stages a tile of CL_TY + 2HY rows x 256 + 2*HALO columns with coalesced
// 16-byte loads (zero outside the map), then each wave walks one band of
// CL_RB output rows with a lane per 4 adjacent columns, reading its window
// from LDS.  The register-only form above keeps two rows of halo reads and a
// ring of centre values per lane, which at a 21-bin window fills 256 VGPRs
// (one wave per SIMD); here the live set is one LDS row and the ring of
// partial sums.
constexpr int CL_NT = 256;
constexpr int CL_TX = 256;            // output columns per tile (a quad per lane)
constexpr int CL_RB = 16;             // output rows per wave
constexpr int CL_TY = 4 * CL_RB;      // output rows per tile

template <int HY, int HX, int EY, int EX, bool EXACT>
__global__ __launch_bounds__(CL_NT) void cfar_lds_kernel(CfarArgs a) {
  constexpr int HALO = (HX + 3) & ~3;
  constexpr int PS = CL_TX + 2 * HALO;  // tile row (floats), a multiple of 4
  constexpr int ROWS = CL_TY + 2 * HY;
  constexpr int QPR = PS / 4;
  constexpr int NQ = 1 + 2 * HALO / 4;
  constexpr int NV = 4 * NQ;
  constexpr int NCELL = (2 * HY + 1) * (2 * HX + 1);
  constexpr int RING = 2 * HY + 1;
  constexpr int NLY = CL_RB + 2 * HY;  // tile rows a band walks
  __shared__ __align__(16) float tile[ROWS * PS];
  __shared__ float thr_tab[NCELL + 1];

  const int c = blockIdx.z, y0 = blockIdx.y * CL_TY, x0 = blockIdx.x * CL_TX;
  const int Nx = a.Nx, Ny = a.Ny;
  for (int n = threadIdx.x; n <= NCELL; n += CL_NT) {
    const float nf = (float)n;
    const float alpha = nf * (powf(a.pfa, -1.0f / nf) - 1.0f);
    thr_tab[n] = EXACT ? alpha : alpha / nf;
  }
  const float *in_c = reinterpret_cast<const float *>(a.in) + (int64_t)c * a.is_c;
  const int isy = (int)a.is_y;
  // every quad of the staged tile is loaded before any is stored, so all of
  // a thread's loads are in flight together (a load-store loop would wait
  // out one memory latency per quad)
  constexpr int NSQ = ROWS * QPR;                 // staged quads
  constexpr int NIT = (NSQ + CL_NT - 1) / CL_NT;  // per thread
  float4 zq[NIT];
#pragma unroll
  for (int it = 0; it < NIT; ++it) {
    const int qi = threadIdx.x + it * CL_NT;
    const int r = qi / QPR, qc = qi - r * QPR;
    const int gy = y0 - HY + r, gx = x0 - HALO + 4 * qc;
    float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
    if (qi < NSQ && gy >= 0 && gy < Ny) {
      const float *rp = in_c + gy * isy;
      if (gx >= 0 && gx + 3 < Nx) {
        z = *reinterpret_cast<const float4 *>(rp + gx);
      } else {
        z.x = (gx >= 0 && gx < Nx) ? rp[gx] : 0.f;
        z.y = (gx + 1 >= 0 && gx + 1 < Nx) ? rp[gx + 1] : 0.f;
        z.z = (gx + 2 >= 0 && gx + 2 < Nx) ? rp[gx + 2] : 0.f;
        z.w = (gx + 3 >= 0 && gx + 3 < Nx) ? rp[gx + 3] : 0.f;
      }
    }
    zq[it] = z;
  }
#pragma unroll
  for (int it = 0; it < NIT; ++it) {
    const int qi = threadIdx.x + it * CL_NT;
    const int r = qi / QPR, qc = qi - r * QPR;
    if (qi < NSQ) *reinterpret_cast<float4 *>(tile + r * PS + 4 * qc) = zq[it];
  }
  __syncthreads();

  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int gx0 = x0 + 4 * lane;
  const int k0 = wv * CL_RB;  // band's first output row, tile-relative
  if (gx0 >= Nx || y0 + k0 >= Ny) return;  // no barrier follows
  int co[4], chh[4];
  float thr_int[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int gx = gx0 + j;
    co[j] = min(gx + HX, Nx - 1) - max(gx - HX, 0) + 1;
    chh[j] = min(gx + EX, Nx - 1) - max(gx - EX, 0) + 1;
    thr_int[j] = thr_tab[min(max((2 * HY + 1) * co[j] - (2 * EY + 1) * chh[j], 0), NCELL)];
  }
  const bool vec_out = a.vec_dets && gx0 + 3 < Nx;
  const int dsy = (int)a.ds_y, dsx = (int)a.ds_x;
  int *const dets_c = a.dets + (int64_t)c * a.ds_c;
  const float *col0 = tile + 4 * lane;  // u[i] of a row is column gx0 - HALO + i

  float full[RING][4], lr[RING][4];
#pragma unroll 1
  for (int ly0 = 0; ly0 < NLY; ly0 += RING)
#pragma unroll
  for (int sl = 0; sl < RING; ++sl) {
    const int ly = ly0 + sl;
    if (ly >= NLY) break;
    const float *row = col0 + (k0 + ly) * PS;
    float u[NV];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const float4 z = *reinterpret_cast<const float4 *>(row + 4 * q);
      u[4 * q] = z.x;
      u[4 * q + 1] = z.y;
      u[4 * q + 2] = z.z;
      u[4 * q + 3] = z.w;
    }
    float L[4], Mh[4], Rt[4];
    win4<HX - EX>(u + HALO - HX, L);
    win4<2 * EX + 1>(u + HALO - EX, Mh);
    win4<HX - EX>(u + HALO + EX + 1, Rt);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      lr[sl][j] = L[j] + Rt[j];
      full[sl][j] = lr[sl][j] + Mh[j];
    }
    if (ly < 2 * HY) continue;
    const int gy = y0 + k0 + ly - 2 * HY;  // row of the cells under test
    if (gy >= Ny) break;
    float thr[4];
    int nrm[4];
    const bool edge = gy < HY || gy + HY >= Ny;
    const int ro = min(gy + HY, Ny - 1) - max(gy - HY, 0) + 1;
    const int rh = min(gy + EY, Ny - 1) - max(gy - EY, 0) + 1;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      nrm[j] = ro * co[j] - rh * chh[j];
      thr[j] = edge ? thr_tab[min(max(nrm[j], 0), NCELL)] : thr_int[j];
    }
    const float4 cut = *reinterpret_cast<const float4 *>(col0 + (k0 + ly - HY) * PS + HALO);
    const float cv[4] = {cut.x, cut.y, cut.z, cut.w};
    int det[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      float sm = 0.f;
#pragma unroll
      for (int dy = -HY; dy <= HY; ++dy) {
        const int kk = (sl - HY + dy + 2 * RING) % RING;  // ring slot of row ly - HY + dy
        const float t = (dy >= -EY && dy <= EY) ? lr[kk][j] : full[kk][j];
        sm = dy == -HY ? t : sm + t;
      }
      if constexpr (EXACT) {
        const float b = sm / (float)nrm[j];
        det[j] = cv[j] > thr[j] * b ? 1 : 0;
        if (gx0 + j < Nx)
          a.ba[(int64_t)c * a.bs_c + (int64_t)gy * a.bs_y + (int64_t)(gx0 + j) * a.bs_x] = b;
      } else {
        det[j] = cv[j] > thr[j] * sm ? 1 : 0;
      }
    }
    int *drow = dets_c + gy * dsy;
    if (vec_out) {
      *reinterpret_cast<int4 *>(drow + gx0) = make_int4(det[0], det[1], det[2], det[3]);
    } else {
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (gx0 + j < Nx) drow[(gx0 + j) * dsx] = det[j];
    }
  }
}

template <int HY, int HX, int EY, int EX, bool CIN, bool EXACT, bool WPOW>
static void launch_cfar3(const CfarArgs &a, dim3 grid, hipStream_t s) {
  cfar_kernel<HY, HX, EY, EX, CIN, EXACT, WPOW><<<grid, dim3(CF_NT), 0, s>>>(a);
}

template <int HY, int HX, int EY, int EX>
static mxr_status launch_cfar(const CfarArgs &a, hipStream_t s) {
  static_assert(HX <= 16 && HY <= 10, "window");
  dim3 grid((a.Nx + CF_TX - 1) / CF_TX, (a.Ny + CF_TY - 1) / CF_TY, a.C);
  if constexpr (HY <= 4) {
    if (!a.in_c64 && a.vec_in && a.pwo == nullptr) {
      dim3 g2((a.Nx + CL_TX - 1) / CL_TX, (a.Ny + CL_TY - 1) / CL_TY, a.C);
      ProfScope ps("cfar_lds_kernel", s);
      if (a.ba) cfar_lds_kernel<HY, HX, EY, EX, true><<<g2, dim3(CL_NT), 0, s>>>(a);
      else cfar_lds_kernel<HY, HX, EY, EX, false><<<g2, dim3(CL_NT), 0, s>>>(a);
      MXR_HIP_TRY(hipGetLastError());
      return MXR_SUCCESS;
    }
  }
  {
    ProfScope ps("cfar_kernel", s);
    const bool ex = a.ba != nullptr, wp = a.pwo != nullptr;
    if (a.in_c64) {
      if (ex) wp ? launch_cfar3<HY, HX, EY, EX, true, true, true>(a, grid, s)
              : launch_cfar3<HY, HX, EY, EX, true, true, false>(a, grid, s);
      else wp ? launch_cfar3<HY, HX, EY, EX, true, false, true>(a, grid, s)
              : launch_cfar3<HY, HX, EY, EX, true, false, false>(a, grid, s);
    } else {
      if (ex) wp ? launch_cfar3<HY, HX, EY, EX, false, true, true>(a, grid, s)
              : launch_cfar3<HY, HX, EY, EX, false, true, false>(a, grid, s);
      else wp ? launch_cfar3<HY, HX, EY, EX, false, false, true>(a, grid, s)
              : launch_cfar3<HY, HX, EY, EX, false, false, false>(a, grid, s);
    }
  }
  MXR_HIP_TRY(hipGetLastError());
  return MXR_SUCCESS;
}

mxr_status cfar_impl(const mxr_tensor *dets, const mxr_tensor *ba_out, const mxr_tensor *pow_out,
                     const mxr_tensor *in, const mxr_cfar_params *pp, hipStream_t stream) {
  mxr_status st;
  MXR_REQUIRE(pp != nullptr, MXR_INVALID_PARAMETER, "cfar: null params");
  if ((st = check_tensor(in, "cfar input", 0, 3, 3))) return st;
  MXR_REQUIRE(in->dtype == MXR_F32 || in->dtype == MXR_C64, MXR_INVALID_TYPE,
              "cfar: input must be f32 power or c64");
  if ((st = check_tensor(dets, "cfar detections", MXR_I32, 3, 3))) return st;
  if (ba_out && (st = check_tensor(ba_out, "cfar ba", MXR_F32, 3, 3))) return st;
  if (pow_out && (st = check_tensor(pow_out, "cfar power", MXR_F32, 3, 3))) return st;
  const int64_t C = in->shape[0], Ny = in->shape[1], Nx = in->shape[2];
  for (const mxr_tensor *o : {dets, ba_out, pow_out}) {
    if (!o) continue;
    MXR_REQUIRE(o->shape[0] == C && o->shape[1] == Ny && o->shape[2] == Nx, MXR_INVALID_SIZE,
                "cfar: outputs must match the input shape");
    MXR_REQUIRE(!views_overlap(*o, *in), MXR_ALIAS_ERROR, "cfar: an output aliases the input");
  }
  const int wy = pp->win_y, wx = pp->win_x, gy = pp->guard_y, gx = pp->guard_x;
  MXR_REQUIRE(wy >= 1 && wx >= 1 && (wy & 1) && (wx & 1), MXR_INVALID_PARAMETER,
              "cfar: window %dx%d must have odd positive sides", wy, wx);
  MXR_REQUIRE(gy >= 1 && gx >= 1 && (gy & 1) && (gx & 1) && gy <= wy && gx <= wx,
              MXR_INVALID_PARAMETER, "cfar: guard %dx%d must be odd and inside the window", gy, gx);
  MXR_REQUIRE(pp->pfa > 0.f && pp->pfa < 1.f, MXR_INVALID_PARAMETER, "cfar: pfa must be in (0,1)");
  if (C == 0 || Ny == 0 || Nx == 0) return MXR_SUCCESS;
  MXR_REQUIRE(C <= 65535 && Ny < (1 << 30) && Nx < (1 << 30), MXR_INVALID_SIZE,
              "cfar: shape too large");
  auto chan_span = [&](const mxr_tensor *t) {
    return std::llabs(t->stride[1]) * (Ny - 1) + std::llabs(t->stride[2]) * (Nx - 1);
  };
  MXR_REQUIRE(chan_span(in) < (1ll << 31) && chan_span(dets) < (1ll << 31), MXR_INVALID_SIZE,
              "cfar: one channel of the input or detections spans more than 2^31 elements");
  CfarArgs a{};
  a.in = in->data;
  a.in_c64 = in->dtype == MXR_C64;
  a.is_c = in->stride[0];
  a.is_y = in->stride[1];
  a.is_x = in->stride[2];
  a.dets = (int *)dets->data;
  a.ds_c = dets->stride[0];
  a.ds_y = dets->stride[1];
  a.ds_x = dets->stride[2];
  if (ba_out) {
    a.ba = (float *)ba_out->data;
    a.bs_c = ba_out->stride[0];
    a.bs_y = ba_out->stride[1];
    a.bs_x = ba_out->stride[2];
  }
  if (pow_out) {
    a.pwo = (float *)pow_out->data;
    a.os_c = pow_out->stride[0];
    a.os_y = pow_out->stride[1];
    a.os_x = pow_out->stride[2];
  }
  a.C = (int)C;
  a.Ny = (int)Ny;
  a.Nx = (int)Nx;
  a.pfa = pp->pfa;
  auto aligned16 = [](const mxr_tensor *t) {
    return ((uintptr_t)t->data % 16 == 0) && t->stride[2] == 1 && t->stride[1] % 4 == 0 &&
           t->stride[0] % 4 == 0;
  };
  a.vec_in = in->dtype == MXR_F32 && aligned16(in);
  a.vec_dets = aligned16(dets);
  const int hy = wy / 2, hx = wx / 2, ey = gy / 2, ex = gx / 2;
#define MXR_CFAR_CASE(HY_, HX_, EY_, EX_) \
  if (hy == HY_ && hx == HX_ && ey == EY_ && ex == EX_) return launch_cfar<HY_, HX_, EY_, EX_>(a, stream);
  MXR_CFAR_CASE(2, 6, 1, 1)   // 5 x 13 window, 3 x 3 hole (MatX mask)
  MXR_CFAR_CASE(2, 6, 1, 2)   // 5 x 13 window, 3 x 5 hole
  MXR_CFAR_CASE(2, 6, 0, 0)
  MXR_CFAR_CASE(1, 1, 0, 0)
  MXR_CFAR_CASE(1, 2, 0, 1)
  MXR_CFAR_CASE(2, 2, 1, 1)
  MXR_CFAR_CASE(2, 4, 1, 1)
  MXR_CFAR_CASE(3, 3, 1, 1)
  MXR_CFAR_CASE(3, 10, 1, 2)  // 7 x 21 window, 3 x 5 hole
  MXR_CFAR_CASE(3, 10, 1, 1)
  MXR_CFAR_CASE(10, 3, 2, 1)  // 21 x 7 window, 5 x 3 hole
  MXR_CFAR_CASE(4, 8, 1, 1)
#undef MXR_CFAR_CASE
  set_error(MXR_NOT_SUPPORTED,
            "cfar: window %dx%d / guard %dx%d has no fused kernel instance (use conv2d + "
            "threshold)", wy, wx, gy, gx);
  return MXR_NOT_SUPPORTED;
}

}  // namespace mxr

extern "C" mxr_status mxr_cfar(const mxr_tensor *dets, const mxr_tensor *ba_out,
                               const mxr_tensor *pow_out, const mxr_tensor *in,
                               const mxr_cfar_params *params, hipStream_t stream) {
  return mxr::guarded("mxr_cfar",
                      [&] { return mxr::cfar_impl(dets, ba_out, pow_out, in, params, stream); });
}
