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
// Tiling: a 256-thread workgroup owns TILE_Y x TILE_X output cells; it stages
// (TILE_Y + 2HY) x (TILE_X + 2*HALO) input cells in LDS (zero outside the
// map), each thread then produces RPT rows x CPT columns.
#include <cmath>

#include "mxr_internal.hpp"

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
  int ey, ex;
  float pfa;
};

constexpr int CF_NT = 256;
constexpr int CF_CPT = 4;                   // columns per thread
constexpr int CF_TX = 64 * CF_CPT;          // tile width (one wave across)
constexpr int CF_RPT = 8;                   // rows per thread
constexpr int CF_TY = (CF_NT / 64) * CF_RPT;  // tile height

template <int HY, int HX>
__global__ __launch_bounds__(CF_NT) void cfar_kernel(CfarArgs a) {
  constexpr int HALO = (HX + 3) & ~3;  // keep the LDS row 16-byte aligned
  constexpr int LW = CF_TX + 2 * HALO;
  constexpr int LH = CF_TY + 2 * HY;
  constexpr int NCELL = (2 * HY + 1) * (2 * HX + 1);
  __shared__ __attribute__((aligned(16))) float tile[LH * LW];
  __shared__ float alpha_tab[NCELL + 1];

  const int c = blockIdx.z;
  const int y0 = blockIdx.y * CF_TY;
  const int x0 = blockIdx.x * CF_TX;

  for (int n = threadIdx.x; n <= NCELL; n += CF_NT) {
    const float nf = (float)n;
    alpha_tab[n] = nf * (powf(a.pfa, -1.0f / nf) - 1.0f);
  }
  // stage the tile (zero padding outside the map)
  for (int e = threadIdx.x; e < LH * LW; e += CF_NT) {
    const int ly = e / LW, lx = e % LW;
    const int gy = y0 - HY + ly, gx = x0 - HALO + lx;
    float v = 0.f;
    if (gy >= 0 && gy < a.Ny && gx >= 0 && gx < a.Nx) {
      const int64_t o = (int64_t)c * a.is_c + (int64_t)gy * a.is_y + (int64_t)gx * a.is_x;
      if (a.in_c64) {
        const float2 z = reinterpret_cast<const float2 *>(a.in)[o];
        v = z.x * z.x + z.y * z.y;
      } else {
        v = reinterpret_cast<const float *>(a.in)[o];
      }
    }
    tile[e] = v;
  }
  __syncthreads();

  const int tx = threadIdx.x % 64, ty = threadIdx.x / 64;
  const int cx0 = tx * CF_CPT;        // first tile column of this thread
  const int ry0 = ty * CF_RPT;        // first tile row of this thread
  const int ex = a.ex, ey = a.ey;

  // horizontal partial sums for RPT + 2HY input rows
  float full[CF_RPT + 2 * HY][CF_CPT], lr[CF_RPT + 2 * HY][CF_CPT];
#pragma unroll
  for (int rr = 0; rr < CF_RPT + 2 * HY; ++rr) {
    const float *row = tile + (ry0 + rr) * LW + HALO + cx0;
    float v[CF_CPT + 2 * HX];
#pragma unroll
    for (int i = 0; i < CF_CPT + 2 * HX; ++i) v[i] = row[i - HX];
#pragma unroll
    for (int j = 0; j < CF_CPT; ++j) {
      float l = 0.f, m = 0.f, r = 0.f;
#pragma unroll
      for (int d = -HX; d <= HX; ++d) {
        const float w = v[j + HX + d];
        if (d < -ex) l += w;
        else if (d <= ex) m += w;
        else r += w;
      }
      lr[rr][j] = l + r;
      full[rr][j] = (l + m) + r;
    }
  }

#pragma unroll
  for (int rr = 0; rr < CF_RPT; ++rr) {
    const int gy = y0 + ry0 + rr;
    if (gy >= a.Ny) break;
    const int ro = min(gy + HY, a.Ny - 1) - max(gy - HY, 0) + 1;
    const int rh = min(gy + ey, a.Ny - 1) - max(gy - ey, 0) + 1;
#pragma unroll
    for (int j = 0; j < CF_CPT; ++j) {
      const int gx = x0 + cx0 + j;
      if (gx >= a.Nx) continue;
      float s = 0.f;
#pragma unroll
      for (int dy = -HY; dy <= HY; ++dy) {
        const bool hole_row = dy >= -ey && dy <= ey;
        s += hole_row ? lr[rr + HY + dy][j] : full[rr + HY + dy][j];
      }
      const int co = min(gx + HX, a.Nx - 1) - max(gx - HX, 0) + 1;
      const int chh = min(gx + ex, a.Nx - 1) - max(gx - ex, 0) + 1;
      const int nrm = ro * co - rh * chh;
      const float normf = (float)nrm;
      const float b = s / normf;
      const float p = tile[(ry0 + rr + HY) * LW + HALO + cx0 + j];
      const int det = p > alpha_tab[nrm] * b ? 1 : 0;
      a.dets[(int64_t)c * a.ds_c + (int64_t)gy * a.ds_y + (int64_t)gx * a.ds_x] = det;
      if (a.ba) a.ba[(int64_t)c * a.bs_c + (int64_t)gy * a.bs_y + (int64_t)gx * a.bs_x] = b;
      if (a.pwo) a.pwo[(int64_t)c * a.os_c + (int64_t)gy * a.os_y + (int64_t)gx * a.os_x] = p;
    }
  }
}

template <int HY, int HX>
static mxr_status launch_cfar(const CfarArgs &a, hipStream_t s) {
  dim3 grid((a.Nx + CF_TX - 1) / CF_TX, (a.Ny + CF_TY - 1) / CF_TY, a.C);
  cfar_kernel<HY, HX><<<grid, dim3(CF_NT), 0, s>>>(a);
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
  a.ey = gy / 2;
  a.ex = gx / 2;
  a.pfa = pp->pfa;
  const int hy = wy / 2, hx = wx / 2;
#define MXR_CFAR_CASE(HY_, HX_) \
  if (hy == HY_ && hx == HX_) return launch_cfar<HY_, HX_>(a, stream);
  MXR_CFAR_CASE(2, 6)   // 5 x 13 (MatX mask)
  MXR_CFAR_CASE(1, 1)
  MXR_CFAR_CASE(1, 2)
  MXR_CFAR_CASE(2, 2)
  MXR_CFAR_CASE(2, 4)
  MXR_CFAR_CASE(3, 3)
  MXR_CFAR_CASE(3, 10)  // 7 x 21
  MXR_CFAR_CASE(10, 3)  // 21 x 7
  MXR_CFAR_CASE(4, 8)
#undef MXR_CFAR_CASE
  set_error(MXR_NOT_SUPPORTED,
            "cfar: window %dx%d has no fused kernel instance (use conv2d + threshold)", wy, wx);
  return MXR_NOT_SUPPORTED;
}

}  // namespace mxr

extern "C" mxr_status mxr_cfar(const mxr_tensor *dets, const mxr_tensor *ba_out,
                               const mxr_tensor *pow_out, const mxr_tensor *in,
                               const mxr_cfar_params *params, hipStream_t stream) {
  return mxr::guarded("mxr_cfar",
                      [&] { return mxr::cfar_impl(dets, ba_out, pow_out, in, params, stream); });
}
