// Sidelobe loss heads of waveform/window design, forward and adjoint (no
// MatX counterpart: MatX has no autodiff).  They are the two scalar
// objectives a design loop puts on top of the differentiable chain:
//
//   ISL_b = (sum_l |r_b[l]|^2 - |r_b[0]|^2) / |r_b[0]|^2     r: [B, n] c64
//           (integrated sidelobe level of a matched-filter output)
//   PSL_t = (sum_{k not in ML_t} pw[c_t, k, r_t]^q)^(1/q) / pw[c_t, kd_t, r_t]
//           ML_t = Doppler bins within ml of kd_t (circular); pw: [C, nd, R]
//           (q-norm peak-sidelobe ratio of a target's Doppler column)
//
// Each is one small kernel forward and one (ISL) or two (PSL: clear, then
// scatter into the columns) backward, in place of the dozens of elementwise
// launches the same expressions cost as separate tensor operations.  Sums
// run in double; q-th powers are formed in double (a q = 8 power of a strong
// sidelobe overflows float).  Gradients follow g = dL/dRe + i dL/dIm.
#include "mxr_cache.hpp"
#include "mxr_internal.hpp"

namespace mxr {

constexpr int LH_NT = 256;

template <class T>
__device__ __forceinline__ T block_sum(T v, T *red) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  __syncthreads();  // red is free (previous use finished)
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  T s = 0;
#pragma unroll
  for (int w = 0; w < LH_NT / 64; ++w) s += red[w];
  return s;
}

struct IslArgs {
  const float2 *r;
  int64_t rs_b;
  int n;
  float *isl;  // forward output [B]
  int64_t is_b;
  const float *gisl;  // adjoint input [B]
  int64_t gs_b;
  float2 *gr;  // adjoint output [B, n]
  int64_t grs_b;
};

// One workgroup per row: S = sum |r|^2 (double), P0 = |r[0]|^2.
__device__ __forceinline__ void isl_row_sums(const float2 *rb, int n, double *red, double *S,
                                             double *P0) {
  double s = 0.0;
  for (int l = threadIdx.x; l < n; l += LH_NT) {
    const float2 v = rb[l];
    s += (double)v.x * v.x + (double)v.y * v.y;
  }
  *S = block_sum(s, red);
  const float2 v0 = rb[0];
  *P0 = (double)v0.x * v0.x + (double)v0.y * v0.y;
}

__global__ __launch_bounds__(LH_NT) void isl_kernel(IslArgs a) {
  __shared__ double red[LH_NT / 64];
  const float2 *rb = a.r + (int64_t)blockIdx.x * a.rs_b;
  double S, P0;
  isl_row_sums(rb, a.n, red, &S, &P0);
  if (threadIdx.x == 0) a.isl[(int64_t)blockIdx.x * a.is_b] = (float)((S - P0) / P0);
}

// d ISL / d|r_l|^2 = 1/P0 (l != 0), -(S - P0)/P0^2 (l = 0); |r|^2 -> 2 r.
// ISL_SPLIT workgroups per row: each re-forms the row sums (the row is 8n
// bytes, served by L2) and writes its share of the row.
constexpr int ISL_SPLIT = 8;
__global__ __launch_bounds__(LH_NT) void isl_adj_kernel(IslArgs a) {
  __shared__ double red[LH_NT / 64];
  const int b = blockIdx.x / ISL_SPLIT, part = blockIdx.x % ISL_SPLIT;
  const float2 *rb = a.r + (int64_t)b * a.rs_b;
  double S, P0;
  isl_row_sums(rb, a.n, red, &S, &P0);
  const double g = (double)a.gisl[(int64_t)b * a.gs_b];
  const float side = (float)(2.0 * g / P0);
  const float lag0 = (float)(-2.0 * g * (S - P0) / (P0 * P0));
  float2 *grb = a.gr + (int64_t)b * a.grs_b;
  const int chunk = (a.n + ISL_SPLIT - 1) / ISL_SPLIT;
  const int l1 = min(a.n, (part + 1) * chunk);
  for (int l = part * chunk + threadIdx.x; l < l1; l += LH_NT)
    grb[l] = cscale(rb[l], l == 0 ? lag0 : side);
}

struct PslArgs {
  const float *pw;
  int64_t ps_c, ps_k, ps_r;
  const int *tg;  // [T, 3]: channel, range bin, Doppler bin
  int64_t ts_t, ts_e;
  int T, nd, ml, q;
  int64_t C, R;  // map extent: a target outside it gives NaN and no gradient
  float *psl;  // forward output [T]
  int64_t ls_t;
  const float *gpsl;  // adjoint input [T]
  int64_t gls_t;
  float *gp;  // adjoint output [C, nd, R]
  int64_t gs_c, gs_k, gs_r;
  int64_t numel;  // of gp (contiguous when cleared as one span)
};

__device__ __forceinline__ bool in_mainlobe(int k, int kd, int nd, int ml) {
  int d = k - kd;
  d = ((d % nd) + nd) % nd;
  return d <= ml || nd - d <= ml;
}

__device__ __forceinline__ double powq(double x, int q) {
  double r = 1.0;
  for (int i = 0; i < q; ++i) r *= x;
  return r;
}

// A_t = (sum_{k not in ML} pw_k^q)^(1/q), M_t = pw_kd.  Target indices are
// device data: one outside the map is reported invalid (no access).  Every
// thread of the workgroup calls this (it holds a barrier).
__device__ __forceinline__ bool psl_terms(const PslArgs &a, int t, double *red, const float **col,
                                          int *kd, double *A, double *M) {
  const int c = a.tg[(int64_t)t * a.ts_t], r = a.tg[(int64_t)t * a.ts_t + a.ts_e];
  *kd = a.tg[(int64_t)t * a.ts_t + 2 * a.ts_e];
  const bool ok = c >= 0 && c < a.C && r >= 0 && r < a.R && *kd >= 0 && *kd < a.nd;
  *col = ok ? a.pw + (int64_t)c * a.ps_c + (int64_t)r * a.ps_r : a.pw;
  double s = 0.0;
  if (ok)
    for (int k = threadIdx.x; k < a.nd; k += LH_NT)
      if (!in_mainlobe(k, *kd, a.nd, a.ml)) s += powq((double)(*col)[(int64_t)k * a.ps_k], a.q);
  s = block_sum(s, red);
  *A = pow(s, 1.0 / a.q);
  *M = ok ? (double)(*col)[(int64_t)(*kd) * a.ps_k] : 0.0;
  return ok;
}

__global__ __launch_bounds__(LH_NT) void psl_kernel(PslArgs a) {
  __shared__ double red[LH_NT / 64];
  const float *col;
  int kd;
  double A, M;
  const bool ok = psl_terms(a, blockIdx.x, red, &col, &kd, &A, &M);
  if (threadIdx.x == 0) a.psl[(int64_t)blockIdx.x * a.ls_t] = ok ? (float)(A / M) : __int_as_float(0x7fc00000);
}

// gp = 0 over the whole map (the loss touches only the target columns).
__global__ __launch_bounds__(LH_NT) void psl_clear_kernel(float4 *gp4, int64_t n4) {
  for (int64_t i = (int64_t)blockIdx.x * LH_NT + threadIdx.x; i < n4; i += (int64_t)gridDim.x * LH_NT)
    gp4[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// d PSL_t / d pw_k = A^(1-q) pw_k^(q-1) / M (sidelobes), -A/M^2 (k = kd).
// One workgroup per target; targets that share a column are summed, in
// target order, by the workgroup of the first of them (the others exit), so
// every column is written by one workgroup and the result is deterministic.
__global__ __launch_bounds__(LH_NT) void psl_adj_kernel(PslArgs a) {
  __shared__ double red[LH_NT / 64];
  const int t0 = blockIdx.x;
  const int c0 = a.tg[(int64_t)t0 * a.ts_t], r0 = a.tg[(int64_t)t0 * a.ts_t + a.ts_e];
  for (int t = 0; t < t0; ++t)
    if (a.tg[(int64_t)t * a.ts_t] == c0 && a.tg[(int64_t)t * a.ts_t + a.ts_e] == r0) return;
  constexpr int KPT = 4;  // Doppler rows per thread (nd <= KPT * LH_NT)
  double acc[KPT] = {0.0, 0.0, 0.0, 0.0};
  bool any = false;
  for (int t = t0; t < a.T; ++t) {
    if (a.tg[(int64_t)t * a.ts_t] != c0 || a.tg[(int64_t)t * a.ts_t + a.ts_e] != r0) continue;
    const float *col;
    int kd;
    double A, M;
    if (!psl_terms(a, t, red, &col, &kd, &A, &M)) continue;
    any = true;
    const double g = (double)a.gpsl[(int64_t)t * a.gls_t];
    const double sc = g * pow(A, 1.0 - a.q) / M;
#pragma unroll
    for (int j = 0; j < KPT; ++j) {
      const int k = threadIdx.x + j * LH_NT;
      if (k >= a.nd) break;
      if (k == kd) acc[j] += -g * A / (M * M);
      else if (!in_mainlobe(k, kd, a.nd, a.ml)) acc[j] += sc * powq((double)col[(int64_t)k * a.ps_k], a.q - 1);
    }
  }
  if (!any) return;
  float *gcol = a.gp + (int64_t)c0 * a.gs_c + (int64_t)r0 * a.gs_r;
#pragma unroll
  for (int j = 0; j < KPT; ++j) {
    const int k = threadIdx.x + j * LH_NT;
    if (k < a.nd) gcol[(int64_t)k * a.gs_k] = (float)acc[j];
  }
}

// The taps gradient of sum_t gpsl_t PSL_t over pw = |FFT_nd(taps .* x)|^2
// along pulses, taken straight from the target columns (the loss touches
// nothing else): one workgroup per target forms the column's d PSL/d pw
// (the rule of psl_adj_kernel, from the saved power map), recomputes the
// column's windowed spectrum D_k = sum_p taps_p x_p e^(-2 pi i k p / nd),
// forms gu_p = sum_k 2 gp_k D_k e^(2 pi i k p / nd) (the pulse-axis adjoint
// of |.|^2 and the transform) and writes the target's share
// Re(conj(gu_p) x_p) of d/dtaps_p; a second kernel sums the shares in target
// order.  Equal to psl_adjoint followed by doppler_taps_adjoint, without
// the map-sized gradient.  The column transforms are direct sums over an
// LDS table of the nd roots of unity (P nd complex products per target).
struct PslTapsArgs {
  const float2 *x;
  int64_t xs_c, xs_p, xs_r;
  const float *taps;
  int64_t ts;
  int P;
  float *part;  // [T, P] shares
  float *gt;    // [P]
  int64_t gts;
};

__global__ __launch_bounds__(LH_NT) void psl_taps_adj_kernel(PslArgs a, PslTapsArgs b) {
  __shared__ double red[LH_NT / 64];
  __shared__ float2 W[4 * LH_NT];   // e^(-2 pi i m / nd)
  __shared__ float2 hx[4 * LH_NT];  // taps_p x_p of the column
  __shared__ float2 D[4 * LH_NT];
  __shared__ float gp[4 * LH_NT];
  const int t = blockIdx.x;
  const float *col;
  int kd;
  double A, M;
  const bool ok = psl_terms(a, t, red, &col, &kd, &A, &M);
  float *share = b.part + (int64_t)t * b.P;
  if (!ok) {  // NaN target: no gradient (uniform over the workgroup)
    for (int p = threadIdx.x; p < b.P; p += LH_NT) share[p] = 0.f;
    return;
  }
  const int c = a.tg[(int64_t)t * a.ts_t], r = a.tg[(int64_t)t * a.ts_t + a.ts_e];
  const float2 *xc = b.x + (int64_t)c * b.xs_c + (int64_t)r * b.xs_r;
  const double g = (double)a.gpsl[(int64_t)t * a.gls_t];
  const double sc = g * pow(A, 1.0 - a.q) / M;
  const int nd = a.nd, P = b.P;
  for (int k = threadIdx.x; k < nd; k += LH_NT) {
    double v = 0.0;
    if (k == kd) v = -g * A / (M * M);
    else if (!in_mainlobe(k, kd, nd, a.ml)) v = sc * powq((double)col[(int64_t)k * a.ps_k], a.q - 1);
    gp[k] = (float)v;
    float sn, cs;
    sincospif(2.0f * (float)k / (float)nd, &sn, &cs);
    W[k] = make_float2(cs, -sn);
  }
  for (int p = threadIdx.x; p < P; p += LH_NT) hx[p] = cscale(xc[(int64_t)p * b.xs_p], b.taps[(int64_t)p * b.ts]);
  __syncthreads();
  // D_k = sum_p hx_p W^(kp)
  for (int k = threadIdx.x; k < nd; k += LH_NT) {
    float2 d = make_float2(0.f, 0.f);
    int m = 0;  // k p mod nd
    for (int p = 0; p < P; ++p) {
      d = cadd(d, cmul(hx[p], W[m]));
      m += k;
      if (m >= nd) m -= nd;
    }
    D[k] = d;
  }
  __syncthreads();
  // gu_p = sum_k 2 gp_k D_k conj(W^(kp)); the share Re(conj(gu_p) x_p)
  for (int p = threadIdx.x; p < P; p += LH_NT) {
    float2 gu = make_float2(0.f, 0.f);
    int m = 0;  // k p mod nd
    for (int k = 0; k < nd; ++k) {
      gu = cadd(gu, cscale(cmulc(D[k], W[m]), 2.0f * gp[k]));
      m += p;
      if (m >= nd) m -= nd;
    }
    const float2 xv = xc[(int64_t)p * b.xs_p];
    share[p] = gu.x * xv.x + gu.y * xv.y;
  }
}

__global__ __launch_bounds__(LH_NT) void psl_taps_sum_kernel(PslTapsArgs b, int T) {
  const int p = blockIdx.x * LH_NT + threadIdx.x;
  if (p >= b.P) return;
  float acc = 0.f;
  for (int t = 0; t < T; ++t) acc += b.part[(int64_t)t * b.P + p];
  b.gt[(int64_t)p * b.gts] = acc;
}

static mxr_status isl_args(const char *fn, const mxr_tensor *r, IslArgs *a) {
  mxr_status st;
  if ((st = check_tensor(r, "r", MXR_C64, 2, 2))) return st;
  MXR_REQUIRE(r->stride[1] == 1, MXR_INVALID_PARAMETER, "%s: r rows need unit stride", fn);
  MXR_REQUIRE(r->shape[1] >= 1 && r->shape[1] < (1ll << 31) && r->shape[0] < (1ll << 27),
              MXR_INVALID_SIZE, "%s: bad shape", fn);
  a->r = (const float2 *)r->data;
  a->rs_b = r->stride[0];
  a->n = (int)r->shape[1];
  return MXR_SUCCESS;
}

static mxr_status isl_impl(const mxr_tensor *isl, const mxr_tensor *r, hipStream_t s) {
  IslArgs a{};
  mxr_status st;
  if ((st = isl_args("isl", r, &a))) return st;
  if ((st = check_tensor(isl, "isl", MXR_F32, 1, 1))) return st;
  MXR_REQUIRE(isl->shape[0] == r->shape[0], MXR_INVALID_SIZE, "isl: output must be [%lld]",
              (long long)r->shape[0]);
  MXR_REQUIRE(!views_overlap(*isl, *r), MXR_ALIAS_ERROR, "isl: output overlaps r");
  if (r->shape[0] == 0) return MXR_SUCCESS;
  a.isl = (float *)isl->data;
  a.is_b = isl->stride[0];
  ProfScope ps("isl_kernel", s);
  isl_kernel<<<dim3((unsigned)r->shape[0]), dim3(LH_NT), 0, s>>>(a);
  MXR_HIP_TRY(hipGetLastError());
  return MXR_SUCCESS;
}

static mxr_status isl_adjoint_impl(const mxr_tensor *gr, const mxr_tensor *gisl,
                                   const mxr_tensor *r, hipStream_t s) {
  IslArgs a{};
  mxr_status st;
  if ((st = isl_args("isl_adjoint", r, &a))) return st;
  if ((st = check_tensor(gisl, "gisl", MXR_F32, 1, 1))) return st;
  if ((st = check_tensor(gr, "gr", MXR_C64, 2, 2))) return st;
  MXR_REQUIRE(gisl->shape[0] == r->shape[0] && gr->shape[0] == r->shape[0] &&
                  gr->shape[1] == r->shape[1] && gr->stride[1] == 1,
              MXR_INVALID_SIZE, "isl_adjoint: gisl [B] and gr [B, n] (unit row stride) must match r");
  MXR_REQUIRE(!views_overlap(*gr, *r) && !views_overlap(*gr, *gisl), MXR_ALIAS_ERROR,
              "isl_adjoint: gr overlaps an input");
  if (r->shape[0] == 0) return MXR_SUCCESS;
  a.gisl = (const float *)gisl->data;
  a.gs_b = gisl->stride[0];
  a.gr = (float2 *)gr->data;
  a.grs_b = gr->stride[0];
  ProfScope ps("isl_adj_kernel", s);
  isl_adj_kernel<<<dim3((unsigned)(r->shape[0] * ISL_SPLIT)), dim3(LH_NT), 0, s>>>(a);
  MXR_HIP_TRY(hipGetLastError());
  return MXR_SUCCESS;
}

static mxr_status psl_args(const char *fn, const mxr_tensor *pw, const mxr_tensor *tg, int ml,
                           int q, PslArgs *a) {
  mxr_status st;
  if ((st = check_tensor(pw, "pw", MXR_F32, 3, 3))) return st;
  if ((st = check_tensor(tg, "targets", MXR_I32, 2, 2))) return st;
  MXR_REQUIRE(tg->shape[1] == 3, MXR_INVALID_SIZE, "%s: targets must be [T, 3]", fn);
  MXR_REQUIRE(q >= 1 && q <= 64 && ml >= 0, MXR_INVALID_PARAMETER, "%s: need 1 <= q <= 64, ml >= 0", fn);
  MXR_REQUIRE(pw->shape[1] >= 1 && pw->shape[1] <= 4 * LH_NT, MXR_INVALID_SIZE,
              "%s: Doppler size %lld not in [1, %d]", fn, (long long)pw->shape[1], 4 * LH_NT);
  a->pw = (const float *)pw->data;
  a->ps_c = pw->stride[0];
  a->ps_k = pw->stride[1];
  a->ps_r = pw->stride[2];
  a->tg = (const int *)tg->data;
  a->ts_t = tg->stride[0];
  a->ts_e = tg->stride[1];
  a->T = (int)tg->shape[0];
  a->nd = (int)pw->shape[1];
  a->ml = ml;
  a->q = q;
  a->C = pw->shape[0];
  a->R = pw->shape[2];
  return MXR_SUCCESS;
}

static mxr_status psl_impl(const mxr_tensor *psl, const mxr_tensor *pw, const mxr_tensor *tg,
                           int ml, int q, hipStream_t s) {
  PslArgs a{};
  mxr_status st;
  if ((st = psl_args("psl", pw, tg, ml, q, &a))) return st;
  if ((st = check_tensor(psl, "psl", MXR_F32, 1, 1))) return st;
  MXR_REQUIRE(psl->shape[0] == a.T, MXR_INVALID_SIZE, "psl: output must be [%d]", a.T);
  MXR_REQUIRE(!views_overlap(*psl, *pw) && !views_overlap(*psl, *tg), MXR_ALIAS_ERROR,
              "psl: output overlaps an input");
  if (a.T == 0) return MXR_SUCCESS;
  a.psl = (float *)psl->data;
  a.ls_t = psl->stride[0];
  ProfScope ps("psl_kernel", s);
  psl_kernel<<<dim3((unsigned)a.T), dim3(LH_NT), 0, s>>>(a);
  MXR_HIP_TRY(hipGetLastError());
  return MXR_SUCCESS;
}

static mxr_status psl_adjoint_impl(const mxr_tensor *gp, const mxr_tensor *gpsl,
                                   const mxr_tensor *pw, const mxr_tensor *tg, int ml, int q,
                                   hipStream_t s) {
  PslArgs a{};
  mxr_status st;
  if ((st = psl_args("psl_adjoint", pw, tg, ml, q, &a))) return st;
  if ((st = check_tensor(gpsl, "gpsl", MXR_F32, 1, 1))) return st;
  if ((st = check_tensor(gp, "gpow", MXR_F32, 3, 3))) return st;
  MXR_REQUIRE(gpsl->shape[0] == a.T, MXR_INVALID_SIZE, "psl_adjoint: gpsl must be [%d]", a.T);
  for (int d = 0; d < 3; ++d)
    MXR_REQUIRE(gp->shape[d] == pw->shape[d], MXR_INVALID_SIZE, "psl_adjoint: gpow must have pw's shape");
  const int64_t n = numel(*gp);
  MXR_REQUIRE(gp->stride[2] == 1 && gp->stride[1] == gp->shape[2] &&
                  gp->stride[0] == gp->shape[1] * gp->shape[2] && n % 4 == 0 &&
                  ((uintptr_t)gp->data & 15) == 0,
              MXR_INVALID_PARAMETER, "psl_adjoint: gpow must be contiguous, 16-byte aligned, numel %% 4 == 0");
  for (const mxr_tensor *i : {pw, tg, gpsl})
    MXR_REQUIRE(!views_overlap(*gp, *i), MXR_ALIAS_ERROR, "psl_adjoint: gpow overlaps an input");
  if (n == 0) return MXR_SUCCESS;
  a.gpsl = (const float *)gpsl->data;
  a.gls_t = gpsl->stride[0];
  a.gp = (float *)gp->data;
  a.gs_c = gp->stride[0];
  a.gs_k = gp->stride[1];
  a.gs_r = gp->stride[2];
  a.numel = n;
  {
    ProfScope ps("psl_clear_kernel", s);
    const int64_t n4 = n / 4;
    int64_t blocks = (n4 + LH_NT - 1) / LH_NT;
    if (blocks > 4096) blocks = 4096;
    psl_clear_kernel<<<dim3((unsigned)blocks), dim3(LH_NT), 0, s>>>((float4 *)gp->data, n4);
  }
  MXR_HIP_TRY(hipGetLastError());
  if (a.T == 0) return MXR_SUCCESS;
  {
    ProfScope ps("psl_adj_kernel", s);
    psl_adj_kernel<<<dim3((unsigned)a.T), dim3(LH_NT), 0, s>>>(a);
  }
  MXR_HIP_TRY(hipGetLastError());
  return MXR_SUCCESS;
}

static mxr_status psl_taps_adjoint_impl(const mxr_tensor *gtaps, const mxr_tensor *gpsl,
                                        const mxr_tensor *pw, const mxr_tensor *x,
                                        const mxr_tensor *taps, const mxr_tensor *tg, int ml, int q,
                                        hipStream_t s) {
  const char *fn = "psl_taps_adjoint";
  PslArgs a{};
  PslTapsArgs b{};
  mxr_status st;
  if ((st = psl_args(fn, pw, tg, ml, q, &a))) return st;
  if ((st = check_tensor(gpsl, "gpsl", MXR_F32, 1, 1))) return st;
  if ((st = check_tensor(x, "x", MXR_C64, 3, 3))) return st;
  if ((st = check_tensor(taps, "taps", MXR_F32, 1, 1))) return st;
  if ((st = check_tensor(gtaps, "gtaps", MXR_F32, 1, 1))) return st;
  MXR_REQUIRE(gpsl->shape[0] == a.T, MXR_INVALID_SIZE, "%s: gpsl must be [%d]", fn, a.T);
  MXR_REQUIRE(x->shape[0] == pw->shape[0] && x->shape[2] == pw->shape[2] && x->shape[1] >= 1 &&
                  x->shape[1] <= pw->shape[1],
              MXR_INVALID_SIZE, "%s: x must be [C, P, R] with P <= nd for pw [C, nd, R]", fn);
  MXR_REQUIRE(taps->shape[0] == x->shape[1] && gtaps->shape[0] == x->shape[1], MXR_INVALID_SIZE,
              "%s: taps and gtaps must be [P]", fn);
  for (const mxr_tensor *i : {gpsl, pw, x, taps, tg})
    MXR_REQUIRE(!views_overlap(*gtaps, *i), MXR_ALIAS_ERROR, "%s: gtaps overlaps an input", fn);
  b.x = (const float2 *)x->data;
  b.xs_c = x->stride[0];
  b.xs_p = x->stride[1];
  b.xs_r = x->stride[2];
  b.taps = (const float *)taps->data;
  b.ts = taps->stride[0];
  b.P = (int)x->shape[1];
  b.gt = (float *)gtaps->data;
  b.gts = gtaps->stride[0];
  a.gpsl = (const float *)gpsl->data;
  a.gls_t = gpsl->stride[0];
  if (a.T == 0) {
    ProfScope ps("psl_taps_sum_kernel", s);
    psl_taps_sum_kernel<<<dim3((unsigned)((b.P + LH_NT - 1) / LH_NT)), dim3(LH_NT), 0, s>>>(b, 0);
    MXR_HIP_TRY(hipGetLastError());
    return MXR_SUCCESS;
  }
  b.part = (float *)stream_workspace("psl_taps_adjoint", sizeof(float) * (size_t)a.T * b.P, s);
  if (!b.part) return workspace_failure(fn);
  {
    ProfScope ps("psl_taps_adj_kernel", s);
    psl_taps_adj_kernel<<<dim3((unsigned)a.T), dim3(LH_NT), 0, s>>>(a, b);
  }
  MXR_HIP_TRY(hipGetLastError());
  {
    ProfScope ps("psl_taps_sum_kernel", s);
    psl_taps_sum_kernel<<<dim3((unsigned)((b.P + LH_NT - 1) / LH_NT)), dim3(LH_NT), 0, s>>>(b, a.T);
  }
  MXR_HIP_TRY(hipGetLastError());
  return MXR_SUCCESS;
}

}  // namespace mxr

extern "C" {

mxr_status mxr_isl(const mxr_tensor *isl, const mxr_tensor *r, hipStream_t stream) {
  return mxr::guarded("mxr_isl", [&] { return mxr::isl_impl(isl, r, stream); });
}

mxr_status mxr_isl_adjoint(const mxr_tensor *gr, const mxr_tensor *gisl, const mxr_tensor *r,
                           hipStream_t stream) {
  return mxr::guarded("mxr_isl_adjoint", [&] { return mxr::isl_adjoint_impl(gr, gisl, r, stream); });
}

mxr_status mxr_psl(const mxr_tensor *psl, const mxr_tensor *pw, const mxr_tensor *targets, int ml,
                   int q, hipStream_t stream) {
  return mxr::guarded("mxr_psl", [&] { return mxr::psl_impl(psl, pw, targets, ml, q, stream); });
}

mxr_status mxr_psl_adjoint(const mxr_tensor *gpow, const mxr_tensor *gpsl, const mxr_tensor *pw,
                           const mxr_tensor *targets, int ml, int q, hipStream_t stream) {
  return mxr::guarded("mxr_psl_adjoint",
                      [&] { return mxr::psl_adjoint_impl(gpow, gpsl, pw, targets, ml, q, stream); });
}

mxr_status mxr_psl_taps_adjoint(const mxr_tensor *gtaps, const mxr_tensor *gpsl, const mxr_tensor *pw,
                                const mxr_tensor *x, const mxr_tensor *taps, const mxr_tensor *targets,
                                int ml, int q, hipStream_t stream) {
  return mxr::guarded("mxr_psl_taps_adjoint", [&] {
    return mxr::psl_taps_adjoint_impl(gtaps, gpsl, pw, x, taps, targets, ml, q, stream);
  });
}

}  // extern "C"
