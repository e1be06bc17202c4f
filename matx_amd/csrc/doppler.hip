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

namespace mxr {

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
  const float2 *tw;
};

template <int ND>
struct DopCfg {
  static constexpr int NT = 256;
  // 32 range bins (256-byte rows) per tile up to 128 Doppler bins, 32 KiB of
  // LDS for 128 pulses: four workgroups per CU keep enough loads in flight
  static constexpr int TR = ND <= 128 ? 32 : 4096 / ND;
};

template <int ND>
__global__ __launch_bounds__(256) void doppler_kernel(DopArgs a, int rows_buf) {
  constexpr int NT = DopCfg<ND>::NT;
  constexpr int TR = DopCfg<ND>::TR;
  using LY = Layout<ND, TR, true>;
  using SQ = Seq<ND, false>;
  constexpr int K = SQ::n;
  extern __shared__ float2 lds[];  // rows_buf * TR
  __shared__ float win[ND];

  const int c = blockIdx.y;
  const int r0 = blockIdx.x * TR;
  for (int i = threadIdx.x; i < ND; i += NT) {
    float h = 0.f;
    if (i < a.rows_t) h = i < a.win_rows ? hamming_f32(i, a.win_rows) : 1.0f;
    win[i] = h;
  }
  // coalesced tile load: rows of TR range bins, 16 loads in flight per lane
  const float2 *yc = a.y + (int64_t)c * a.ys_c;
  {
    constexpr int RPI = NT / TR;  // pulses covered by one sweep of the workgroup
    constexpr int BATCH = 16;
    const int t = threadIdx.x % TR, p0 = threadIdx.x / TR;
    const int r = r0 + t;
    const bool rok = r < a.R;
    const float2 *col = yc + (int64_t)r * a.ys_r;
    for (int pb = 0; pb < a.P; pb += RPI * BATCH) {
      float2 v[BATCH];
#pragma unroll
      for (int u = 0; u < BATCH; ++u) {
        const int p = pb + p0 + u * RPI;
        v[u] = (rok && p < a.P) ? col[(int64_t)p * a.ys_p] : make_float2(0.f, 0.f);
      }
#pragma unroll
      for (int u = 0; u < BATCH; ++u) {
        const int p = pb + p0 + u * RPI;
        if (p < a.P) lds[p * TR + t] = v[u];
      }
    }
  }
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
static mxr_status launch_doppler(const DopArgs &a, hipStream_t s) {
  constexpr int TR = DopCfg<ND>::TR;
  const int rows_buf = a.P > ND ? a.P : ND;
  const size_t lds = (size_t)rows_buf * TR * sizeof(float2);
  MXR_REQUIRE(lds + ND * sizeof(float) <= (size_t)max_lds_per_block(), MXR_NOT_SUPPORTED,
              "doppler: %d pulses x %d range bins do not fit in LDS", rows_buf, TR);
  if (lds > 65536)
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

mxr_status doppler_impl(const mxr_tensor *pow_out, const mxr_tensor *rd_out, const mxr_tensor *y,
                        int tpc_mode, int64_t win_rows, int64_t nd, hipStream_t stream) {
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

}  // namespace mxr

extern "C" mxr_status mxr_doppler(const mxr_tensor *pow_out, const mxr_tensor *rd_out,
                                  const mxr_tensor *y, int tpc_mode, int64_t win_rows, int64_t nd,
                                  hipStream_t stream) {
  return mxr::guarded("mxr_doppler", [&] {
    return mxr::doppler_impl(pow_out, rd_out, y, tpc_mode, win_rows, nd, stream);
  });
}
