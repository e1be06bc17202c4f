// Waveform gradient of pulse compression (shared by the plan entry point in
// pulse.hip and the kernels in pcgrad.hip).
#pragma once

#include "mxr_internal.hpp"

namespace mxr {

// One call: every row of x ([rows, S], row offsets bi.sa) against its lags
// gy ([rows, nout], offsets bi.sb) over the overlap-save blocks (M, V, nblk)
// of the forward plan; w ([L]) is the raw waveform the forward windowed and
// normalised, gw ([L]) receives dL/dw.
struct WgradJob {
  const float2 *x, *gy, *w;
  float2 *gw;
  int64_t wstride, gwstride;
  BatchIdx bi;
  int64_t rows, V, nblk;
  int M, n, seff, nout, L;
  int window, normalize;
};

mxr_status pc_waveform_grad_run(const WgradJob &j, hipStream_t s);

}  // namespace mxr
