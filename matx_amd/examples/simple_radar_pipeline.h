// Radar pipeline on the matx_amd C++ front end: the public interface of
// MatX's RadarPipeline (examples/simple_radar_pipeline.h:157-520 -- the
// constructor, PulseCompression, ThreePulseCanceller, DopplerProcessing,
// CFARDetections, the getters, sync) with stage bodies written against the
// operator API of include/matx_amd.h.
//
// The stages are expressed with MatX's own operators: windowing and
// normalisation as elementwise expressions plus sum, the waveform spectrum as
// fft + conj, pulse compression as fft -> multiply -> ifft, the canceller as
// conv1d over permuted views, the Doppler stage as a windowed fft along the
// pulse axis, CFAR as conv2d + divide + a threshold expression (the
// expression form of calcDets; MatX's user-operator form of it runs from
// tests/cpp/matx_radar_pipeline.h).
#pragma once

#include "matx_amd.h"

namespace matx_amd_examples {
using namespace matx;

class RadarPipeline {
 public:
  RadarPipeline(index_t P, index_t S, index_t L, index_t C, hipStream_t stream)
      : numPulses(P), numSamples(S), waveformLength(L), numChannels(C), exec(stream) {
    numSamplesRnd = 1;
    while (numSamplesRnd < numSamples) numSamplesRnd *= 2;
    numPulsesRnd = 1;
    while (numPulsesRnd <= numPulses) numPulsesRnd *= 2;  // strictly above P
    R = numSamples - waveformLength + 1;
    make_tensor(waveformView, {numSamplesRnd});
    make_tensor(norms);
    make_tensor(inputView, {C, P, numSamplesRnd});
    make_tensor(tpcView, {C, numPulsesRnd, R});
    make_tensor(cancelMask, {3});
    make_tensor(normT, {C, numPulsesRnd + mY - 1, R + mX - 1});
    make_tensor(ba, {C, numPulsesRnd + mY - 1, R + mX - 1});
    make_tensor(dets, {C, numPulsesRnd, R});
    make_tensor(xPow, {C, numPulsesRnd, R});
    make_tensor(cfarMask, {mY, mX});
    auto zero = [&](void *p, size_t n) { detail::hip_check(hipMemset(p, 0, n), "hipMemset"); };
    zero(waveformView.Data(), sizeof(cfloat) * waveformView.TotalSize());
    zero(inputView.Data(), sizeof(cfloat) * inputView.TotalSize());
    zero(tpcView.Data(), sizeof(cfloat) * tpcView.TotalSize());
    cancelMask.SetVals({1.f, -2.f, 1.f});
    // 5 (Doppler) x 13 (range) training window with a 3 x 3 hole around
    // the cell under test
    std::vector<float> m(mY * mX, 1.f);
    for (int y = 1; y < 4; ++y)
      for (int x = 5; x < 8; ++x) m[y * mX + x] = 0.f;
    detail::hip_check(hipMemcpy(cfarMask.Data(), m.data(), sizeof(float) * m.size(), hipMemcpyDefault),
                      "hipMemcpy");
  }

  void sync() { exec.sync(); }

  void PulseCompression() {
    auto wPart = slice(waveformView, {0}, {waveformLength});
    auto wFull = slice(waveformView, {0}, {numSamplesRnd});
    auto wT = waveformView.template Clone<3>({numChannels, numPulses, matxKeepDim});
    (wPart = wPart * hamming<0, 1>({waveformLength})).run(exec);
    (norms = sum(abs2(wPart))).run(exec);
    (norms = sqrt(norms)).run(exec);
    (wPart = wPart / norms).run(exec);
    (wFull = fft(wPart, numSamplesRnd)).run(exec);
    (wFull = conj(wFull)).run(exec);
    (inputView = fft(inputView)).run(exec);
    (inputView = inputView * wT).run(exec);
    (inputView = ifft(inputView)).run(exec);
  }

  void ThreePulseCanceller() {
    auto x = slice(inputView.Permute({0, 2, 1}), {0, 0, 0}, {numChannels, R, numPulses});
    auto xo = slice(tpcView.Permute({0, 2, 1}), {0, 0, 0}, {numChannels, R, numPulses});
    (xo = conv1d(x, cancelMask, MATX_C_MODE_SAME)).run(exec);
  }

  void DopplerProcessing() {
    const index_t cp = numPulses - 2;
    auto xc = slice(tpcView, {0, 0, 0}, {numChannels, cp, R});
    (xc = xc * hamming<1, 3>({numChannels, cp, R})).run(exec);
    (tpcView = fft(tpcView, std::array<int, 1>{1})).run(exec);
  }

  void CFARDetections() {
    ComputeCFARNorm();
    (xPow = abs2(tpcView)).run(exec);
    (ba = conv2d(xPow, cfarMask, MATX_C_MODE_FULL)).run(exec);
    auto normTrim = slice(normT, {0, mY / 2, mX / 2}, {numChannels, numPulsesRnd + mY / 2, R + mX / 2});
    auto baTrim = slice(ba, {0, mY / 2, mX / 2}, {numChannels, numPulsesRnd + mY / 2, R + mX / 2});
    (baTrim = baTrim / normTrim).run(exec);
    (dets = xPow > normTrim * (pow(pfa, -1.0f / normTrim) - 1.0f) * baTrim).run(exec);
  }

  tensor_t<cfloat, 3> GetInputView() { return inputView; }
  tensor_t<cfloat, 1> GetwaveformView() { return waveformView; }
  tensor_t<cfloat, 3> GetTPCView() { return tpcView; }
  tensor_t<int, 3> GetDetections() { return dets; }
  tensor_t<float, 3> GetBackgroundAverages() { return ba; }
  tensor_t<float, 3> GetnormT() {
    ComputeCFARNorm();
    return normT;
  }

 private:
  void ComputeCFARNorm() {
    if (cfarNormReady) return;
    (normT = conv2d(ones({numChannels, numPulsesRnd, R}), cfarMask, MATX_C_MODE_FULL)).run(exec);
    exec.sync();
    cfarNormReady = true;
  }

  index_t numPulses, numSamples, waveformLength, numChannels, numSamplesRnd, numPulsesRnd, R;
  static constexpr index_t mX = 13, mY = 5;
  static constexpr float pfa = 1e-5f;
  tensor_t<float, 3> normT, ba, xPow;
  tensor_t<int, 3> dets;
  tensor_t<float, 1> cancelMask;
  tensor_t<cfloat, 1> waveformView;
  tensor_t<float, 0> norms;
  tensor_t<cfloat, 3> inputView, tpcView;
  tensor_t<float, 2> cfarMask;
  HIPExecutor exec;
  bool cfarNormReady = false;
};

}  // namespace matx_amd_examples
