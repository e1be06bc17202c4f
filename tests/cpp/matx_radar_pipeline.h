// MatX's radar pipeline as MatX writes it: the calcDets user operator and the
// RadarPipeline class of examples/simple_radar_pipeline.h:50-520 (MatX
// repository), statement for statement, with three substitutions only:
// cudaExecutor -> HIPExecutor, cudaStream_t -> hipStream_t and
// cudaMemset -> hipMemset.  The doc comments are dropped.  It is caller code,
// kept here to show that MatX's own stage bodies compile and run unchanged
// against include/matx_amd.h (tests/cpp/radar_stage_test.cpp runs it and
// tests/test_cpp_frontend.py checks every stage against the FP64 oracle).
#pragma once

#include "matx_amd.h"

using namespace matx;

// examples/simple_radar_pipeline.h:50-150
template <class O, class I1, class I2, class I3, class I4>
class calcDets : public BaseOp<calcDets<O, I1, I2, I3, I4>> {

private:
  O out_;
  I1 xpow_;
  I2 ba_;
  I3 norm_;
  I4 pfa_;

public:
  using matxop = bool;
  using value_type = typename O::value_type;
  using calc_dets_xform_op = bool;

  calcDets(O out, I1 xpow, I2 ba, I3 norm, I4 pfa)
      : out_(out), xpow_(xpow), ba_(ba), norm_(norm), pfa_(pfa)
  {
  }

  template <typename CapType>
  __device__ inline void operator()(index_t idz, index_t idy, index_t idx)
  {
    if constexpr (CapType::ept == detail::ElementsPerThread::ONE) {
      typename I1::type xpow = xpow_(idz, idy, idx);
      typename I2::type ba = ba_(idz, idy, idx);
      typename I2::type norm = norm_(idz, idy, idx);
      using calc_type = typename I2::type;
      calc_type alpha = norm * (cuda::std::pow(static_cast<calc_type>(pfa_), calc_type(-1) / norm) - calc_type(1));
      out_(idz, idy, idx) = (xpow > alpha * ba) ? 1 : 0;

    }
  }

  __host__ __device__ inline index_t Size(uint32_t i) const
  {
    return out_.Size(i);
  }

  static inline constexpr __host__ __device__ int32_t Rank()
  {
    return O::Rank();
  }


  template <detail::OperatorCapability Cap, typename InType>
  __MATX_INLINE__ __MATX_HOST__ auto get_capability(const InType &in) const {
    if constexpr (Cap == detail::OperatorCapability::ELEMENTS_PER_THREAD) {
      const auto my_cap = cuda::std::array<detail::ElementsPerThread, 2>{detail::ElementsPerThread::ONE, detail::ElementsPerThread::ONE};
      auto g = combine_capabilities<Cap>(my_cap,
        detail::get_operator_capability<Cap>(xpow_, in),
        detail::get_operator_capability<Cap>(ba_, in),
        detail::get_operator_capability<Cap>(norm_, in),
        detail::get_operator_capability<Cap>(pfa_, in)
      );
      MATX_LOG_DEBUG("get_capability {} {}", static_cast<int>(g[0]), static_cast<int>(g[1]));
      return combine_capabilities<Cap>(my_cap,
        detail::get_operator_capability<Cap>(xpow_, in),
        detail::get_operator_capability<Cap>(ba_, in),
        detail::get_operator_capability<Cap>(norm_, in),
        detail::get_operator_capability<Cap>(pfa_, in)
      );
    }
    else {
      auto self_has_cap = detail::capability_attributes<Cap>::default_value;
      return detail::combine_capabilities<Cap>(self_has_cap,
        detail::get_operator_capability<Cap>(xpow_, in),
        detail::get_operator_capability<Cap>(ba_, in),
        detail::get_operator_capability<Cap>(norm_, in),
        detail::get_operator_capability<Cap>(pfa_, in)
      );
    }
  }
};

// examples/simple_radar_pipeline.h:157-520
template <typename ComplexType = cuda::std::complex<float>>
class RadarPipeline {
public:
  RadarPipeline() = delete;
  ~RadarPipeline()
  {

  }

  RadarPipeline(const index_t _numPulses, const index_t _numSamples,
                index_t _wfLen, index_t _numChannels, hipStream_t _stream)
      : numPulses(_numPulses), numSamples(_numSamples), waveformLength(_wfLen),
        numChannels(_numChannels), stream(_stream), exec(_stream)
  {
    numSamplesRnd = 1;
    while (numSamplesRnd < numSamples) {
      numSamplesRnd *= 2;
    }

    numPulsesRnd = 1;
    while (numPulsesRnd <= numPulses) {
      numPulsesRnd *= 2;
    }

    numCompressedSamples = numSamples - waveformLength + 1;

    make_tensor(waveformView, {numSamplesRnd});
    make_tensor(norms);
    make_tensor(inputView,
        {numChannels, numPulses, numSamplesRnd});
    make_tensor(tpcView,
        {numChannels, numPulsesRnd, numCompressedSamples});
    make_tensor(cancelMask, {3});
    make_tensor(normT,
        {numChannels, numPulsesRnd + cfarMaskY - 1,
         numCompressedSamples + cfarMaskX - 1});
    make_tensor(ba,
        {numChannels, numPulsesRnd + cfarMaskY - 1,
         numCompressedSamples + cfarMaskX - 1});
    make_tensor(dets,
        {numChannels, numPulsesRnd, numCompressedSamples});
    make_tensor(xPow,
        {numChannels, numPulsesRnd, numCompressedSamples});

    hipMemset(waveformView.Data(), 0, numSamplesRnd * sizeof(ComplexType));
    hipMemset(inputView.Data(), 0,
               inputView.TotalSize() * sizeof(ComplexType));
    hipMemset(tpcView.Data(), 0, tpcView.TotalSize() * sizeof(ComplexType));

    cancelMask.SetVals({1, -2, 1});

    make_tensor(cfarMaskView,
        {cfarMaskY, cfarMaskX});
    cfarMaskView.SetVals({{1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1},
                           {1, 1, 1, 1, 1, 0, 0, 0, 1, 1, 1, 1, 1},
                           {1, 1, 1, 1, 1, 0, 0, 0, 1, 1, 1, 1, 1},
                           {1, 1, 1, 1, 1, 0, 0, 0, 1, 1, 1, 1, 1},
                           {1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1}});

    cancelMask.PrefetchDevice(stream);
    ba.PrefetchDevice(stream);
    normT.PrefetchDevice(stream);
    cfarMaskView.PrefetchDevice(stream);
    dets.PrefetchDevice(stream);
    waveformView.PrefetchDevice(stream);
    norms.PrefetchDevice(stream);
    inputView.PrefetchDevice(stream);
    tpcView.PrefetchDevice(stream);
    xPow.PrefetchDevice(stream);
  }

  void sync() {
    exec.sync();
  }

  void PulseCompression()
  {
    auto waveformPart = slice(waveformView, {0}, {waveformLength});
    auto waveformT =
        waveformView.template Clone<3>({numChannels, numPulses, matxKeepDim});

    auto waveformFull = slice(waveformView, {0}, {numSamplesRnd});

    auto x = inputView;
    (waveformPart = waveformPart * hamming<0, waveformPart.Rank(), typename ComplexType::value_type>({waveformLength})).run(exec);

    (norms = sum(abs2(waveformPart))).run(exec);
    (norms = sqrt(norms)).run(exec);

    (waveformPart = waveformPart / norms).run(exec);
    (waveformFull = fft(waveformPart, numSamplesRnd)).run(exec);
    (waveformFull = conj(waveformFull)).run(exec);

    (x = fft(x)).run(exec);
    (x = x * waveformT).run(exec);
    (x = ifft(x)).run(exec);
  }

  void ThreePulseCanceller()
  {
    auto x = slice(inputView.Permute({0, 2, 1}),
        {0, 0, 0}, {numChannels, numCompressedSamples, numPulses});
    auto xo = slice(tpcView.Permute({0, 2, 1}),
        {0, 0, 0}, {numChannels, numCompressedSamples, numPulses});
    (xo = conv1d(x, cancelMask, matxConvCorrMode_t::MATX_C_MODE_SAME)).run(exec);
  }

  void DopplerProcessing()
  {
    const index_t cpulses = numPulses - (cancelMask.Size(0) - 1);

    auto xc =
        slice(tpcView, {0, 0, 0}, {numChannels, cpulses, numCompressedSamples});

    auto xf = tpcView.Permute({0, 2, 1});

    (xc = xc * hamming<1, xc.Rank(), typename ComplexType::value_type>({numChannels, numPulses - (cancelMask.Size(0) - 1),
                          numCompressedSamples}))
        .run(exec);
    (xf = fft(xf)).run(exec);
  }

  void CFARDetections()
  {
    ComputeCFARNorm();

    (xPow = abs2(tpcView)).run(exec);

    (ba = conv2d(xPow, cfarMaskView, matxConvCorrMode_t::MATX_C_MODE_FULL)).run(exec);

    auto normTrim = slice(normT, {0, cfarMaskY / 2, cfarMaskX / 2},
                                 {numChannels, numPulsesRnd + cfarMaskY / 2,
                                  numCompressedSamples + cfarMaskX / 2});

    auto baTrim = slice(ba, {0, cfarMaskY / 2, cfarMaskX / 2},
                            {numChannels, numPulsesRnd + cfarMaskY / 2,
                             numCompressedSamples + cfarMaskX / 2});
    (baTrim = baTrim / normTrim).run(exec);

#if 0
    IFELSE(xPow > normTrim*(pow(pfa, -1.0f/normTrim) - 1.0f)*baTrim,
                dets = 1, dets = 0).run(exec);
#else
    calcDets(dets, xPow, baTrim, normTrim, pfa).run(exec);
#endif
  }

  auto GetInputView() { return inputView; }

  auto GetwaveformView() { return waveformView; }

  auto GetTPCView() { return tpcView; }

  auto GetDetections() { return dets; }

  auto GetBackgroundAverages() { return ba; }

  auto GetnormT()
  {
    ComputeCFARNorm();
    return normT;
  }

private:
  void ComputeCFARNorm()
  {
    if (!cfarNormReady) {
      (normT = conv2d(ones({numChannels, numPulsesRnd, numCompressedSamples}),
             cfarMaskView, matxConvCorrMode_t::MATX_C_MODE_FULL)).run(exec);
      cfarNormReady = true;
    }
  }

  index_t numPulses;
  index_t numSamples;
  index_t waveformLength;
  index_t numSamplesRnd;
  index_t numPulsesRnd;
  index_t numCompressedSamples;
  index_t numChannels;
  const index_t cfarMaskX = 13;
  const index_t cfarMaskY = 5;

  static const constexpr float pfa = 1e-5f;

  tensor_t<typename ComplexType::value_type, 3> normT;
  tensor_t<typename ComplexType::value_type, 3> ba;
  tensor_t<int, 3> dets;
  tensor_t<typename ComplexType::value_type, 1> cancelMask;
  tensor_t<typename ComplexType::value_type, 3> xPow;
  tensor_t<ComplexType, 1> waveformView;
  tensor_t<typename ComplexType::value_type, 0> norms;
  tensor_t<ComplexType, 3> inputView;
  tensor_t<ComplexType, 3> tpcView;
  tensor_t<typename ComplexType::value_type, 2> cfarMaskView;

  hipStream_t stream;
  HIPExecutor exec;
  bool cfarNormReady = false;
};
