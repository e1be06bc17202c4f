// Driver for the radar pipeline example on the HIP executor: builds the
// pipeline for 16 channels x 128 pulses x 9000 samples with a 1000-sample
// waveform, runs the four stages `iterations` times on one stream and prints
// the rate (as one JSON line with --json).
// Usage: simple_radar_pipeline [iterations] [--json]
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "simple_radar_pipeline.h"

using namespace matx;
using matx_amd_examples::RadarPipeline;

int main(int argc, char **argv) {
  const index_t C = 16, P = 128, S = 9000, L = 1000;
  const int iterations = argc > 1 ? atoi(argv[1]) : 100;
  hipStream_t stream;
  detail::hip_check(hipStreamCreate(&stream), "hipStreamCreate");
  try {
    // user ranges around the set-up and each stage, as MatX's driver
    // (examples/simple_radar_pipeline.cu:59-97); a profiler shows them with
    // rocprofv3 --marker-trace
    MATX_NVTX_SET_LOG_LEVEL(matx_nvxtLogLevels::MATX_NVTX_LOG_ALL)
    MATX_NVTX_START_RANGE("Pipeline Initialize", matx_nvxtLogLevels::MATX_NVTX_LOG_USER, 1)
    RadarPipeline p(P, S, L, C, stream);
    p.sync();
    MATX_NVTX_END_RANGE(1)
    HIPExecutor ex(stream);
    auto run = [&] {
      MATX_NVTX_START_RANGE("PulseCompression", matx_nvxtLogLevels::MATX_NVTX_LOG_USER, 21)
      p.PulseCompression();
      MATX_NVTX_END_RANGE(21)
      MATX_NVTX_START_RANGE("ThreePulseCanceller", matx_nvxtLogLevels::MATX_NVTX_LOG_USER, 22)
      p.ThreePulseCanceller();
      MATX_NVTX_END_RANGE(22)
      MATX_NVTX_START_RANGE("DopplerProcessing", matx_nvxtLogLevels::MATX_NVTX_LOG_USER, 23)
      p.DopplerProcessing();
      MATX_NVTX_END_RANGE(23)
      MATX_NVTX_START_RANGE("CFARDetections", matx_nvxtLogLevels::MATX_NVTX_LOG_USER, 24)
      p.CFARDetections();
      MATX_NVTX_END_RANGE(24)
    };
    run();
    MATX_NVTX_START_RANGE("Pipeline Test", matx_nvxtLogLevels::MATX_NVTX_LOG_USER, 2)
    ex.start_timer();
    for (int i = 0; i < iterations; ++i) run();
    ex.stop_timer();
    MATX_NVTX_END_RANGE(2)
    const float ms = ex.get_time_ms();
    const bool json = argc > 2 && std::strcmp(argv[2], "--json") == 0;
    if (json)
      printf("{\"cubes_per_s\": %.3f, \"ms_per_cube\": %.4f, \"iterations\": %d, "
             "\"channels\": %lld, \"pulses\": %lld, \"samples\": %lld, \"waveform\": %lld}\n",
             iterations / (ms * 1e-3), ms / iterations, iterations, (long long)C, (long long)P,
             (long long)S, (long long)L);
    else
      printf("Pipeline finished in %.2f ms, rate: %.2f pulses/channel/sec, %.2f cubes/s\n", ms,
             (double)iterations * C * P / (ms * 1e-3), iterations / (ms * 1e-3));
  } catch (const matxException &e) {
    fprintf(stderr, "matx error %d: %s\n", e.code(), e.what());
    return 1;
  }
  (void)hipStreamDestroy(stream);
  return 0;
}
