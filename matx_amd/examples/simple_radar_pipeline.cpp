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
    RadarPipeline p(P, S, L, C, stream);
    p.sync();
    HIPExecutor ex(stream);
    auto run = [&] {
      p.PulseCompression();
      p.ThreePulseCanceller();
      p.DopplerProcessing();
      p.CFARDetections();
    };
    run();
    ex.start_timer();
    for (int i = 0; i < iterations; ++i) run();
    ex.stop_timer();
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
