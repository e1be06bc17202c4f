// Stage test of the C++ front end: runs a radar pipeline stage by stage on
// the input written by tests/test_cpp_frontend.py and dumps every stage's
// output, which the Python side compares with the FP64 oracle (the per-stage
// checks of MatX's test/01_radar/MultiChannelRadarPipeline.cu).
//   mode "matx":    MatX's own RadarPipeline and calcDets source
//                   (matx_radar_pipeline.h), unchanged but for the executor
//                   and stream types;
//   mode "example": this library's example (matx_amd/examples).
// Usage: radar_stage_test <dir> <channels> [matx|example]
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../matx_amd/examples/simple_radar_pipeline.h"
#include "matx_radar_pipeline.h"

using namespace matx;

template <typename T>
static std::vector<T> read_file(const std::string &path, size_t n) {
  std::vector<T> v(n);
  FILE *f = fopen(path.c_str(), "rb");
  if (!f || fread(v.data(), sizeof(T), n, f) != n) {
    fprintf(stderr, "cannot read %s\n", path.c_str());
    exit(2);
  }
  fclose(f);
  return v;
}

template <typename T, int R>
static void dump(const std::string &path, const tensor_t<T, R> &t, HIPExecutor &ex) {
  ex.sync();
  // the tensors here are contiguous (device or managed memory)
  std::vector<T> h((size_t)t.TotalSize());
  detail::hip_check(hipMemcpy(h.data(), t.Data(), sizeof(T) * h.size(), hipMemcpyDefault), "copy");
  FILE *f = fopen(path.c_str(), "wb");
  fwrite(h.data(), sizeof(T), h.size(), f);
  fclose(f);
}

template <typename Pipeline>
static void run_stages(Pipeline &p, const std::string &dir, index_t C, index_t P, index_t S, index_t L,
                       HIPExecutor &ex) {
  auto x = read_file<cfloat>(dir + "/x.bin", (size_t)P * S);
  auto w = read_file<cfloat>(dir + "/w.bin", (size_t)L);
  auto in = p.GetInputView();
  const index_t N = in.Size(2);
  std::vector<cfloat> h((size_t)C * P * N, cfloat(0.f, 0.f));
  for (index_t c = 0; c < C; ++c)
    for (index_t q = 0; q < P; ++q)
      for (index_t s = 0; s < S; ++s) h[(c * P + q) * N + s] = x[q * S + s];
  detail::hip_check(hipMemcpy(in.Data(), h.data(), sizeof(cfloat) * h.size(), hipMemcpyDefault), "copy");
  detail::hip_check(hipMemcpy(p.GetwaveformView().Data(), w.data(), sizeof(cfloat) * L, hipMemcpyDefault),
                    "copy");
  p.PulseCompression();
  dump(dir + "/pc.bin", p.GetInputView(), ex);
  p.ThreePulseCanceller();
  dump(dir + "/tpc.bin", p.GetTPCView(), ex);
  p.DopplerProcessing();
  dump(dir + "/dop.bin", p.GetTPCView(), ex);
  p.CFARDetections();
  dump(dir + "/dets.bin", p.GetDetections(), ex);
  dump(dir + "/ba.bin", p.GetBackgroundAverages(), ex);
  dump(dir + "/normT.bin", p.GetnormT(), ex);
}

int main(int argc, char **argv) {
  if (argc < 3) return 2;
  const std::string dir = argv[1];
  const index_t C = atoi(argv[2]), P = 128, S = 9000, L = 1000;
  const std::string mode = argc > 3 ? argv[3] : "matx";
  hipStream_t stream;
  detail::hip_check(hipStreamCreate(&stream), "hipStreamCreate");
  HIPExecutor ex(stream);
  try {
    if (mode == "matx") {
      RadarPipeline<> p(P, S, L, C, stream);
      run_stages(p, dir, C, P, S, L, ex);
    } else {
      matx_amd_examples::RadarPipeline p(P, S, L, C, stream);
      run_stages(p, dir, C, P, S, L, ex);
    }
    auto x = read_file<cfloat>(dir + "/x.bin", (size_t)P * S);
    // size checks throw before launch (FFT.cu:661-719)
    auto tc = make_tensor<cfloat, 1>({16});
    auto t2 = make_tensor<cfloat, 1>({32});
    bool threw = false;
    try {
      (t2 = fft(tc)).run(ex);
    } catch (const matxException &) {
      threw = true;
    }
    printf(threw ? "size-check ok\n" : "size-check MISSING\n");

    // transform operators on the first 4 x 100 input samples (real parts for
    // the real-input forms), dumped for the FP64 comparison
    auto xr = make_tensor<float, 2>({4, 100});
    auto xc = make_tensor<cfloat, 2>({4, 100});
    auto fh = make_tensor<cfloat, 1>({9});
    {
      std::vector<float> hr(400);
      std::vector<cfloat> hc(400);
      for (int i = 0; i < 400; ++i) {
        hr[i] = x[i].real();
        hc[i] = x[i];
      }
      std::vector<cfloat> hf(x.begin() + 400, x.begin() + 409);
      detail::hip_check(hipMemcpy(xr.Data(), hr.data(), sizeof(float) * 400, hipMemcpyDefault), "copy");
      detail::hip_check(hipMemcpy(xc.Data(), hc.data(), sizeof(cfloat) * 400, hipMemcpyDefault), "copy");
      detail::hip_check(hipMemcpy(fh.Data(), hf.data(), sizeof(cfloat) * 9, hipMemcpyDefault), "copy");
    }
    auto rf = make_tensor<cfloat, 2>({4, 51});
    auto ir = make_tensor<float, 2>({4, 100});
    auto f2 = make_tensor<cfloat, 2>({4, 100});
    auto if2 = make_tensor<cfloat, 2>({4, 100});
    auto rf2 = make_tensor<cfloat, 2>({4, 51});
    auto ir2 = make_tensor<float, 2>({4, 100});
    auto cv = make_tensor<cfloat, 2>({4, 108});
    auto cr = make_tensor<cfloat, 2>({4, 100});
    (rf = rfft(xr)).run(ex);
    (ir = irfft(rf, 100)).run(ex);
    (f2 = fft2(xc)).run(ex);
    (if2 = ifft2(f2)).run(ex);
    (rf2 = rfft2(xr)).run(ex);
    (ir2 = irfft2(rf2)).run(ex);
    (cv = conv1d(xc, fh, MATX_C_MODE_FULL, MATX_C_METHOD_FFT)).run(ex);
    (cr = corr(xc, fh, MATX_C_MODE_SAME, MATX_C_METHOD_FFT)).run(ex);
    dump(dir + "/op_rfft.bin", rf, ex);
    dump(dir + "/op_irfft.bin", ir, ex);
    dump(dir + "/op_fft2.bin", f2, ex);
    dump(dir + "/op_ifft2.bin", if2, ex);
    dump(dir + "/op_rfft2.bin", rf2, ex);
    dump(dir + "/op_irfft2.bin", ir2, ex);
    dump(dir + "/op_conv.bin", cv, ex);
    dump(dir + "/op_corr.bin", cr, ex);
    // .npy round trips: write two results; read the file the Python side wrote
    // and dump it back
    io::write_npy(rf, dir + "/op_rfft.npy");
    io::write_npy(ir, dir + "/op_irfft.npy");
    auto npin = make_tensor<cfloat, 2>({3, 7});
    io::read_npy(npin, dir + "/npy_in.npy");
    dump(dir + "/npy_in.bin", npin, ex);
    bool npy_threw = false;
    try {
      auto wrong = make_tensor<cfloat, 2>({7, 3});
      io::read_npy(wrong, dir + "/npy_in.npy");
    } catch (const matxException &) {
      npy_threw = true;
    }
    printf(npy_threw ? "npy-shape-check ok\n" : "npy-shape-check MISSING\n");
    // DLPack export of a permuted view and import back: same memory, shape
    // and strides; the imported tensor owns the export and releases it
    {
      auto t = make_tensor<float, 2>({6, 9}, MATX_DEVICE_MEMORY);
      auto v = t.Permute({1, 0});
      mxr_dl_managed *m = v.ToDlPack();
      bool ok = m->dl_tensor.data == (void *)v.Data() && m->dl_tensor.ndim == 2 &&
                m->dl_tensor.shape[0] == 9 && m->dl_tensor.shape[1] == 6 &&
                m->dl_tensor.strides[0] == 1 && m->dl_tensor.strides[1] == 9 &&
                m->dl_tensor.dtype_code == 2 && m->dl_tensor.device_type == 10;
      auto back = make_tensor_from_dlpack<float, 2>(m);
      ok = ok && back.Data() == v.Data() && back.Size(0) == 9 && back.Stride(1) == 9;
      printf(ok ? "dlpack ok\n" : "dlpack MISMATCH\n");
    }
  } catch (const matxException &e) {
    fprintf(stderr, "matx error %d: %s\n", e.code(), e.what());
    return 1;
  }
  printf("stages ok\n");
  return 0;
}
