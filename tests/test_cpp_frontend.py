"""The C++ front end (include/matx_amd.h) end to end: two radar pipelines
run stage by stage on MatX's seed-1234 test input through the compiled
tests/cpp/radar_stage_test --

  * "matx": MatX's own RadarPipeline and calcDets user operator
    (tests/cpp/matx_radar_pipeline.h: examples/simple_radar_pipeline.h:50-520
    with only the executor, stream and memset names changed);
  * "example": this library's example (matx_amd/examples/);

every stage is compared with the FP64 oracle in MatX's own dataflow (the
per-stage checks of test/01_radar/MultiChannelRadarPipeline.cu, end to
end).  A CPU test compiles a host-only program against the header with g++."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "matx_amd", "_build", "bin")


def test_cpp_programs_built():
    for name in ("simple_radar_pipeline", "radar_stage_test"):
        assert os.path.exists(os.path.join(BIN, name)), f"{name} not built"


HOST_ONLY = r"""
#include "matx_amd.h"
using namespace matx;
int main() {
  auto a = make_tensor<cfloat, 2>({4, 64});
  auto b = make_tensor<cfloat, 2>({4, 64});
  auto p = make_tensor<float, 2>({4, 64});
  tensor_t<float, 0> n;
  make_tensor(n);
  HIPExecutor ex(nullptr);
  (b = fft(a)).run(ex);
  (p = abs2(b) * hamming<1, 2, float>({4, 64})).run(ex);
  (n = sum(abs2(a))).run(ex);
  auto c = make_tensor<float, 2>({8, 76});
  (c = conv2d(ones({4, 64}), p, MATX_C_MODE_FULL)).run(ex);
  return 0;
}
"""


def test_cpp_header_host_only_gxx(tmp_path):
    """Code without user-defined operators needs no device compiler: the
    header compiles with g++ (syntax and semantics, no link)."""
    src = tmp_path / "host_only.cpp"
    src.write_text(HOST_ONLY)
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-D__HIP_PLATFORM_AMD__",
                        "-I" + os.path.join(ROOT, "include"), "-I/opt/rocm/include", str(src)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["matx", "example"])
def test_cpp_radar_stages(orc, matx_inputs, golden, tmp_path, mode):
    x, w = matx_inputs
    C, P, S, L, R, Pr = 2, 128, 9000, 1000, 8001, 256
    x.astype(np.complex64).tofile(tmp_path / "x.bin")
    npy_in = (np.arange(21) + 1j * np.arange(21)[::-1]).astype(np.complex64).reshape(3, 7)
    np.save(tmp_path / "npy_in.npy", npy_in)
    w.astype(np.complex64).tofile(tmp_path / "w.bin")
    r = subprocess.run([os.path.join(BIN, "radar_stage_test"), str(tmp_path), str(C), mode],
                       capture_output=True, text=True, timeout=180)
    assert r.returncode == 0, r.stderr
    assert "size-check ok" in r.stdout and "stages ok" in r.stdout
    assert "npy-shape-check ok" in r.stdout
    assert "dlpack ok" in r.stdout
    st = orc.radar_pipeline(x[None], w, mode="matx", stages=True)

    def rel(a, b):
        return np.linalg.norm(a - b) / np.linalg.norm(b)
    pc = np.fromfile(tmp_path / "pc.bin", np.complex64).reshape(C, P, 16384)
    tpc = np.fromfile(tmp_path / "tpc.bin", np.complex64).reshape(C, Pr, R)
    dop = np.fromfile(tmp_path / "dop.bin", np.complex64).reshape(C, Pr, R)
    dets = np.fromfile(tmp_path / "dets.bin", np.int32).reshape(C, Pr, R)
    ba = np.fromfile(tmp_path / "ba.bin", np.float32).reshape(C, Pr + 4, R + 12)
    nt = np.fromfile(tmp_path / "normT.bin", np.float32).reshape(C, Pr + 4, R + 12)
    for c in range(C):
        assert rel(pc[c, :, :R], st["y"][0]) < 1e-5
        assert rel(tpc[c, :P], st["tpc"][0]) < 1e-5
        assert rel(dop[c], st["D"][0]) < 1e-5
        assert rel(ba[c, 2:Pr + 2, 6:R + 6], st["ba"][0]) < 1e-5
        assert np.array_equal(nt[c, 2:Pr + 2, 6:R + 6], st["norm"][0])
        near = np.abs(st["margin"][0]) < 1e-5
        assert ((dets[c] != st["dets"][0]) & ~near).sum() == 0
        assert np.array_equal(np.argwhere(dets[c] == 1), golden["matx_det_idx"])
    # transform operators (rfft/irfft, fft2/ifft2, rfft2/irfft2, conv1d by the
    # FFT method, corr) on the first 4 x 100 samples, against numpy FP64
    xs = x.reshape(-1)[:400].astype(np.complex128).reshape(4, 100)
    xr, fh = xs.real.astype(np.float32).astype(np.float64), x.reshape(-1)[400:409].astype(np.complex128)

    def op(name, dt, shape):
        return np.fromfile(tmp_path / f"op_{name}.bin", dt).reshape(shape)
    assert rel(op("rfft", np.complex64, (4, 51)), np.fft.rfft(xr)) < 1e-5
    assert rel(op("irfft", np.float32, (4, 100)), xr) < 1e-5
    assert rel(op("fft2", np.complex64, (4, 100)), np.fft.fft2(xs)) < 1e-5
    assert rel(op("ifft2", np.complex64, (4, 100)), xs) < 1e-5
    assert rel(op("rfft2", np.complex64, (4, 51)), np.fft.rfft2(xr)) < 1e-5
    assert rel(op("irfft2", np.float32, (4, 100)), xr) < 1e-5
    conv = np.array([np.convolve(r, fh) for r in xs])
    assert rel(op("conv", np.complex64, (4, 108)), conv) < 1e-5
    corr = np.array([np.convolve(r, np.conj(fh[::-1]))[4:104] for r in xs])  # SAME start (9-1)/2
    assert rel(op("corr", np.complex64, (4, 100)), corr) < 1e-5
    # io::write_npy / io::read_npy: files numpy reads back exactly, and a
    # numpy-written file read by the C++ side
    assert np.array_equal(np.load(tmp_path / "op_rfft.npy"), op("rfft", np.complex64, (4, 51)))
    assert np.array_equal(np.load(tmp_path / "op_irfft.npy"), op("irfft", np.float32, (4, 100)))
    assert np.array_equal(np.fromfile(tmp_path / "npy_in.bin", np.complex64).reshape(3, 7), npy_in)
