"""Re-entrancy of the C ABI from several host threads (MatX's contract:
plan caches keyed and locked per thread and device, work enqueued on the
caller's stream, include/matx/core/cache.h:221-257).  Four threads, each on
its own stream, run the same sequence of library calls on their own inputs:
short and long FFTs, a Bluestein FFT, fft_conv, a pulse-compression plan,
the fused pulse-axis + CFAR pass, a generated elementwise expression and a
sum.  Every thread's outputs must be bit-identical to the same sequence run
on one thread (every kernel is deterministic), and the first calls of new
sizes (tables, plans, workspaces and hipRTC modules built under the locks)
happen concurrently."""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

NTHREADS = 4


@pytest.fixture(scope="module")
def mx():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import matx_amd
    matx_amd.lib()
    return matx_amd


def inputs(seed):
    g = torch.Generator(device="cpu").manual_seed(seed)
    x = torch.randn((2, 32, 3000), dtype=torch.complex64, generator=g).cuda()
    w = torch.randn((400,), dtype=torch.complex64, generator=g).cuda()
    big = torch.randn((2, 65536), dtype=torch.complex64, generator=g).cuda()
    odd = torch.randn((3, 12345), dtype=torch.complex64, generator=g).cuda()
    return x, w, big, odd


def sequence(mx, seed, stream):
    """The calls one thread makes; returns every output (on the stream)."""
    from matx_amd import expr as E
    x, w, big, odd = inputs(seed)
    out = {}
    with torch.cuda.stream(stream):
        out["fft"] = mx.fft(x, 4096)
        out["ifft_axis1"] = mx.ifft(x, axis=1)
        out["fft_long"] = mx.fft(big)
        out["fft_bluestein"] = mx.fft(odd)
        out["fft_conv"] = mx.fft_conv(x, w, 4096)
        plan = mx.ops.PulseCompressionPlan(3000, 400, 4096)
        plan.set_waveform(w)
        y = plan(x)
        out["pc"] = y
        dets, _, pw = mx.doppler_cfar(y, 32, tpc="valid", want_pow=True)
        out["dets"], out["pw"] = dets, pw
        t = torch.empty((2, 30, y.shape[2]), dtype=torch.float32, device="cuda")
        E.run(t, E.abs2(E.wrap(y[:, :30, :])) * E.hamming(1, 30) + 0.5)
        out["ew"] = t
        out["sum"] = E.sum(big)
    stream.synchronize()
    plan.close()
    return {k: v.clone() for k, v in out.items()}


def test_threads_bit_identical_to_single_thread(mx):
    torch.cuda.synchronize()
    ref = [sequence(mx, 100 + i, torch.cuda.Stream()) for i in range(NTHREADS)]
    mx.lib().mxr_clear_caches()   # the threads build tables, plans and modules afresh
    results, errors = [None] * NTHREADS, []
    barrier = threading.Barrier(NTHREADS)

    def worker(i):
        try:
            s = torch.cuda.Stream()
            barrier.wait()
            results[i] = sequence(mx, 100 + i, s)
        except Exception as e:  # reported below with the thread's index
            errors.append((i, repr(e)))

    th = [threading.Thread(target=worker, args=(i,)) for i in range(NTHREADS)]
    for t in th:
        t.start()
    for t in th:
        t.join(timeout=120)
    assert not any(t.is_alive() for t in th), "a worker thread did not finish"
    assert not errors, errors
    torch.cuda.synchronize()
    for i in range(NTHREADS):
        for k, v in ref[i].items():
            assert torch.equal(results[i][k], v), (i, k)
