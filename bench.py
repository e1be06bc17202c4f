#!/usr/bin/env python3
"""Radar-pipeline throughput on one MI355X.

Headline workload (one step = one cube): 16 channels x 128 pulses x 9000
samples, complex64, resident in HBM; 1000-sample LFM (bandwidth 0.4 fs,
Hamming applied by the matched filter); noise sigma = 1 plus 20 point targets
(seeded range, Doppler, SNR 0-20 dB).  Pulse compression (N = 16384 circular
correlation, lags 0..8000) -> three-pulse canceller [1,-2,1] (126 rows) ->
Hamming(126) + 128-point Doppler FFT -> |.|^2 -> CA-CFAR (13 range x 5
Doppler window, 2 x 1 guard cells around the cell under test, i.e. a 5 x 3
hole, Pfa 1e-5) -> int32 detections.  The waveform is prepared once
(pulse-compression plan); every cube then runs two kernels: the overlap-save
pulse-compression kernel and the fused pulse-axis + CFAR kernel.  Four
distinct cubes stay resident and are processed in turn, so no cube's samples
can be served from the 256 MB last-level cache.

The cube sits in HBM in MatX's inputView layout {16, 128, 16384} (pulses
padded to the FFT size; only the 9000 samples are read); intermediate rows
are padded to 64 elements.

Prints one JSON line: cubes/s from cubes replayed out of a captured hipGraph
(no instrumentation in that timed region), the dominant kernel's roofline
from HIP events recorded around every launch of each kernel in a second,
eager pass, the CPU baseline (MatX's host dataflow restated on torch CPU with
MKL FFTs, all host cores) with the FP64 oracle beside it, and -- with the
default --workload all -- the other terms of the metric as sub-measurements:
the ragged-batch speedup, the gradient steps/s, fft_conv us per call, the
64 x 512 x 32768 pipeline with its backward, and MatX's stage calls on the
C++ front end.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload all]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import subprocess
import sys
import time

HBM_PEAK_GBS = 8000.0
ROOT = os.path.dirname(os.path.abspath(__file__))


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1440)
    ap.add_argument("--warmup", type=int, default=40)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="wall budget of each CPU baseline leg of the headline workload; the "
                         "sub-measurements of --workload all get a third of it (0 disables them)")
    ap.add_argument("--no-profile", action="store_true",
                    help="time without per-kernel events (roofline omitted)")
    ap.add_argument("--workload", default="all",
                    choices=["all", "radar", "fftconv", "large", "grad", "ragged", "stage", "conv",
                             "ew"],
                    help="all (default): the 16x128x9000 pipeline as the headline value, with the "
                         "fft_conv, ragged-batch, gradient, 64x512x32768 and C++ stage-form "
                         "workloads as sub-measurements of the same line; radar: the "
                         "16x128x9000 pipeline alone; fftconv: fft_conv on 4x32x2048 with a "
                         "256-sample chirp at N=4096 (us per call); large: the 64x512x32768 "
                         "pipeline (8.6 GB cube) and its backward; grad: waveform/window design "
                         "loss forward + backward (grad steps/s); ragged: 256 uneven "
                         "pulse-compression + Doppler jobs, batched vs per-job launches; stage: "
                         "MatX's stage calls on the C++ front end; conv: conv1d by the FFT "
                         "method on 16 x 1048576 c64 with 209 taps and 1000-point Bluestein "
                         "FFTs; ew: fused elementwise expressions at the radar cube's size")
    ap.add_argument("--streams", type=int, default=None,
                    help="pipelines run side by side as branches of the timed graph of the "
                         "radar workloads, each on its own HIP stream (default 6 for "
                         "16x128x9000, 1 for 64x512x32768)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run 5 plain 512 MiB device copies (scale reference for the "
                         "FETCH_SIZE / WRITE_SIZE counters) and exit")
    return ap.parse_args()


class HipStream:
    """A non-blocking HIP stream created for this process (not one of
    torch's pooled streams, which share the device's hardware queues
    round-robin), usable as a torch stream."""

    def __init__(self):
        import ctypes
        import torch
        try:
            self._hip = ctypes.CDLL("libamdhip64.so")
        except OSError:
            self._hip = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so")
        h = ctypes.c_void_p()
        rc = self._hip.hipStreamCreateWithFlags(ctypes.byref(h), 1)
        if rc != 0:
            raise RuntimeError(f"hipStreamCreateWithFlags failed: {rc}")
        self._h = h
        self.stream = torch.cuda.ExternalStream(h.value)

    def close(self):
        if self._h is not None:
            self._hip.hipStreamDestroy(self._h)
            self._h = None


def synth_cube(C, P, S, L, nd, n_targets, snr_db, seed, device, bw=0.4):
    """Noise (E|n|^2 = 1) plus point-target echoes of an unwindowed LFM."""
    import torch
    g = torch.Generator(device="cpu").manual_seed(seed)
    # MatX's inputView layout {C, P, N}: pulses padded to the FFT size, only
    # the first S samples hold data (and only those are read)
    N = 1
    while N < S:
        N *= 2
    xbuf = torch.zeros((C, P, N), dtype=torch.complex64, device=device)
    x = xbuf[:, :, :S]
    x.copy_(torch.randn((C, P, S), dtype=torch.complex64, device=device,
                        generator=torch.Generator(device=device).manual_seed(seed)))
    n = torch.arange(L, dtype=torch.float64)
    ph = 2 * math.pi * (-bw / 2 * n + 0.5 * bw / L * n * n)
    lfm = torch.polar(torch.ones(L, dtype=torch.float64), ph)
    targets = []
    p = torch.arange(P, dtype=torch.float64)
    for _ in range(n_targets):
        rng_bin = int(torch.randint(0, S - L + 1, (1,), generator=g))
        dop = int(torch.randint(0, nd, (1,), generator=g))
        snr = float(torch.rand(1, generator=g)) * (snr_db[1] - snr_db[0]) + snr_db[0]
        phase = float(torch.rand(1, generator=g)) * 2 * math.pi
        amp = 10 ** (snr / 20)
        ramp = torch.polar(torch.full((P,), amp, dtype=torch.float64),
                           2 * math.pi * dop * p / nd + phase)
        echo = (ramp[:, None] * lfm[None, :]).to(torch.complex64).to(device)
        x[:, :, rng_bin:rng_bin + L] += echo[None]
        targets.append((rng_bin, dop, snr))
    return x, lfm.to(torch.complex64).to(device), targets


def cfar_mask(win, guard):
    import numpy as np
    m = np.ones(win)
    cy, cx, gy, gx = win[0] // 2, win[1] // 2, guard[0] // 2, guard[1] // 2
    m[cy - gy:cy + gy + 1, cx - gx:cx + gx + 1] = 0
    return m


def cpu_baseline_leg(x, w, pw, dets, nd, budget_s, C, mask, all_channels=True):
    """Times the FP64 oracle (C, OpenMP) on channel 0 of the cube, and checks
    the GPU's power map and detections against it: every channel of the cube
    when all_channels (one more oracle pass over the whole cube), else
    channel 0."""
    import numpy as np
    sys.path.insert(0, ROOT)
    from oracle import oracle as orc
    orc.build()
    cores = int(os.environ.get("OMP_NUM_THREADS", len(os.sched_getaffinity(0))))
    reps, t0 = 0, time.perf_counter()
    st = None
    while True:
        st = orc.radar_pipeline(x[0][None], w, Nd=nd, mode="model", mask=mask, stages=True)
        reps += 1
        if time.perf_counter() - t0 >= budget_s or reps >= 2000:
            break
    el = time.perf_counter() - t0
    cpu = {"value": reps / C / el, "unit": "cubes/s", "cores": cores, "kind": "port",
           "sample": f"FP64 oracle (oracle/mxr_oracle.c, OpenMP) on channel 0 of the "
                     f"same cube (1/{C} cube), {reps} runs in {el:.1f} s"}
    chans = C if all_channels else 1
    if chans > 1:
        st = orc.radar_pipeline(x[:chans], w, Nd=nd, mode="model", mask=mask, stages=True)
    pw_ref = st["pow"]
    pw_g = pw[:chans].astype(np.float64)
    rel = float(np.linalg.norm(pw_g - pw_ref) / np.linalg.norm(pw_ref))
    near = np.abs(st["margin"]) < 1e-5
    mism = int(((dets[:chans] != st["dets"]) & ~near).sum())
    # the CFAR stage alone: the FP64 CFAR applied to the GPU's own power map
    cf = orc.cfar(pw_g, mask)
    near_cf = np.abs(cf["margin"]) < 1e-5
    mism_cf = int(((dets[:chans] != cf["dets"]) & ~near_cf).sum())
    parity = {"channels": chans, "pow_rel_l2_vs_fp64": rel, "det_mismatch_outside_1e-5": mism,
              "det_mismatch_cfar_stage": mism_cf,
              "dets_gpu": int(dets[:chans].sum()), "dets_oracle": int(st["dets"].sum())}
    return cpu, parity


def cpu_standin_leg(x0, w, nd, win, guard, budget_s, C):
    """MatX's host-executor dataflow (HostExecutor + FFTW, which cannot be
    built here) restated stage by stage with torch on the CPU: FP32, MKL
    FFTs, all host threads.  Padded N-point FFT, multiply by the conjugate
    waveform spectrum, inverse FFT, three-pulse canceller (valid), window,
    Doppler FFT, |.|^2, window sums and cell counts as 2-D convolutions with
    the CFAR mask, average and threshold.  Channel 0 of the cube, scaled to
    cubes/s."""
    import torch
    cores = len(os.sched_getaffinity(0))
    torch.set_num_threads(cores)
    P, S = x0.shape
    L = len(w)
    R = S - L + 1
    N = 1
    while N < S:
        N *= 2
    X = torch.from_numpy(x0)
    wt = torch.from_numpy(w)
    hl = torch.hamming_window(L, periodic=False, dtype=torch.float32)
    hp = torch.hamming_window(P - 2, periodic=False, dtype=torch.float32)[:, None]
    mask = torch.tensor(cfar_mask(win, guard), dtype=torch.float32)[None, None]
    pad = (win[0] // 2, win[1] // 2)
    nrm = torch.nn.functional.conv2d(torch.ones((1, 1, nd, R)), mask, padding=pad)[0, 0]
    alpha = nrm * (1e-5 ** (-1.0 / nrm) - 1.0)

    def cube():
        w1 = wt * hl
        Wc = torch.conj(torch.fft.fft(w1 / torch.sqrt((w1.abs() ** 2).sum()), N))
        y = torch.fft.ifft(torch.fft.fft(X, N, dim=-1) * Wc, dim=-1)[:, :R]
        t = (y[:-2] - 2 * y[1:-1] + y[2:]) * hp
        D = torch.fft.fft(t, nd, dim=0)
        pw = D.real ** 2 + D.imag ** 2
        s = torch.nn.functional.conv2d(pw[None, None], mask, padding=pad)[0, 0]
        return pw > alpha * (s / nrm)

    cube()
    reps, t0 = 0, time.perf_counter()
    while True:
        cube()
        reps += 1
        if time.perf_counter() - t0 >= budget_s or reps >= 5000:
            break
    el = time.perf_counter() - t0
    return {"value": round(reps / C / el, 3), "unit": "cubes/s", "cores": cores,
            "kind": "port (stand-in for MatX HostExecutor + FFTW)",
            "sample": f"channel 0 of the cube (1/{C}), torch CPU FP32 with MKL FFTs, "
                      f"{reps} runs in {el:.1f} s"}


def load_traffic(kernel, tag):
    p = os.path.join(ROOT, "profiles", f"{kernel}_{tag}_traffic.json")
    if os.path.exists(p):
        with open(p) as f:
            return json.load(f).get("hbm_bytes_per_launch")
    return None


def profiled_kernels(fn, calls, names):
    """Runs fn() `calls` times with per-launch HIP events (recorded by the
    library on the launch stream) and returns {kernel: (launches per call,
    average us per launch)} for the names that ran."""
    import torch
    from matx_amd import _lib
    lib = _lib.lib()
    # untimed warm-up: callers often come here after seconds of CPU work
    # (the oracle), with the GPU idle in between
    for _ in range(min(calls, 10)):
        fn()
    torch.cuda.synchronize()
    _lib.check(lib.mxr_profile_reset())
    _lib.check(lib.mxr_profile_enable(1))
    # the launches queue behind a fixed-length spin kernel on the same
    # stream, so the GPU runs them back to back: on an idle GPU the events
    # around a kernel shorter than its host-side launch would also time the
    # launch (and any descheduling of the host thread)
    if hasattr(torch.cuda, "_sleep"):
        torch.cuda._sleep(5_000_000)
    for _ in range(calls):
        fn()
    torch.cuda.synchronize()
    _lib.check(lib.mxr_profile_enable(0))
    out = {}
    for name in names:
        cnt, tot = _lib.profile_query(name)
        if cnt:
            out[name] = (cnt / calls, 1000.0 * tot / cnt)
    return out


def kernel_roofline(kern, bytes_per_launch):
    """Roofline object of the kernel with the largest time per call among
    those with a byte model (HBM-bound)."""
    best = None
    for name, (per_call, avg_us) in kern.items():
        if name not in bytes_per_launch:
            continue
        if best is None or per_call * avg_us > best[1] * best[2]:
            best = (name, per_call, avg_us)
    if best is None:
        return None
    name, _, avg_us = best
    b = bytes_per_launch[name]
    ach = b / (avg_us * 1e-6) / 1e9
    return {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None, "kernel": name,
            "algorithmic_bytes_per_launch": b, "avg_us": round(avg_us, 2)}


def bench_fftconv(args):
    """fft_conv workload: x {4, 32, 2048} c64 (noise sigma = 1 plus three
    delayed copies of the waveform at 10 dB, delays uniform in [0, 1792)),
    w = 256-sample LFM (bandwidth 0.5 fs, unit amplitude, Hamming-windowed),
    out = ifft(fft(x, 4096) * conj(fft(w, 4096))) {4, 32, 4096}.  Each call
    prepares the waveform spectrum and filters the cube (two kernels); timed
    as plain back-to-back calls and as calls replayed from one hipGraph."""
    import numpy as np
    import torch
    import matx_amd
    dev = "cuda:0"
    C, P, S, L, N = 4, 32, 2048, 256, 4096
    g = torch.Generator(device="cpu").manual_seed(args.seed)
    n = torch.arange(L, dtype=torch.float64)
    bw = 0.5
    lfm = torch.polar(torch.ones(L, dtype=torch.float64),
                      2 * math.pi * (-bw / 2 * n + 0.5 * bw / L * n * n))
    ham = 0.54 - 0.46 * torch.cos(2 * math.pi * n / (L - 1))
    w64 = lfm * ham
    x64 = torch.complex(torch.randn((C, P, S), generator=g, dtype=torch.float64),
                        torch.randn((C, P, S), generator=g, dtype=torch.float64)) / math.sqrt(2)
    amp = 10 ** (10 / 20)
    for c in range(C):
        for p in range(P):
            for d in torch.randint(0, 1792, (3,), generator=g).tolist():
                x64[c, p, d:d + L] += amp * lfm
    x = x64.to(torch.complex64).to(dev)
    w = w64.to(torch.complex64).to(dev)
    out = torch.empty((C, P, N), dtype=torch.complex64, device=dev)
    steps, warm = max(args.steps, 200), max(args.warmup, 20)
    for _ in range(warm):
        matx_amd.fft_conv(x, w, N, out=out)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        matx_amd.fft_conv(x, w, N, out=out)
    e1.record()
    torch.cuda.synchronize()
    us_plain = 1000.0 * e0.elapsed_time(e1) / steps
    # the same calls captured once (CALLS per graph) and replayed
    CALLS = 50
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        for _ in range(3):
            matx_amd.fft_conv(x, w, N, out=out)
    s.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=s):
        for _ in range(CALLS):
            matx_amd.fft_conv(x, w, N, out=out)
    reps = max(1, steps // CALLS)
    graph.replay()
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        graph.replay()
    e1.record()
    torch.cuda.synchronize()
    us_graph = 1000.0 * e0.elapsed_time(e1) / (reps * CALLS)
    # parity of the replayed result against the FP64 oracle
    sys.path.insert(0, ROOT)
    from oracle import oracle as orc
    orc.build()
    xh, wh = x.cpu().numpy(), w.cpu().numpy()
    t0 = time.perf_counter()
    cpu_reps = 0
    while True:
        ref = orc.fft_conv(xh.reshape(C * P, S), wh, N).reshape(C, P, N)
        cpu_reps += 1
        if time.perf_counter() - t0 >= max(args.cpu_seconds, 0.5):
            break
    cpu_s = (time.perf_counter() - t0) / cpu_reps
    got = out.cpu().numpy()
    rel = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
    B = 8 * C * P * S + 8 * C * P * N
    best = min(us_plain, us_graph)
    # the filter kernel reads x and writes the full circular output
    kern = profiled_kernels(lambda: matx_amd.fft_conv(x, w, N, out=out), 50,
                            ["pc4096_wf_kernel", "pc4096_kernel", "pc_ols_kernel", "wfspec_kernel"])
    roof = kernel_roofline(kern, {"pc4096_wf_kernel": B, "pc4096_kernel": B, "pc_ols_kernel": B})
    return ({
        "metric": "fft_conv us per call (4x32x2048 c64, L=256, N=4096)",
        "value": round(best, 3), "unit": "us/call", "n_gpus": 1, "steps": steps,
        "warmup": warm, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "c64", "data": "synthetic (noise + 3 chirp echoes per pulse at 10 dB, seeded)",
        "config": {"workload": "fft_conv circular cross-correlation, x {4,32,2048} c64, "
                               "w 256-sample Hamming LFM (0.5 fs), N=4096, full {4,32,4096} "
                               "output; per call: one kernel, each workgroup transforming the waveform "
                               "and filtering one pulse",
                   "us_per_call_plain": round(us_plain, 3),
                   "us_per_call_graph": round(us_graph, 3),
                   "calls_per_s": round(1e6 / best, 1),
                   "algorithmic_bytes_per_call": B,
                   "hbm_gbs_at_best": round(B / (best * 1e-6) / 1e9, 1),
                   "per_kernel_us": {k: round(v[1], 2) for k, v in kern.items()},
                   "rel_l2_vs_fp64": rel},
        "roofline": roof,
        "cpu_baseline": {"value": round(cpu_s * 1e6, 1), "unit": "us/call", "cores":
                         int(os.environ.get("OMP_NUM_THREADS", len(os.sched_getaffinity(0)))),
                         "kind": "port",
                         "sample": f"FP64 oracle fft_conv on the same input, {cpu_reps} calls"},
    })


def bench_conv(args):
    """FFT-method conv1d at the size of MatX's CorrelationConvolutionLargeTest
    (test/00_transform/ConvCorr.cu:61-63: 8*122880 + 2*32768 samples, 209
    taps, FULL), batched over 16 rows of unit complex noise (sigma = 1), and
    8192 Bluestein FFTs of 1000 points.  Timed as back-to-back calls with the
    inputs resident; parity of row 0 (conv) and of 64 rows (FFT) against the
    FP64 oracle."""
    import numpy as np
    import torch
    import matx_amd
    dev = "cuda:0"
    rows, N, L = 16, 8 * 122880 + 2 * 32768, 209
    g = torch.Generator(device="cpu").manual_seed(args.seed)

    def crandn(*shape):
        return (torch.complex(torch.randn(shape, generator=g), torch.randn(shape, generator=g))
                / math.sqrt(2)).to(torch.complex64)

    x, f = crandn(rows, N).to(dev), crandn(L).to(dev)
    out = torch.empty((rows, N + L - 1), dtype=torch.complex64, device=dev)
    rows2, n2 = 8192, 1000
    x2 = crandn(rows2, n2).to(dev)
    out2 = torch.empty_like(x2)
    conv = lambda: matx_amd.conv1d(x, f, "full", method="fft", out=out)  # noqa: E731
    ftf = lambda: matx_amd.fft(x2, out=out2)  # noqa: E731

    def timed(fn, k):
        for _ in range(max(args.warmup, 3)):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(k):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return 1000.0 * e0.elapsed_time(e1) / k

    k = max(args.steps, 20)
    us_conv = timed(conv, k)
    us_fft = timed(ftf, k)
    # algorithmic bytes: read the signal, write the output (the filter is
    # negligible); the overlap-save kernel writes outputs L-1 .. N+L-2, the
    # fused head kernel the first L-1
    B_conv = 8 * rows * N + 8 * rows * (N + L - 1)
    B_ols = 8 * rows * N + 8 * rows * N
    B_fft = 2 * 8 * rows2 * n2
    kern = profiled_kernels(conv, 10, ["pc4096_kernel", "pc8192_kernel", "pc_ols_kernel",
                                       "wfspec_kernel", "fconv_kernel", "fconv_spec_kernel"])
    roof = kernel_roofline(kern, {"pc4096_kernel": B_ols, "pc8192_kernel": B_ols,
                                  "pc_ols_kernel": B_ols})
    kfft = profiled_kernels(ftf, 10, ["czt_kernel"])
    roof_fft = kernel_roofline(kfft, {"czt_kernel": B_fft})
    sys.path.insert(0, ROOT)
    from oracle import oracle as orc
    orc.build()
    xh, fh = x[0].cpu().numpy(), f.cpu().numpy()
    t0 = time.perf_counter()
    ref = orc.conv1d(xh, fh, "full")
    cpu_row_s = time.perf_counter() - t0
    rel_conv = float(np.linalg.norm(out[0].cpu().numpy() - ref) / np.linalg.norm(ref))
    x2h = x2[:64].cpu().numpy()
    ref2 = orc.fft(x2h)
    rel_fft = float(np.linalg.norm(out2[:64].cpu().numpy() - ref2) / np.linalg.norm(ref2))
    return ({
        "metric": "conv1d FFT method us per call (16 x 1048576 c64, 209 taps, FULL)",
        "value": round(us_conv, 2), "unit": "us/call", "n_gpus": 1, "steps": k,
        "warmup": max(args.warmup, 3), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "c64", "data": "synthetic (unit complex noise, seeded)",
        "config": {"workload": "conv1d(x, f, FULL, MATX_C_METHOD_FFT), x {16, 1048576} c64, "
                               "f {209} c64 (MatX ConvCorr large-test size, 16 rows)",
                   "algorithmic_bytes_per_call": B_conv,
                   "hbm_gbs": round(B_conv / (us_conv * 1e-6) / 1e9, 1),
                   "per_kernel_us": {kk: round(v[1], 2) for kk, v in kern.items()},
                   "rel_l2_vs_fp64_row0": rel_conv,
                   "bluestein_fft": {"workload": "fft of 8192 rows x 1000 points c64 "
                                                 "(2048-point chirp-z convolution)",
                                     "us_per_call": round(us_fft, 2),
                                     "hbm_gbs": round(B_fft / (us_fft * 1e-6) / 1e9, 1),
                                     "roofline": roof_fft,
                                     "rel_l2_vs_fp64_64rows": rel_fft}},
        "roofline": roof,
        "cpu_baseline": {"value": round(cpu_row_s * rows * 1e6, 1), "unit": "us/call",
                         "cores": 1, "kind": "port",
                         "sample": "FP64 oracle direct convolution of row 0, scaled by 16 rows"},
    })


def bench_ew(args):
    """Fused elementwise expressions (mxr_ew_eval) at the 16 x 128 x 8001
    radar cube's size: the calcDets threshold of
    examples/simple_radar_pipeline.h:86-98, dets = |x|^2 > n (pfa^(-1/n) - 1) ba
    (x c64, n and ba f32, dets i32), and the in-place Doppler window
    x *= hamming<1>(126) (:353-355).  Generated kernels by default; the
    interpreter kernel when MXR_EW_JIT=0."""
    import numpy as np
    import torch
    from matx_amd import expr as E
    dev = "cuda:0"
    C, P, R = 16, 128, 8001
    g = torch.Generator(device="cpu").manual_seed(args.seed)
    x = torch.complex(torch.randn((C, P, R), generator=g), torch.randn((C, P, R), generator=g)).to(dev)
    nrm = torch.randint(20, 57, (C, P, R), generator=g).to(torch.float32).to(dev)
    ba = (torch.rand((C, P, R), generator=g) * 2).to(dev)
    dets = torch.empty((C, P, R), dtype=torch.int32, device=dev)
    xw = x.clone()[:, :126, :]  # the window rewrites its operand: a copy
    thr = lambda: E.run(dets, E.abs2(x) > E.wrap(nrm) * (E.pow(1e-5, -1.0 / E.wrap(nrm)) - 1.0) * ba)  # noqa: E731
    win = lambda: E.run(xw, E.wrap(xw) * E.hamming(1, 126))  # noqa: E731
    t0 = time.perf_counter()
    thr()
    torch.cuda.synchronize()
    first_ms = (time.perf_counter() - t0) * 1000.0
    k = max(args.steps, 50)

    def timed(fn):
        for _ in range(3):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(k):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return 1000.0 * e0.elapsed_time(e1) / k

    us_thr_eager, us_win = timed(thr), timed(win)
    # value: the threshold replayed from a hipGraph of 20 captured calls (the
    # eager figure above includes the Python front end's per-call lowering)
    G = 20
    gs = torch.cuda.Stream()
    gs.wait_stream(torch.cuda.current_stream())
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=gs):
        for _ in range(G):
            thr()
    torch.cuda.synchronize()
    graph.replay()
    torch.cuda.synchronize()
    reps = max(1, k // G)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        graph.replay()
    e1.record()
    torch.cuda.synchronize()
    us_thr = 1000.0 * e0.elapsed_time(e1) / (reps * G)
    del graph
    B_thr = (8 + 4 + 4 + 4) * C * P * R
    B_win = 2 * 8 * C * 126 * R
    jit = os.environ.get("MXR_EW_JIT", "1") != "0"
    name = "ew_jit_kernel" if jit else "ew_kernel"
    kern = profiled_kernels(thr, 10, [name])
    roof = kernel_roofline(kern, {name: B_thr})
    # check one channel of the threshold against FP64
    xs = x[0].cpu().numpy().astype(np.complex128)
    ns, bs = (t[0].cpu().numpy().astype(np.float64) for t in (nrm, ba))
    ref = (np.abs(xs) ** 2 > ns * (np.power(1e-5, -1.0 / ns) - 1.0) * bs)
    got = dets[0].cpu().numpy().astype(bool)
    pw, th = np.abs(xs) ** 2, ns * (np.power(1e-5, -1.0 / ns) - 1.0) * bs
    near = np.abs(pw - th) <= 1e-5 * th
    mismatch = int(((got != ref) & ~near).sum())
    return ({
        "metric": "fused elementwise expression GB/s (calcDets threshold, 16x128x8001)",
        "value": round(B_thr / (us_thr * 1e-6) / 1e9, 1), "unit": "GB/s", "n_gpus": 1, "steps": k,
        "warmup": 3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "c64/f32", "data": "synthetic (complex noise, integer norms, uniform averages, seeded)",
        "config": {"workload": "dets = |x|^2 > n (pfa^(-1/n) - 1) ba over 16 x 128 x 8001 cells; "
                               "x *= hamming<1>(126) in place over 16 x 126 x 8001",
                   "path": "generated kernels (hipRTC)" if jit else "interpreter kernel",
                   "us_threshold": round(us_thr, 2), "us_threshold_eager": round(us_thr_eager, 2),
                   "us_window_eager": round(us_win, 2),
                   "window_gbs": round(B_win / (us_win * 1e-6) / 1e9, 1),
                   "first_call_ms_incl_compile": round(first_ms, 1),
                   "mismatches_vs_fp64_channel0": mismatch},
        "roofline": roof,
        "cpu_baseline": None,
    })


def bench_grad(args):
    """Waveform/window design: 64 phase-only 1024-sample waveforms + a
    128-tap Doppler window (Hamming start); loss = mean integrated sidelobe
    level of each waveform's matched filter (fft_conv over 8192 lags) + mean
    peak-sidelobe penalty of the Doppler spectra at 24 target cells of a
    16x128x4096 cube.  One step = loss forward + backward (gradients w.r.t.
    phases and taps) through torch.autograd over the HIP kernels; timed eager
    and replayed from a captured hipGraph."""
    import numpy as np
    import torch
    from matx_amd import autograd as ag
    dev = "cuda:0"
    n, nd = 8192, 128
    phi0, h0, X0, tg0 = ag.make_design_problem(args.seed)
    p = torch.tensor(phi0, device=dev, requires_grad=True)
    t = torch.tensor(h0, device=dev, requires_grad=True)
    X = torch.tensor(X0, device=dev)
    tg = torch.tensor(np.array(tg0), dtype=torch.int32, device=dev)

    def fwd():
        return ag.design_loss(p, t, X, tg, n, nd)

    def step():
        p.grad = None
        t.grad = None
        L = fwd()
        L.backward()
        return L

    steps, warm = max(20, min(args.steps, 200)), max(3, min(args.warmup, 10))
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        for _ in range(warm):
            step()
        with torch.no_grad():
            for _ in range(warm):
                fwd()
    s.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def timed(fn, k):
        with torch.cuda.stream(s):
            e0.record(s)
            for _ in range(k):
                fn()
            e1.record(s)
        s.synchronize()
        return e0.elapsed_time(e1) / k

    def fwd_nograd():
        with torch.no_grad():
            return fwd()

    eager_fwd = timed(fwd_nograd, steps)
    eager_step = timed(step, steps)
    # whole step captured once, replayed
    g_step, g_fwd = torch.cuda.CUDAGraph(), torch.cuda.CUDAGraph()
    p.grad = None
    t.grad = None
    with torch.cuda.graph(g_step, stream=s):
        L_static = fwd()
        L_static.backward()
    # backward under capture wrote fresh gradient tensors from the graph's
    # pool; every replay rewrites them
    gp_static, gt_static = p.grad, t.grad
    with torch.cuda.graph(g_fwd, stream=s):
        with torch.no_grad():
            Lf_static = fwd()
    graph_step = timed(g_step.replay, steps)
    graph_fwd = timed(g_fwd.replay, steps)
    # eager gradients, then replays (with eager work in between) checked
    # against them
    with torch.cuda.stream(s):
        L_e = step()
    s.synchronize()
    gp_e, gt_e = p.grad.detach().clone(), t.grad.detach().clone()
    replay_dev = 0.0
    for _ in range(3):
        with torch.cuda.stream(s):
            fwd_nograd()
            g_step.replay()
        s.synchronize()
        replay_dev = max(replay_dev, float((gp_static - gp_e).norm() / gp_e.norm()),
                         float((gt_static - gt_e).norm() / gt_e.norm()))
    sys.path.insert(0, ROOT)
    from oracle import grad as og
    t0 = time.perf_counter()
    cpu_runs = 0
    while True:
        L64, gp64, gt64 = og.loss_and_grad(phi0, h0, X0, tg0, n, nd)
        cpu_runs += 1
        if time.perf_counter() - t0 >= max(args.cpu_seconds, 0.1):
            break
    cpu_s = (time.perf_counter() - t0) / cpu_runs
    gp, gt = p.grad.detach().cpu().numpy(), t.grad.detach().cpu().numpy()
    best = min(eager_step, graph_step)
    fwd_best = min(eager_fwd, graph_fwd)
    return ({
        "metric": "grad steps/s (waveform/window design: 64x1024 phases + 128 taps, "
                  "16x128x4096 Doppler cube)",
        "value": round(1000.0 / best, 2), "unit": "steps/s", "n_gpus": 1, "steps": steps,
        "warmup": warm, "ms_per_step": round(best, 5), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32/c64",
        "data": "synthetic (random phases, Hamming start, noise + 24 Doppler tones, seeded)",
        "config": {"workload": "loss = mean ISL of |fft_conv(s, s, 8192)|^2 over 64 phase-only "
                               "1024-sample waveforms + mean peak-sidelobe (q=8 norm) penalty "
                               "of |FFT_128(taps .* X)|^2 at 24 targets of a 16x128x4096 cube; "
                               "fwd + bwd w.r.t. phases and taps",
                   "fwd_ms_eager": round(eager_fwd, 4), "step_ms_eager": round(eager_step, 4),
                   "fwd_ms_graph": round(graph_fwd, 4), "step_ms_graph": round(graph_step, 4),
                   "fwd_bwd_over_fwd": round(best / fwd_best, 3),
                   "graph_replay_vs_eager_rel": replay_dev,
                   "loss_rel_vs_fp64": abs(float(L_e.detach()) - L64) / abs(L64),
                   "grad_phases_rel_vs_fp64": float(np.linalg.norm(gp - gp64) / np.linalg.norm(gp64)),
                   "grad_taps_rel_vs_fp64": float(np.linalg.norm(gt - gt64) / np.linalg.norm(gt64))},
        "cpu_baseline": {"value": round(1.0 / cpu_s, 3), "unit": "steps/s",
                         "cores": int(os.environ.get("OMP_NUM_THREADS",
                                                     len(os.sched_getaffinity(0)))),
                         "kind": "port",
                         "sample": f"FP64 numpy oracle loss + analytic gradient (oracle/grad.py), "
                                   f"{cpu_runs} runs"},
    })


def stage_dataflow_leg(jobs, shapes, ref_pws, s, reps):
    """Every ragged job as MatX's stage dataflow would run it
    (examples/simple_radar_pipeline.h:276-300 and 344-357,398, canceller off):
    waveform window, unit-energy normalisation and padded FFT; the cube's
    padded FFT, conj-multiply, IFFT; Doppler window into a zero-padded
    nd-row buffer, FFT along pulses, |.|^2 -- each step one launch of this
    library's generic operators (the norm's scalar through torch).  Returns
    (ms eager, ms graph-replayed, worst rel L2 against the fused per-job
    power maps)."""
    import torch
    import matx_amd
    from matx_amd import expr as E
    from matx_amd.ragged import next_pow2
    dev = jobs[0][0].device
    bufs = []
    for (x, w), (C, P, S, L) in zip(jobs, shapes):
        n, nd, R = next_pow2(S), next_pow2(P), S - L + 1
        bufs.append(dict(
            w1=torch.empty((L,), dtype=torch.complex64, device=dev),
            W=torch.empty((n,), dtype=torch.complex64, device=dev),
            X=torch.empty((C, P, n), dtype=torch.complex64, device=dev),
            T=torch.zeros((C, nd, R), dtype=torch.complex64, device=dev),
            pw=torch.empty((C, nd, R), dtype=torch.float32, device=dev)))

    def run():
        for (x, w), (C, P, S, L), b in zip(jobs, shapes, bufs):
            n, nd, R = next_pow2(S), next_pow2(P), S - L + 1
            E.run(b["w1"], E.wrap(w) * E.hamming(0, L))
            nrm = torch.linalg.vector_norm(b["w1"]).to(torch.complex64)
            E.run(b["w1"], E.wrap(b["w1"]) / nrm)
            matx_amd.fft(b["w1"], n, out=b["W"])
            matx_amd.fft(x, n, out=b["X"])
            E.run(b["X"], E.wrap(b["X"]) * E.conj(b["W"]))
            matx_amd.ifft(b["X"], out=b["X"])
            T = b["T"][:, :P, :]
            E.run(T, E.wrap(b["X"][:, :, :R]) * E.hamming(1, P))
            matx_amd.fft(b["T"], axis=1, out=b["T"])
            E.run(b["pw"], E.abs2(b["T"]))
            # the next pass needs rows P..nd-1 of T zero again
            b["T"][:, P:, :].zero_()

    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    with torch.cuda.stream(s):
        run()
        s.synchronize()
        e0.record(s)
        for _ in range(reps):
            run()
        e1.record(s)
    s.synchronize()
    ms_eager = e0.elapsed_time(e1) / reps
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=s):
        run()
    graph.replay()
    s.synchronize()
    with torch.cuda.stream(s):
        e0.record(s)
        for _ in range(reps):
            graph.replay()
        e1.record(s)
    s.synchronize()
    ms_graph = e0.elapsed_time(e1) / reps
    worst = max(float((b["pw"] - r).norm() / r.norm()) for b, r in zip(bufs, ref_pws))
    del graph
    return ms_eager, ms_graph, worst


def bench_ragged(args):
    """Ragged batch: 256 independent jobs, C uniform in {1..8}, P in
    {16, 32, 64, 128, 256}, S uniform in [1024, 16384], L uniform in
    [64, min(2048, S)]; x = unit complex noise + 5 echoes (uniform delays in
    [0, S-L], 10 dB) of the job's LFM (0.5 fs, Hamming-windowed), w = that
    LFM.  Per job: pulse compression (n = next_pow2(S)) then Hamming +
    nd-point Doppler FFT (nd = next_pow2(P)) + |.|^2.  Plain: per-job entry
    calls back to back; batched: the ragged plan replayed from one
    hipGraph.  value = plain time / batched time."""
    import numpy as np
    import torch
    import matx_amd
    from matx_amd.ragged import RaggedBatch, next_pow2
    dev = "cuda:0"
    rng = np.random.default_rng(args.seed)
    g = torch.Generator(device=dev).manual_seed(args.seed)
    jobs, shapes = [], []
    for _ in range(256):
        C = int(rng.integers(1, 9))
        P = int(rng.choice([16, 32, 64, 128, 256]))
        S = int(rng.integers(1024, 16385))
        L = int(rng.integers(64, min(2048, S) + 1))
        n = np.arange(L)
        lfm = np.exp(2j * np.pi * (-0.25 * n + 0.25 / L * n * n)) * np.hamming(L)
        w = torch.tensor(lfm.astype(np.complex64), device=dev)
        x = torch.randn((C, P, S), dtype=torch.complex64, device=dev, generator=g)
        for d in rng.integers(0, S - L + 1, 5):
            x[:, :, int(d):int(d) + L] += np.sqrt(10.0) * w
        jobs.append((x, w))
        shapes.append((C, P, S, L))
    # plain path: per-job entry calls, buffers allocated up front
    ys = [torch.empty((C, P, S - L + 1), dtype=torch.complex64, device=dev) for C, P, S, L in shapes]
    pws = [torch.empty((C, next_pow2(P), S - L + 1), dtype=torch.float32, device=dev)
           for C, P, S, L in shapes]

    def plain():
        for (x, w), y, pw, (C, P, S, L) in zip(jobs, ys, pws, shapes):
            matx_amd.pulse_compress(x, w, next_pow2(S), out=y)
            matx_amd.doppler(y, next_pow2(P), tpc="none", win_rows=P, pw=pw)

    s = torch.cuda.Stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def timed(fn, k):
        with torch.cuda.stream(s):
            fn()
            s.synchronize()
            e0.record(s)
            for _ in range(k):
                fn()
            e1.record(s)
        s.synchronize()
        return e0.elapsed_time(e1) / k

    reps = 3
    t_plain = timed(plain, reps)
    t_tune0 = time.perf_counter()
    rb = RaggedBatch(jobs, autotune=True, stream=s)
    t_tune = time.perf_counter() - t_tune0
    t_eager = timed(rb.execute, reps)
    graph = rb.graph(stream=s)
    t_graph = timed(graph.replay, reps)
    rbp = RaggedBatch(jobs, autotune=True, persistent=True, stream=s)
    t_persist = timed(rbp.execute, reps)
    rbp.close()
    # parity of the batched power maps against the per-job calls
    with torch.cuda.stream(s):
        plain()
        graph.replay()
    s.synchronize()
    worst = max(float((a - b).norm() / b.norm()) for a, b in zip(rb.pw, pws))
    sd_eager, sd_graph, sd_worst = stage_dataflow_leg(jobs, shapes, pws, s, reps)
    B = sum(8 * C * P * S + 16 * C * P * (S - L + 1) + 4 * C * next_pow2(P) * (S - L + 1)
            for C, P, S, L in shapes)
    Ms = sorted({rb.block_size(i) for i in range(len(jobs))})
    best = min(t_graph, t_eager, t_persist)
    # CPU baseline: the FP64 oracle on the first jobs within the budget
    sys.path.insert(0, ROOT)
    from oracle import oracle as orc
    orc.build()
    t0, done = time.perf_counter(), 0
    for (x, w), (C, P, S, L) in zip(jobs, shapes):
        xh, wh = x.cpu().numpy(), w.cpu().numpy()
        n = next_pow2(S)
        Wc = orc.waveform_spectrum(wh, n, True, True)
        y = orc.pulse_compress(xh.reshape(C * P, S), Wc, S - L + 1).reshape(C, P, S - L + 1)
        orc.abs2(orc.doppler(y, P, next_pow2(P)))
        done += 1
        if time.perf_counter() - t0 >= max(args.cpu_seconds, 0.5):
            break
    cpu_jobs_s = done / (time.perf_counter() - t0)
    return ({
        "metric": "ragged-batch speedup (256 uneven pulse-compression + Doppler jobs, "
                  "hipGraph of the batched launches vs per-job launches)",
        "value": round(t_plain / t_graph, 3), "unit": "x", "n_gpus": 1, "steps": reps,
        "warmup": 1, "ms_per_step": round(t_graph, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "c64",
        "data": "synthetic (noise + 5 LFM echoes per job at 10 dB, seeded)",
        "config": {"workload": "256 jobs: C~U{1..8}, P in {16,32,64,128,256}, S~U[1024,16384], "
                               "L~U[64,min(2048,S)]; pulse compression (n=next_pow2(S)) + "
                               "Hamming + next_pow2(P)-pt Doppler FFT + |.|^2 per job",
                   "plain_ms": round(t_plain, 3), "batched_eager_ms": round(t_eager, 3),
                   "batched_graph_ms": round(t_graph, 3),
                   "batched_persistent_ms": round(t_persist, 3),
                   "speedup_best": round(t_plain / best, 3),
                   "autotune_s": round(t_tune, 2), "block_sizes_used": Ms,
                   "algorithmic_bytes": B,
                   "batched_gbs": round(B / (t_graph * 1e-3) / 1e9, 1),
                   "plain_gbs": round(B / (t_plain * 1e-3) / 1e9, 1),
                   "max_rel_vs_per_job": worst,
                   "jobs_per_s_batched": round(256 / (t_graph * 1e-3), 1),
                   "matx_stage_dataflow": {
                       "what": "each job as MatX's stage dataflow (about 10 launches of this "
                               "library's generic operators per job); value above compares "
                               "with this library's fused per-job entries instead",
                       "eager_ms": round(sd_eager, 3), "graph_ms": round(sd_graph, 3),
                       "speedup_batched_graph_vs_dataflow_graph": round(sd_graph / t_graph, 3),
                       "speedup_batched_graph_vs_dataflow_eager": round(sd_eager / t_graph, 3),
                       "max_rel_vs_per_job": sd_worst}},
        "cpu_baseline": {"value": round(cpu_jobs_s, 4), "unit": "jobs/s",
                         "cores": int(os.environ.get("OMP_NUM_THREADS",
                                                     len(os.sched_getaffinity(0)))),
                         "kind": "port",
                         "sample": f"FP64 oracle pulse compression + Doppler on the first {done} "
                                   "jobs of the batch"},
    })


METRIC = ("cubes/s (16x128x9000 c64) & % HBM peak; ragged-batch speedup; grad steps/s")


def bench_all(args):
    """The headline line: the 16 x 128 x 9000 pipeline is `value`; every other
    workload of the metric runs after it in the same process and lands in
    config.sub_measurements, each with its own parity check and baseline."""
    import torch
    head = bench_pipeline(args, WORKLOADS["radar"])
    sub_args = argparse.Namespace(**vars(args))
    sub_args.cpu_seconds = args.cpu_seconds / 3.0
    sub_args.steps, sub_args.warmup = min(args.steps, 200), min(args.warmup, 20)
    legs = [("ragged_batch", bench_ragged), ("grad_design", bench_grad),
            ("fft_conv", bench_fftconv),
            ("large_pipeline", lambda a: bench_pipeline(a, WORKLOADS["large"])),
            ("matx_stage_form", bench_stage_form)]
    subs = {}
    for name, fn in legs:
        print(f"bench: {name}", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        try:
            subs[name] = fn(sub_args)
        except Exception as e:  # a failing leg is reported, the headline stands
            subs[name] = {"error": f"{type(e).__name__}: {e}"[:400]}
        subs[name]["wall_s"] = round(time.perf_counter() - t0, 1)
        torch.cuda.synchronize()
        torch.cuda.empty_cache()
    head["metric"] = METRIC
    cfg = head["config"]
    rg, gd = subs.get("ragged_batch", {}), subs.get("grad_design", {})
    cfg["ragged_speedup"] = rg.get("value")
    cfg["grad_steps_per_s"] = gd.get("value")
    cfg["grad_fwd_bwd_over_fwd"] = gd.get("config", {}).get("fwd_bwd_over_fwd")
    cfg["sub_measurements"] = subs
    return head


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the path shards by channel with no exchange between GPUs)", file=sys.stderr)
        sys.exit(2)
    import torch
    if args.calibrate_copy:
        a = torch.ones(128 << 20, dtype=torch.float32, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes": a.numel() * 4, "copies": 5}))
        return
    runners = {"all": bench_all, "fftconv": bench_fftconv, "grad": bench_grad,
               "ragged": bench_ragged, "conv": bench_conv, "ew": bench_ew,
               "stage": bench_stage_form,
               "radar": lambda a: bench_pipeline(a, WORKLOADS["radar"]),
               "large": lambda a: bench_pipeline(a, WORKLOADS["large"])}
    print(json.dumps(runners[args.workload](args)))


# (Doppler, range) window and guard sizes.  radar: 13 range x 5 Doppler
# window with 2 x 1 guard cells around the cell under test; large: 21 x 7.
WORKLOADS = {
    "radar": dict(C=16, P=128, S=9000, L=1000, nd=128, bw=0.4, win=(5, 13), guard=(3, 5),
                  targets=20, snr=(0.0, 20.0), steps=None, cubes=4, streams=6,
                  tag="radar16x128x9000",
                  metric="cubes/s (16x128x9000 c64 radar pipeline)",
                  desc="radar pipeline 16ch x 128 pulses x 9000 samples c64, L=1000 LFM (0.4 fs), "
                       "N=16384 pulse compression, [1,-2,1] canceller (126 rows), Hamming + "
                       "128-pt Doppler FFT, CA-CFAR 13x5 window / 2x1 guard cells (5x3 hole), "
                       "Pfa 1e-5; waveform spectrum prepared once per waveform"),
    "large": dict(C=64, P=512, S=32768, L=4096, nd=512, bw=0.5, win=(7, 21), guard=(3, 5),
                  targets=200, snr=(0.0, 25.0), steps=20, cubes=1, streams=1,
                  tag="radar64x512x32768",
                  metric="cubes/s (64x512x32768 c64 radar pipeline)",
                  desc="radar pipeline 64ch x 512 pulses x 32768 samples c64 (8.6 GB), L=4096 "
                       "LFM (0.5 fs), N=32768 pulse compression, [1,-2,1] canceller (510 rows), "
                       "Hamming + 512-pt Doppler FFT, CA-CFAR 21x7 window / 2x1 guard cells, "
                       "Pfa 1e-5; waveform spectrum prepared once per waveform"),
}


def bench_pipeline(args, cfg):
    import torch
    import matx_amd
    from matx_amd import _lib

    dev = "cuda:0"
    cpu_s = args.cpu_seconds
    C, P, S, L, nd = cfg["C"], cfg["P"], cfg["S"], cfg["L"], cfg["nd"]
    win, guard = cfg["win"], cfg["guard"]
    steps = args.steps if cfg["steps"] is None else min(args.steps, cfg["steps"])
    warmup = args.warmup if cfg["steps"] is None else min(args.warmup, 3)
    # distinct resident cubes, replayed in turn: together they exceed the
    # 256 MB last-level cache, so every cube's samples come from HBM
    ncubes = cfg["cubes"]
    cubes = [synth_cube(C, P, S, L, nd, cfg["targets"], cfg["snr"], args.seed + i, dev, bw=cfg["bw"])
             for i in range(ncubes)]
    xs = [c[0] for c in cubes]
    x, w, targets = cubes[0]
    fused = nd <= 128
    pipe = matx_amd.FusedRadarPipeline(C, P, S, L, nd=nd, tpc="valid", win=win, guard=guard,
                                       want_pow=not fused)
    pipe.set_waveform(w)
    for i in range(warmup):
        pipe.step(xs[i % ncubes])
    pipe.sync()
    torch.cuda.synchronize()

    lib = _lib.lib()
    stream = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    # throughput: cubes replayed from one captured hipGraph, with no
    # per-kernel instrumentation in the timed region.  As in MatX's radar
    # example (examples/simple_radar_pipeline.cu:44-137: num_streams
    # pipelines, each with its own stream), `streams` pipelines run side by
    # side: the graph forks into one branch per pipeline (captured on its own
    # HIP stream, G cubes cycling through the resident cubes) and joins at
    # the end, so one cube's pulse compression can run beside another's
    # pulse-axis + CFAR pass
    ns = cfg["streams"] if args.streams is None else max(1, args.streams)
    pipes = [pipe]
    for _ in range(ns - 1):
        q = matx_amd.FusedRadarPipeline(C, P, S, L, nd=nd, tpc="valid", win=win, guard=guard,
                                        want_pow=not fused)
        q.set_waveform(w)
        for i in range(min(warmup, ncubes)):
            q.step(xs[i % ncubes])
        q.sync()
        pipes.append(q)
    G = max(ncubes, min(steps, 48) // ncubes * ncubes)
    reps = max(1, round(steps / (G * ns)))
    branches = [HipStream() for _ in pipes]
    gs = torch.cuda.Stream()
    gs.wait_stream(stream)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=gs):
        fork = torch.cuda.Event()
        fork.record(gs)
        for br in branches:
            br.stream.wait_event(fork)
        for k, (p, br) in enumerate(zip(pipes, branches)):
            with torch.cuda.stream(br.stream):
                for i in range(G):
                    p.step(xs[(i + k) % ncubes])
        for br in branches:
            join = torch.cuda.Event()
            join.record(br.stream)
            gs.wait_event(join)
    torch.cuda.synchronize()
    graph.replay()
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(reps):
        graph.replay()
    e1.record(stream)
    torch.cuda.synchronize()
    steps_timed = reps * G * ns
    ms = e0.elapsed_time(e1) / steps_timed
    del graph, pipes[1:]
    for br in branches:
        br.close()
    # eager loop (instrumented when profiling): per-kernel durations from
    # HIP events recorded around every launch on its own stream
    prof = not args.no_profile
    if prof:
        _lib.check(lib.mxr_profile_reset())
        _lib.check(lib.mxr_profile_enable(1))
    torch.cuda.synchronize()
    e0.record(stream)
    for i in range(steps):
        pipe.step(xs[i % ncubes])
    pipe.sync()
    e1.record(stream)
    torch.cuda.synchronize()
    ms_eager = e0.elapsed_time(e1) / steps
    if prof:
        _lib.check(lib.mxr_profile_enable(0))

    B = pipe.algorithmic_bytes()
    value = 1000.0 / ms
    R = S - L + 1
    # algorithmic bytes per launch of each kernel: pulse compression reads x
    # and writes the compressed cube; the fused pulse-axis + CFAR pass reads
    # the compressed cube and writes the int32 detections; unfused, the
    # pulse-axis pass writes the power map and the CFAR kernel reads it
    kbytes = {"pc4096_kernel": 8 * C * P * S + 8 * C * P * R,
              "pc8192_kernel": 8 * C * P * S + 8 * C * P * R,
              "pc_ols_kernel": 8 * C * P * S + 8 * C * P * R,
              "dopcfar_kernel": 8 * C * P * R + 4 * C * nd * R,
              "doppler_kernel": 8 * C * P * R + 4 * C * nd * R,
              "cfar_kernel": 4 * C * nd * R + 4 * C * nd * R,
              "cfar_lds_kernel": 4 * C * nd * R + 4 * C * nd * R}
    if fused:
        kbytes.pop("doppler_kernel")
        kbytes.pop("cfar_kernel")   # the name also matches dopcfar_kernel
        kbytes.pop("cfar_lds_kernel")
    kern = {}
    if prof:
        for k in ["wfspec_kernel"] + list(kbytes):
            n, tot = _lib.profile_query(k)
            kern[k] = {"launches": n, "avg_us": 1000.0 * tot / max(n, 1)}
    rooflines = {}
    for k, b in kbytes.items():
        if prof and kern.get(k, {}).get("launches"):
            ach = b / (kern[k]["avg_us"] * 1e-6) / 1e9
            rooflines[k] = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                            "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                            "traffic": load_traffic(k, cfg["tag"]), "kernel": k,
                            "algorithmic_bytes_per_launch": b,
                            "avg_us": round(kern[k]["avg_us"], 2)}
    dominant = max(rooflines, key=lambda k: kern[k]["avg_us"]) if rooflines else None
    roofline = rooflines.get(dominant)

    # one more cube with the power map written, for the parity check
    pipe.set_outputs(want_pow=True, want_ba=False)
    pipe.step(x)
    pipe.sync()
    torch.cuda.synchronize()
    dets = pipe.dets
    det_cells = int(dets.sum().item())
    # count the injected targets that land on a detection at (Doppler, range)
    hit = sum(int(dets[:, dop, rb].any().item()) for rb, dop, _ in targets)

    oracle, parity = None, None
    if cpu_s > 0:
        # every channel of the headline cube; channel 0 of the 8.6 GB cube
        full = cfg["tag"] == "radar16x128x9000"
        nch = C if full else 1
        oracle, parity = cpu_baseline_leg(x[:nch].cpu().numpy(), w.cpu().numpy(),
                                          pipe.pw[:nch].cpu().numpy(), dets[:nch].cpu().numpy(),
                                          nd, cpu_s, C, cfar_mask(win, guard), all_channels=full)
    standin = None
    if cpu_s > 0 and cfg["tag"] == "radar16x128x9000":
        standin = cpu_standin_leg(x[0].cpu().numpy(), w.cpu().numpy(), nd, win, guard, cpu_s, C)
    fwd_bwd = None
    if cfg["tag"] == "radar64x512x32768":
        del pipe, cubes, xs
        torch.cuda.empty_cache()
        fwd_bwd = bench_large_grad(args, cfg, x, w, targets)

    return {
        "metric": cfg["metric"],
        "value": round(value, 2),
        "unit": "cubes/s",
        "n_gpus": 1,
        "steps": steps_timed,
        "warmup": warmup,
        "ms_per_step": round(ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64",
        "data": f"synthetic ({ncubes} resident cube(s), each noise + {len(targets)} LFM point "
                f"targets, seeded)",
        "config": {"workload": cfg["desc"],
                   "channels": C, "pulses": P, "samples": S, "waveform": L, "doppler": nd,
                   "resident_cubes": ncubes,
                   "streams": ns,
                   "timing": (f"value: {steps_timed} cubes, {reps} replay(s) of one hipGraph "
                              f"that forks into {ns} branch(es), one pipeline per branch on its "
                              f"own HIP stream, {G} captured cubes per branch cycling over "
                              f"{ncubes} resident cube(s); HIP events around the replays, no "
                              f"instrumentation; eager loop of {steps} cubes on one stream with "
                              f"per-launch HIP events: {round(ms_eager, 5)} ms per cube"),
                   "ms_per_step_eager_instrumented": round(ms_eager, 5),
                   "algorithmic_bytes_per_cube": B,
                   "pipeline_hbm_gbs": round(B * value / 1e9, 1),
                   "pipeline_hbm_frac": round(B * value / (HBM_PEAK_GBS * 1e9), 4),
                   "per_kernel_us": {k: round(v["avg_us"], 2) for k, v in kern.items() if v["launches"]},
                   "detections": det_cells, "targets_detected": f"{hit}/{len(targets)}",
                   "kernel_rooflines": {k: v for k, v in rooflines.items() if k != dominant},
                   "parity": parity,
                   **({"cpu_fp64_oracle": oracle} if oracle else {}),
                   **({"fwd_bwd": fwd_bwd} if fwd_bwd else {})},
        "roofline": roofline,
        "cpu_baseline": standin if standin else oracle,
    }


def bench_stage_form(args):
    """MatX's stage-by-stage dataflow through the C++ operator API
    (matx_amd/examples: PulseCompression, ThreePulseCanceller,
    DopplerProcessing, CFARDetections written with fft / ifft / conv1d /
    conv2d / expressions, MatX's sizes: N = 16384, 256-point Doppler FFT over
    all 128 pulses, 13 x 5 CFAR mask) on the 16 x 128 x 9000 cube, run by the
    compiled example program in a child process."""
    exe = os.path.join(ROOT, "matx_amd", "_build", "bin", "simple_radar_pipeline")
    if not os.path.exists(exe):
        return {"error": "simple_radar_pipeline is not built"}
    iters = max(20, min(args.steps, 100))
    r = subprocess.run([exe, str(iters), "--json"], capture_output=True, text=True, timeout=300)
    if r.returncode != 0:
        return {"error": r.stderr.strip()[-400:]}
    line = [l for l in r.stdout.splitlines() if l.startswith("{")][-1]
    out = json.loads(line)
    out["what"] = ("MatX's four stage calls on the C++ front end (generic operators, about 20 "
                   "launches per cube, MatX's buffer sizes), timed with the executor's events")
    return out


def bench_large_grad(args, cfg, x, w, targets):
    """Forward + backward of the large cube with respect to the waveform
    phases: pow = doppler(pulse_compress(x, exp(i phi)), 512) (valid
    canceller), L = mean clutter power - mean target power over the
    injected target cells.  The backward recomputes the Doppler spectrum
    from the compressed cube (the only activation kept) and sums the
    waveform gradient over all 32768 pulses.  The FP64 oracle gradient of
    channel 0 is compared with the GPU gradient of that channel."""
    import numpy as np
    import torch
    import matx_amd
    from matx_amd import _lib
    from matx_amd import autograd as ag

    C, P, S, L, nd = cfg["C"], cfg["P"], cfg["S"], cfg["L"], cfg["nd"]
    R = S - L + 1
    n = 1
    while n < S:
        n *= 2
    cells = [(dop, rb) for rb, dop, _ in targets]
    wmap = ag.contrast_weights(nd, R, cells, C, device=x.device)
    phi0 = torch.angle(w).to(torch.float32)
    plan = matx_amd.ops.PulseCompressionPlan(S, L, n)

    def step(xc=x, p0=phi0):
        p = p0.clone().requires_grad_()
        wv = torch.polar(torch.ones_like(p), p)
        loss = ag.weighted_power_sum(ag.radar_power(wv, xc, plan, nd, tpc="valid"), wmap)
        loss.backward()
        return loss.detach(), p.grad

    def fwd_only():
        with torch.no_grad():
            wv = torch.polar(torch.ones_like(phi0), phi0)
            return ag.weighted_power_sum(ag.radar_power(wv, x, plan, nd, tpc="valid"), wmap)

    def timed(fn, k):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(k):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / k * 1000.0

    k = max(1, min(args.steps, 5))
    print("large fwd+bwd: warm-up", file=sys.stderr, flush=True)
    step()
    fwd_only()
    lib = _lib.lib()
    prof = not args.no_profile
    if prof:
        _lib.check(lib.mxr_profile_reset())
        _lib.check(lib.mxr_profile_enable(1))
    ms_fb = timed(step, k)
    if prof:
        _lib.check(lib.mxr_profile_enable(0))
    ms_f = timed(fwd_only, k)
    kern = {}
    if prof:
        for name in ["pc8192_kernel", "pc_ols_kernel", "doppler_kernel", "doppler_adj_kernel", "wgrad8192_kernel", "pc_wgrad_kernel",
                     "wgrad_reduce_kernel", "wgrad_finish_kernel", "wfspec_kernel"]:
            cnt, tot = _lib.profile_query(name)
            if cnt:
                kern[name] = round(1000.0 * tot / cnt, 1)
    # algorithmic bytes of one forward + backward: forward reads x, writes
    # and reads the compressed cube, writes and reads the power map; the
    # backward reads the compressed cube, writes the lag gradient, then reads
    # x and the lag gradient (the weight map is 1/C of a power map, resident)
    B = 2 * 8 * C * P * S + 5 * 8 * C * P * R + 2 * 4 * C * nd * R
    out = {"grad_steps_per_s": round(1000.0 / ms_fb, 3), "ms_fwd_bwd": round(ms_fb, 2),
           "ms_fwd": round(ms_f, 2), "fwd_bwd_over_fwd": round(ms_fb / ms_f, 3),
           "algorithmic_bytes_fwd_bwd": B,
           "hbm_gbs_fwd_bwd": round(B / (ms_fb * 1e-3) / 1e9, 1),
           "hbm_frac_fwd_bwd": round(B / (ms_fb * 1e-3) / (HBM_PEAK_GBS * 1e9), 4),
           "per_kernel_us": kern, "steps": k,
           "loss": "mean clutter power - mean target power over the injected target cells",
           "checkpointing": "Doppler spectrum recomputed in the backward; saved: compressed cube"}
    # size-independent checks on the GPU: determinism, and the gradient of
    # the cube equals the sum of its channel-block gradients
    print(f"large fwd+bwd: {ms_fb:.1f} ms/step; checks", file=sys.stderr, flush=True)
    _, g_all = step()
    _, g_again = step()
    blocks = [(c0, min(c0 + 16, C)) for c0 in range(0, C, 16)]
    g_sum = sum(step(x[a:b])[1].double() for a, b in blocks)
    out["deterministic"] = bool(torch.equal(g_all, g_again))
    out["channel_block_sum_rel"] = float((g_all.double() - g_sum).norm() / g_sum.norm())
    if args.cpu_seconds > 0:
        sys.path.insert(0, ROOT)
        from oracle import grad as og
        print("large fwd+bwd: FP64 oracle on channel 0", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        _, g0 = step(x[0:1])
        L64, g64 = og.radar_loss_and_grad(phi0.double().cpu().numpy(), x[0:1].cpu().numpy(),
                                          wmap.double().cpu().numpy(), n, nd, "valid")
        g0 = g0.double().cpu().numpy()
        out["fp64_oracle"] = {"channel": 0,
                              "grad_rel_l2": float(np.linalg.norm(g0 - g64) / np.linalg.norm(g64)),
                              "oracle_s": round(time.perf_counter() - t0, 1)}
    plan.close()
    return out


if __name__ == "__main__":
    main()
