#!/usr/bin/env python3
"""Writes the committed golden fixtures under tests/golden/.

This is an independent numpy/scipy restatement, written from the text of the
MatX test generator ``test/test_vectors/generators/01_radar.py:24-125`` and of
``matx_common.randn_ndarray`` (``matx_common.py:19-23``).  It does not import
anything from the reference, and it does not use the C oracle: the point of the
fixtures is to pin that oracle with library routines (numpy's pocketfft,
scipy.signal.convolve2d) that share no code with it.

Inputs are MatX's own test input: ``np.random.seed(1234)``; x = randn(128,9000)
+ 1j*randn(128,9000) then waveform = randn(1000) + 1j*randn(1000), rounded to
complex64 (the radar test instantiates complex<float>,
``test/01_radar/MultiChannelRadarPipeline.cu:79``).  The legacy RandomState
stream is deterministic, so tests regenerate the inputs and use the fixture
only for the expected outputs.

Two chains are restated:

* ``model`` -- exactly 01_radar.py: VALID [1,-2,1] (rows 1..126), Hamming(126),
  256-point Doppler FFT, 5x13 CFAR mask with a 3x3 hole, Pfa 1e-5.  The
  reference test asserts 17 detections per channel for it
  (MultiChannelRadarPipeline.cu:187-197).
* ``matx`` -- MatX's end-to-end stage dataflow
  (examples/simple_radar_pipeline.h:273-436): SAME canceller over all 128
  pulses, Hamming(126) on rows 0..125 only, rows 126..127 unwindowed, rows
  128..255 zero, 256-point FFT, CFAR over all 256 rows.

Run:  python tests/golden/make_golden.py
"""
import json
import os

import numpy as np
from scipy import signal

HERE = os.path.dirname(os.path.abspath(__file__))
P, S, L = 128, 9000, 1000
R = S - L + 1
NFFT = 16384
ND = 256
PFA = 1e-5


def matx_test_inputs():
    np.random.seed(1234)
    x = (np.random.randn(P, S) + 1j * np.random.randn(P, S)).astype(np.complex64)
    w = (np.random.randn(L) + 1j * np.random.randn(L)).astype(np.complex64)
    return x, w


def cfar_mask():
    # 01_radar.py:86-98 writes the mask transposed (13x5) and transposes it.
    m = np.ones((13, 5))
    m[5:8, 1:4] = 0
    return np.transpose(m)


def chain(x, w, mode):
    x = x.astype(np.complex128)
    w = w.astype(np.complex128)
    wwin = w * signal.windows.hamming(L)
    wn = wwin / np.linalg.norm(wwin)
    W = np.conj(np.fft.fft(wn, NFFT))
    y = np.fft.ifft(np.fft.fft(x, NFFT, 1) * W, NFFT, 1)[:, :R]
    k = np.array([[1.0], [-2.0], [1.0]])
    if mode == "model":
        t = signal.convolve2d(y, k, "valid")                     # rows 1..126
        tw = t * signal.windows.hamming(P - 2)[:, None]
    else:
        t = signal.convolve2d(y, k, "same")                      # all 128 rows
        tw = t.copy()
        tw[: P - 2] *= signal.windows.hamming(P - 2)[:, None]
    D = np.fft.fft(tw, ND, 0)
    pw = np.abs(D) ** 2
    mask = cfar_mask()
    norm = signal.convolve2d(np.ones(pw.shape), mask, "same")
    ba = signal.convolve2d(pw, mask, "same") / norm
    alpha = norm * (np.power(PFA, -1.0 / norm) - 1)
    thr = alpha * ba
    dets = (pw > thr).astype(np.int32)
    return dict(W=W, y=y, t=t, D=D, pw=pw, norm=norm, ba=ba, dets=dets,
                margin=(pw - thr) / thr)


def main():
    x, w = matx_test_inputs()
    rng = np.random.RandomState(99)
    out = {}
    meta = {"inputs": "np.random.seed(1234); x=randn(128,9000)+1j*randn(128,9000); "
                      "w=randn(1000)+1j*randn(1000); both complex64",
            "N": NFFT, "Nd": ND, "R": R, "pfa": PFA}
    for mode in ("model", "matx"):
        r = chain(x, w, mode)
        det_idx = np.argwhere(r["dets"] == 1).astype(np.int32)
        # sample positions for stage values
        iy = rng.randint(0, P, 64)
        ir = rng.randint(0, R, 64)
        it = rng.randint(0, r["t"].shape[0], 64)
        idp = rng.randint(0, ND, 64)
        out[f"{mode}_det_idx"] = det_idx
        out[f"{mode}_y_idx"] = np.stack([iy, ir], 1).astype(np.int32)
        out[f"{mode}_y_val"] = r["y"][iy, ir]
        out[f"{mode}_t_idx"] = np.stack([it, ir], 1).astype(np.int32)
        out[f"{mode}_t_val"] = r["t"][it, ir]
        out[f"{mode}_D_idx"] = np.stack([idp, ir], 1).astype(np.int32)
        out[f"{mode}_D_val"] = r["D"][idp, ir]
        out[f"{mode}_ba_val"] = r["ba"][idp, ir]
        out[f"{mode}_pw_sum"] = np.array([r["pw"].sum()])
        out[f"{mode}_y_energy"] = np.array([np.sum(np.abs(r["y"]) ** 2)])
        out[f"{mode}_min_abs_margin"] = np.array([np.abs(r["margin"]).min()])
        meta[f"{mode}_dets"] = int(r["dets"].sum())
        if mode == "model":
            out["W_head"] = r["W"][:64]
            out["norm_rows"] = r["norm"][:3, :]
            out["norm_last_rows"] = r["norm"][-3:, :]
    np.savez_compressed(os.path.join(HERE, "radar_seed1234.npz"), **out)
    with open(os.path.join(HERE, "radar_seed1234.json"), "w") as f:
        json.dump(meta, f, indent=1)
    print(meta)


if __name__ == "__main__":
    main()
