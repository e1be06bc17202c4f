"""Pins the FP64 oracle before anything is compared against it (CPU only).

Each oracle routine is checked against something that shares no code with it:
numpy's pocketfft, an O(n^2) DFT, scipy.signal.convolve2d, np.hamming, the
ConvCorr.cu known-answer tables, closed-form CFAR counts, and the reference
radar test's own invariant (17 detections per channel on its seed-1234 input,
test/01_radar/MultiChannelRadarPipeline.cu:187-197).  Each family also has a
"teeth" check: a plausible mistake (wrong sign, un-flipped filter, dropped
term) is planted in a numpy copy of the computation on the CPU and must fail
the same comparison.
"""
import json
import os

import numpy as np
import pytest
from scipy import signal

HERE = os.path.dirname(os.path.abspath(__file__))
KAT = json.load(open(os.path.join(HERE, "golden", "kat_convcorr.json")))


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300)


# ------------------------------------------------------------------- FFT
@pytest.mark.parametrize("n", [1, 2, 4, 8, 16, 64, 256, 1024, 4096, 16384])
def test_fft_matches_numpy(orc, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    assert rel(orc.fft(x), np.fft.fft(x)) < 1e-13
    assert rel(orc.ifft(x), np.fft.ifft(x)) < 1e-13


@pytest.mark.parametrize("n", [3, 5, 12, 17, 100])
def test_fft_non_pow2_uses_dft(orc, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    assert rel(orc.fft(x), np.fft.fft(x)) < 1e-12


@pytest.mark.parametrize("n", [8, 32, 128])
def test_dft_closed_forms(orc, n):
    # impulse -> all-ones; shifted impulse -> phase ramp; constant -> n*delta
    d = np.zeros(n, complex); d[0] = 1
    assert np.allclose(orc.dft(d), np.ones(n), atol=1e-14)
    s = 3
    d = np.zeros(n, complex); d[s] = 1
    k = np.arange(n)
    assert np.allclose(orc.dft(d), np.exp(-2j * np.pi * s * k / n), atol=1e-13)
    assert np.allclose(orc.idft_unscaled(d), np.exp(2j * np.pi * s * k / n), atol=1e-13)
    c = orc.dft(np.ones(n))
    e = np.zeros(n, complex); e[0] = n
    assert np.allclose(c, e, atol=1e-11)


def test_fft_parseval_linearity_roundtrip(orc):
    rng = np.random.default_rng(7)
    n = 2048
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    y = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    X = orc.fft(x)
    assert abs(np.sum(abs(X) ** 2) / n - np.sum(abs(x) ** 2)) < 1e-9 * np.sum(abs(x) ** 2)
    assert rel(orc.fft(2 * x - 3j * y), 2 * X - 3j * orc.fft(y)) < 1e-13
    assert rel(orc.ifft(X), x) < 1e-13


def test_fft_norm_and_size_rules(orc):
    rng = np.random.default_rng(3)
    x = rng.standard_normal(100) + 1j * rng.standard_normal(100)
    # padding and truncation (fft_common.h:39-140)
    assert rel(orc.fft(x, 256), np.fft.fft(x, 256)) < 1e-13
    assert rel(orc.fft(x, 64), np.fft.fft(x, 64)) < 1e-13
    for norm in ("backward", "ortho", "forward"):
        assert rel(orc.fft(x, 128, norm=norm), np.fft.fft(x, 128, norm=norm)) < 1e-13
        assert rel(orc.ifft(x, 128, norm=norm), np.fft.ifft(x, 128, norm=norm)) < 1e-13


def test_fft_axis_equivalence(orc):
    # fft along axis 0 == fft of the transposed rows (FFT.cu:174-291)
    rng = np.random.default_rng(5)
    x = rng.standard_normal((64, 10)) + 1j * rng.standard_normal((64, 10))
    a = orc.fft(x, axis=0)
    b = orc.fft(x.T).T
    assert np.array_equal(a, b)
    assert rel(a, np.fft.fft(x, axis=0)) < 1e-13


def test_fft_teeth():
    # a wrong twiddle sign is caught by the same comparison
    rng = np.random.default_rng(11)
    x = rng.standard_normal(64) + 1j * rng.standard_normal(64)
    wrong = np.fft.ifft(x) * 64
    assert rel(wrong, np.fft.fft(x)) > 0.5


# --------------------------------------------------------------- window
@pytest.mark.parametrize("n", [2, 5, 126, 1000])
def test_hamming(orc, n):
    assert np.allclose(orc.hamming(n), np.hamming(n), atol=1e-15)
    assert np.allclose(orc.hamming(n), signal.windows.hamming(n), atol=1e-15)


# ----------------------------------------------------------------- conv
def test_conv1d_kat(orc):
    k = KAT["conv1d"]
    for mode in ("full", "same", "valid"):
        out = orc.conv1d(np.array(k["in"], float), np.array(k["filter"], float), mode)
        assert np.allclose(out.real, k[mode], atol=1e-12), mode
        assert np.allclose(out.imag, 0)


def test_conv1d_batched_formula_and_numpy(orc):
    k = KAT["conv1d_batched_strided"]
    B, n, fl = k["batches"], k["signal_len"], k["filter_len"]
    i = np.arange(n)
    f = np.arange(fl)
    for b in range(B):
        x = ((i % 17) - 8) * 0.05 + (b + 1) * 0.125
        h = ((f % 9) - 4) * 0.025 + (b + 1) * 0.01
        for mode in ("full", "same", "valid"):
            out = orc.conv1d(x, h, mode).real
            start = {"full": 0, "same": (fl - 1) // 2, "valid": fl - 1}[mode]
            exp = np.array([sum(x[o + start - t] * h[t] for t in range(fl)
                                if 0 <= o + start - t < n) for o in range(out.size)])
            assert np.allclose(out, exp, atol=1e-12), mode
            assert np.allclose(out, np.convolve(x, h, mode), atol=1e-12), mode


def test_conv1d_complex_even_filter(orc):
    rng = np.random.default_rng(2)
    x = rng.standard_normal((3, 50)) + 1j * rng.standard_normal((3, 50))
    h = rng.standard_normal(6) + 1j * rng.standard_normal(6)
    full = np.stack([np.convolve(r, h, "full") for r in x])
    assert rel(orc.conv1d(x, h, "full"), full) < 1e-13
    # SAME start for even filter is L/2-1 (conv.cuh:159-163)
    assert rel(orc.conv1d(x, h, "same"), full[:, 2:52]) < 1e-13
    assert rel(orc.conv1d(x, h, "valid"), full[:, 5:50]) < 1e-13


def test_conv_teeth():
    k = KAT["conv1d"]
    # correlation instead of convolution (un-flipped filter) fails the KAT
    wrong = np.correlate(np.array(k["in"], float), np.array(k["filter"], float), "full")
    assert not np.allclose(wrong, k["full"])


def test_conv2d_kat(orc):
    k = KAT["conv2d"]
    for mode in ("full", "same", "valid"):
        out = orc.conv2d(np.array(k["in"], float), np.array(k["filter"], float), mode)
        assert np.allclose(out.real, k[mode], atol=1e-12), mode


def test_conv2d_scipy(orc):
    rng = np.random.default_rng(4)
    x = rng.standard_normal((2, 20, 31)) + 1j * rng.standard_normal((2, 20, 31))
    f = rng.standard_normal((5, 13)) + 1j * rng.standard_normal((5, 13))
    for mode in ("full", "same", "valid"):
        exp = np.stack([signal.convolve2d(p, f, mode) for p in x])
        assert rel(orc.conv2d(x, f, mode), exp) < 1e-13, mode


# ------------------------------------------------------------------ stages
def test_tpc_modes(orc):
    rng = np.random.default_rng(6)
    y = rng.standard_normal((2, 12, 9)) + 1j * rng.standard_normal((2, 12, 9))
    k = np.array([[1.0], [-2.0], [1.0]])
    for mode in ("same", "valid"):
        exp = np.stack([signal.convolve2d(p, k, mode) for p in y])
        assert rel(orc.tpc(y, mode), exp) < 1e-14
    # closed form t[p] = y[p-1] - 2y[p] + y[p+1]
    t = orc.tpc(y, "same")
    assert np.allclose(t[:, 5], y[:, 4] - 2 * y[:, 5] + y[:, 6])


def test_doppler_window_rows(orc):
    rng = np.random.default_rng(8)
    t = rng.standard_normal((1, 10, 4)) + 1j * rng.standard_normal((1, 10, 4))
    D = orc.doppler(t, 8, 16)
    tw = t[0].copy()
    tw[:8] *= np.hamming(8)[:, None]          # rows 8, 9 unwindowed
    assert rel(D[0], np.fft.fft(tw, 16, 0)) < 1e-13


def test_cfar_norm_closed_form(orc):
    mask = orc.MATX_CFAR_MASK
    assert mask.sum() == 56
    Ny, Nx = 20, 40
    r = orc.cfar(np.ones((1, Ny, Nx)), mask)
    nm = r["norm"][0]
    assert np.all(nm[2:-2, 6:-6] == 56)

    # rectangle-intersection count: outer 5x13 box minus 3x3 hole, clipped
    def cnt(y, x):
        oy = min(y + 2, Ny - 1) - max(y - 2, 0) + 1
        ox = min(x + 6, Nx - 1) - max(x - 6, 0) + 1
        hy = min(y + 1, Ny - 1) - max(y - 1, 0) + 1
        hx = min(x + 1, Nx - 1) - max(x - 1, 0) + 1
        return oy * ox - hy * hx
    exp = np.array([[cnt(y, x) for x in range(Nx)] for y in range(Ny)])
    assert np.array_equal(nm, exp)
    assert nm[0, 0] == 3 * 7 - 2 * 2


def test_cfar_vs_scipy(orc):
    rng = np.random.default_rng(9)
    pw = rng.exponential(1.0, (2, 32, 50))
    pw[0, 10, 20] = 400.0
    r = orc.cfar(pw)
    mask = orc.MATX_CFAR_MASK
    for c in range(2):
        nm = signal.convolve2d(np.ones(pw[c].shape), mask, "same")
        ba = signal.convolve2d(pw[c], mask, "same") / nm
        alpha = nm * (np.power(1e-5, -1.0 / nm) - 1)
        assert np.allclose(r["ba"][c], ba, rtol=1e-13)
        assert np.array_equal(r["dets"][c], (pw[c] > alpha * ba).astype(np.int32))
    assert r["dets"][0, 10, 20] == 1


def test_cfar_teeth():
    # using the un-normalised sum (dropping /norm) changes detections
    rng = np.random.default_rng(10)
    pw = rng.exponential(1.0, (32, 50))
    pw[10, 20] = 60.0
    mask = np.ones((5, 13)); mask[1:4, 5:8] = 0
    nm = signal.convolve2d(np.ones(pw.shape), mask, "same")
    s = signal.convolve2d(pw, mask, "same")
    alpha = nm * (np.power(1e-5, -1.0 / nm) - 1)
    good = pw > alpha * s / nm
    wrong = pw > alpha * s
    assert good.sum() != wrong.sum()


# ------------------------------------------------- fft_conv / pulse compress
def test_fft_conv_is_circular_correlation(orc):
    rng = np.random.default_rng(12)
    N, S, L = 64, 40, 9
    x = rng.standard_normal((2, S)) + 1j * rng.standard_normal((2, S))
    w = rng.standard_normal(L) + 1j * rng.standard_normal(L)
    y = orc.fft_conv(x, w, N)
    xp = np.zeros((2, N), complex); xp[:, :S] = x
    exp = np.array([[sum(xp[b, (n + l) % N] * np.conj(w[l]) for l in range(L))
                     for n in range(N)] for b in range(2)])
    assert rel(y, exp) < 1e-13
    assert rel(y, np.fft.ifft(np.fft.fft(x, N) * np.conj(np.fft.fft(w, N)))) < 1e-13


def test_pulse_compress_linear_part(orc):
    # the first R = S-L+1 lags are the linear matched filter (no wrap)
    rng = np.random.default_rng(13)
    S, L, N = 300, 40, 512
    x = rng.standard_normal((3, S)) + 1j * rng.standard_normal((3, S))
    w = rng.standard_normal(L) + 1j * rng.standard_normal(L)
    Wc, w2 = orc.waveform_spectrum(w, N, True, True, return_prepared=True)
    hw = w * np.hamming(L)
    assert rel(w2, hw / np.linalg.norm(hw)) < 1e-14
    R = S - L + 1
    y = orc.pulse_compress(x, Wc, R)
    exp = np.stack([np.correlate(r, w2, "valid") for r in x])
    assert rel(y, exp) < 1e-13


# ------------------------------------------- end-to-end vs reference invariant
def test_model_chain_17_detections(orc, matx_inputs, golden):
    x, w = matx_inputs
    st = orc.radar_pipeline(x[None], w, Nd=256, mode="model", stages=True)
    assert int(st["dets"].sum()) == 17          # MultiChannelRadarPipeline.cu:197
    assert np.array_equal(np.argwhere(st["dets"][0] == 1), golden["model_det_idx"])
    yi, ti, di = golden["model_y_idx"], golden["model_t_idx"], golden["model_D_idx"]
    assert rel(st["y"][0][yi[:, 0], yi[:, 1]], golden["model_y_val"]) < 1e-12
    assert rel(st["tpc"][0][ti[:, 0], ti[:, 1]], golden["model_t_val"]) < 1e-12
    assert rel(st["D"][0][di[:, 0], di[:, 1]], golden["model_D_val"]) < 1e-12
    assert rel(st["ba"][0][di[:, 0], di[:, 1]], golden["model_ba_val"]) < 1e-12
    assert rel(st["Wc"][:64], golden["W_head"]) < 1e-12
    assert np.array_equal(st["norm"][0][:3], golden["norm_rows"])
    assert np.array_equal(st["norm"][0][-3:], golden["norm_last_rows"])


def test_matx_chain_matches_restatement(orc, matx_inputs, golden):
    x, w = matx_inputs
    st = orc.radar_pipeline(x[None], w, mode="matx", stages=True)
    assert st["D"].shape == (1, 256, 8001)      # numPulsesRnd = 256
    assert np.array_equal(np.argwhere(st["dets"][0] == 1), golden["matx_det_idx"])
    ti, di = golden["matx_t_idx"], golden["matx_D_idx"]
    assert rel(st["tpc"][0][ti[:, 0], ti[:, 1]], golden["matx_t_val"]) < 1e-12
    assert rel(st["D"][0][di[:, 0], di[:, 1]], golden["matx_D_val"]) < 1e-12
    assert rel(st["ba"][0][di[:, 0], di[:, 1]], golden["matx_ba_val"]) < 1e-12
