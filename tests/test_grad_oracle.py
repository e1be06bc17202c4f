"""Pins the FP64 gradient oracle (oracle/grad.py): its analytic gradients of
the waveform/window design loss agree with FP64 central finite differences
(h = 1e-3), and plausible mistakes in the adjoint formulas do not."""
import numpy as np
import pytest

from oracle import grad as og


def small_problem(seed=0):
    rng = np.random.default_rng(seed)
    B, Lw, n = 3, 32, 128
    C, P, R, nd = 2, 16, 20, 16
    phi = rng.uniform(0, 2 * np.pi, (B, Lw))
    h = np.hamming(P) + 0.05 * rng.standard_normal(P)
    X = (rng.standard_normal((C, P, R)) + 1j * rng.standard_normal((C, P, R))) / np.sqrt(2)
    targets = [(0, 3, 5), (1, 11, 2), (1, 17, 14)]
    for c, r, kd in targets:
        X[c, :, r] += 20 * np.exp(2j * np.pi * kd * np.arange(P) / nd)
    return phi, h, X, targets, n, nd


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_gradient_matches_central_differences():
    phi, h, X, targets, n, nd = small_problem()
    L, g_phi, g_h = og.loss_and_grad(phi, h, X, targets, n, nd)
    assert np.isclose(L, og.loss(phi, h, X, targets, n, nd))
    rng = np.random.default_rng(1)
    idx = rng.choice(phi.size, 24, replace=False)
    fd_phi = og.central_fd(lambda p: og.loss(p, h, X, targets, n, nd), phi, idx)
    assert rel(g_phi.flat[idx], fd_phi) < 1e-4
    fd_h = og.central_fd(lambda t: og.loss(phi, t, X, targets, n, nd), h, range(h.size))
    assert rel(g_h, fd_h) < 1e-4


def test_isl_of_a_barker_code():
    # Barker-13 (phases 0 / pi): aperiodic autocorrelation sidelobes are 1 at
    # the 6 even lags and 0 at the odd ones on each side, so ISL = 12 / 13^2
    b13 = np.array([1, 1, 1, 1, 1, -1, -1, 1, 1, -1, 1, -1, 1])
    phi = np.where(b13 > 0, 0.0, np.pi)[None, :]
    isl, _, _ = og.isl_terms(phi, 64)
    assert np.isclose(isl[0], 12 / 169)


def test_doppler_power_matches_direct_dft():
    rng = np.random.default_rng(3)
    X = rng.standard_normal((1, 6, 4)) + 1j * rng.standard_normal((1, 6, 4))
    h = rng.standard_normal(6)
    Pw, _ = og.doppler_power(X, h, 8)
    k = np.arange(8)[:, None]
    p = np.arange(6)[None, :]
    F = np.exp(-2j * np.pi * k * p / 8)
    ref = np.abs(np.einsum("kp,cpr->ckr", F, h[None, :, None] * X)) ** 2
    assert np.allclose(Pw, ref)


@pytest.mark.parametrize("mistake", ["no_gw", "half_gr", "conj_gx"])
def test_planted_adjoint_mistakes_fail(mistake):
    """Each mistake rebuilds the phase gradient with one adjoint term wrong;
    the finite differences must reject it."""
    phi, h, X, targets, n, nd = small_problem(2)
    B, Lw = phi.shape
    s = og.waveforms(phi)
    _, r, pr = og.isl_terms(phi, n)
    g_pr = np.broadcast_to(1.0 / (B * pr[:, :1]), pr.shape).copy()
    g_pr[:, 0] = -(pr.sum(-1) - pr[:, 0]) / (B * pr[:, 0] ** 2)
    g_r = (1.0 if mistake == "half_gr" else 2.0) * g_pr * r
    S = np.fft.fft(s, n, axis=-1)
    GR = np.fft.fft(g_r, axis=-1)
    gx = np.fft.ifft(GR * (np.conj(S) if mistake == "conj_gx" else S), axis=-1)[:, :Lw]
    gw = np.fft.ifft(S * np.conj(GR), axis=-1)[:, :Lw]
    g_s = gx if mistake == "no_gw" else gx + gw
    g_phi = np.real(np.conj(g_s) * 1j * s)
    idx = np.arange(0, phi.size, 5)
    fd = og.central_fd(lambda p: og.loss(p, h, X, targets, n, nd), phi, idx)
    assert rel(g_phi.flat[idx], fd) > 1e-2


# ------------------------------------------------------ radar-chain gradient
def radar_problem(seed=0):
    rng = np.random.default_rng(seed)
    C, P, S, Lw, n, nd = 2, 8, 48, 12, 64, 8
    R = S - Lw + 1
    X = (rng.standard_normal((C, P, S)) + 1j * rng.standard_normal((C, P, S))) / np.sqrt(2)
    phi = rng.uniform(0, 2 * np.pi, Lw)
    wmap = rng.standard_normal((nd, R))
    return phi, X, wmap, n, nd


@pytest.mark.parametrize("tpc", ["valid", "same", "none"])
def test_radar_gradient_matches_central_differences(tpc):
    """dL/dphi (phase-only waveform) and dL/dw (free complex waveform, which
    also exercises the normalisation term that phase-only waveforms cancel)
    against FP64 central differences with h = 1e-3."""
    phi, X, wmap, n, nd = radar_problem(1)
    L, g_phi = og.radar_loss_and_grad(phi, X, wmap, n, nd, tpc)
    assert np.isclose(L, og.radar_loss(phi, X, wmap, n, nd, tpc))
    fd = og.central_fd(lambda p: og.radar_loss(p, X, wmap, n, nd, tpc), phi, range(phi.size))
    assert rel(g_phi, fd) < 1e-5
    rng = np.random.default_rng(2)
    w = og.waveforms(phi) * (1 + 0.3 * rng.standard_normal(phi.size))
    _, gw = og.radar_grad_w(w, X, wmap, n, nd, tpc)
    Lw = w.size
    fd = og.central_fd(lambda v: og.radar_loss_w(v[:Lw] + 1j * v[Lw:], X, wmap, n, nd, tpc),
                       np.concatenate([w.real, w.imag]), range(2 * Lw))
    assert rel(gw, fd[:Lw] + 1j * fd[Lw:]) < 1e-5


@pytest.mark.parametrize("tpc,mistake", [("valid", "no_norm"), ("same", "no_norm"),
                                         ("valid", "tpc_shift"), ("valid", "conj_corr"),
                                         ("none", "conj_corr"), ("same", "window_all")])
def test_radar_planted_mistakes_fail(tpc, mistake):
    phi, X, wmap, n, nd = radar_problem(3)
    rng = np.random.default_rng(4)
    w = og.waveforms(phi) * (1 + 0.3 * rng.standard_normal(phi.size))
    _, gw = og.radar_grad_w(w, X, wmap, n, nd, tpc, mistake=mistake)
    Lw = w.size
    fd = og.central_fd(lambda v: og.radar_loss_w(v[:Lw] + 1j * v[Lw:], X, wmap, n, nd, tpc),
                       np.concatenate([w.real, w.imag]), range(2 * Lw))
    assert rel(gw, fd[:Lw] + 1j * fd[Lw:]) > 1e-2


def test_radar_gradient_is_additive_over_channels():
    """The loss is a sum over channels, so the gradient of the whole cube is
    the sum of the per-channel gradients (the size-independent check the GPU
    path is held to at full size)."""
    phi, X, wmap, n, nd = radar_problem(5)
    _, g = og.radar_loss_and_grad(phi, X, wmap, n, nd)
    parts = [og.radar_loss_and_grad(phi, X[c:c + 1], wmap, n, nd)[1] for c in range(X.shape[0])]
    assert rel(sum(parts), g) < 1e-12


def test_contrast_weights_sum():
    wm = og.contrast_weights(8, 10, [(1, 2), (3, 4)], C=4)
    assert np.isclose(wm.sum(), 0.0)                      # mean clutter - mean target of 1s
    assert np.isclose(wm[1, 2], -1.0 / 8) and np.isclose(wm[0, 0], 1.0 / (4 * 78))
