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
