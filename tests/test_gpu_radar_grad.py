"""GPU parity of the whole-cube waveform gradient: the pulse-axis adjoint
(mxr_doppler_adjoint) and the pulse-compression waveform gradient
(mxr_pc_plan_waveform_grad) against FP64 numpy formulas (rel L2 <= 1e-5),
and torch.autograd through both against the FP64 oracle gradient of
oracle/grad.py (rel <= 1e-4; that gradient is pinned to FP64 central
differences in tests/test_grad_oracle.py).  At sizes the oracle cannot reach
in a test, the gradient of the cube is checked against the sum of its
per-channel gradients."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

RTOL = 1e-5
GTOL = 1e-4


def rel(a, b):
    a = np.asarray(a, dtype=np.complex128)
    b = np.asarray(b, dtype=np.complex128)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def dev(a, dtype=torch.complex64):
    return torch.as_tensor(np.ascontiguousarray(a)).to(dtype).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def crandn(rng, *shape):
    return ((rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) /
            np.sqrt(2)).astype(np.complex64)


@pytest.fixture(scope="module")
def mx():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import matx_amd
    matx_amd.lib()
    return matx_amd


def hamming(n):
    if n <= 1:
        return np.ones(n)
    return 0.54 - 0.46 * np.cos(2 * np.pi * np.arange(n) / (n - 1))


def np_doppler_adjoint(gp, y, nd, tpc, win_rows):
    """FP64: gy for pow = |FFT_nd(hw .* canceller(y))|^2."""
    y = y.astype(np.complex128)
    C, P, R = y.shape
    rows = P - 2 if tpc == "valid" else P
    hw = np.ones(rows)
    hw[:win_rows] = hamming(win_rows)
    if tpc == "none":
        t = y
    elif tpc == "same":
        yp = np.pad(y, ((0, 0), (1, 1), (0, 0)))
        t = yp[:, :-2] - 2 * yp[:, 1:-1] + yp[:, 2:]
    else:
        t = y[:, :-2] - 2 * y[:, 1:-1] + y[:, 2:]
    D = np.fft.fft(hw[None, :, None] * t, nd, axis=1)
    s = hw[None, :, None] * (nd * np.fft.ifft(2 * gp * D, axis=1)[:, :rows])
    if tpc == "none":
        return s
    o = 1 if tpc == "valid" else 0
    sp = np.zeros((C, rows + 2, R), dtype=np.complex128)
    sp[:, 1:-1] = s                                      # sp[:, i + 1] = s[:, i]
    gy = np.zeros_like(y)
    for p in range(P):
        for d, cf in ((1, 1.0), (0, -2.0), (-1, 1.0)):
            i = p - o + d
            if 0 <= i < rows:
                gy[:, p] += cf * sp[:, i + 1]
    return gy


@pytest.mark.parametrize("C,P,R,nd,tpc", [
    (2, 128, 300, 128, "same"), (2, 128, 300, 128, "valid"), (3, 20, 77, 32, "none"),
    (1, 130, 50, 128, "valid"), (2, 512, 40, 512, "valid"), (1, 64, 1000, 1024, "same"),
    (2, 6, 33, 8, "same")])
def test_doppler_adjoint(mx, C, P, R, nd, tpc):
    rng = np.random.default_rng(C * P + R + nd)
    y = crandn(rng, C, P, R)
    gp = rng.standard_normal((C, nd, R)).astype(np.float32)
    rows = P - 2 if tpc == "valid" else P
    win_rows = P - 2 if tpc != "none" else rows
    gy = mx.ops.doppler_adjoint(dev(gp, torch.float32), dev(y), nd, tpc=tpc)
    assert rel(host(gy), np_doppler_adjoint(gp, y, nd, tpc, win_rows)) < RTOL


def test_doppler_adjoint_reads_an_expanded_gradient(mx):
    """gpow given as a stride-0 view over channels (the loss head's form)."""
    rng = np.random.default_rng(11)
    C, P, R, nd = 4, 64, 200, 64
    y = crandn(rng, C, P, R)
    wmap = rng.standard_normal((nd, R)).astype(np.float32)
    gpe = dev(wmap, torch.float32).expand(C, nd, R)
    assert gpe.stride(0) == 0
    gy = mx.ops.doppler_adjoint(gpe, dev(y), nd, tpc="valid")
    ref = np_doppler_adjoint(np.broadcast_to(wmap, (C, nd, R)), y, nd, "valid", P - 2)
    assert rel(host(gy), ref) < RTOL


def test_doppler_adjoint_identity(mx):
    """pow = |A y|^2 with A linear (canceller, window, DFT); with gpow = 1/2
    the adjoint is A^H A y, so <y, gy> = ||A y||^2 = sum(pow) exactly."""
    rng = np.random.default_rng(12)
    C, P, R, nd = 2, 100, 64, 128
    y = crandn(rng, C, P, R)
    yd = dev(y)
    pw, _ = mx.ops.doppler(yd, nd, tpc="same")
    gp = torch.full((C, nd, R), 0.5, dtype=torch.float32, device="cuda")
    gy = host(mx.ops.doppler_adjoint(gp, yd, nd, tpc="same"))
    lhs = float(host(pw).astype(np.float64).sum())
    rhs = float(np.real(np.vdot(y.astype(np.complex128), gy.astype(np.complex128))))
    assert abs(lhs - rhs) / abs(lhs) < 1e-5


def np_waveform_grad(gy, x, w, n, window=True, normalize=True):
    x = x.astype(np.complex128).reshape(-1, x.shape[-1])
    gy = gy.astype(np.complex128).reshape(-1, gy.shape[-1])
    w = w.astype(np.complex128)
    L = w.size
    G = (np.fft.fft(x, n, axis=-1) * np.conj(np.fft.fft(gy, n, axis=-1))).sum(0)
    g2 = np.fft.ifft(G)[:L]
    h = hamming(L) if window else np.ones(L)
    w1 = h * w
    if normalize:
        nu = np.linalg.norm(w1)
        g2 = g2 / nu - w1 * np.real(np.vdot(g2, w1)) / nu ** 3
    return h * g2


@pytest.mark.parametrize("rows,S,L,n,flags", [
    ((2, 16), 9000, 1000, 16384, (True, True)),      # 4096-point blocks
    ((4, 32), 2048, 256, 4096, (True, True)),
    ((3, 5), 2000, 1500, 2048, (False, True)),       # one 2048-point block
    ((2, 8), 32768, 4096, 32768, (True, True)),      # 16384-point plan, 8192-point gradient
    ((1, 7), 5000, 700, 8192, (True, False)),
    ((16,), 1200, 64, 2048, (False, False))])
def test_pc_waveform_grad(mx, rows, S, L, n, flags):
    rng = np.random.default_rng(S + L)
    window, normalize = flags
    x = crandn(rng, *rows, S)
    w = crandn(rng, L)
    nout = S - L + 1
    gy = crandn(rng, *rows, nout)
    plan = mx.ops.PulseCompressionPlan(S, L, n, window=window, normalize=normalize)
    gw = plan.waveform_grad(dev(gy), dev(x), dev(w))
    assert rel(host(gw), np_waveform_grad(gy, x, w, n, window, normalize)) < RTOL
    # adjoint pairing with the forward: Re <y(w + e d), gy> is linear in the
    # filter, so Re<d w2, g2> = Re <dy, gy>; checked on the unnormalised filter
    if not normalize:
        d = crandn(rng, L)
        plan.set_waveform(dev(d))
        yd = host(plan(dev(x)))
        lhs = np.real(np.vdot(yd.astype(np.complex128), gy.astype(np.complex128)))
        g2 = np_waveform_grad(gy, x, w, n, window, False)
        h = hamming(L) if window else np.ones(L)
        rhs = np.real(np.vdot(h * d.astype(np.complex128), g2 / h))
        assert abs(lhs - rhs) / abs(lhs) < 1e-4
    plan.close()


def test_pc_waveform_grad_is_deterministic(mx):
    rng = np.random.default_rng(21)
    S, L, n = 9000, 1000, 16384
    x, gy, w = dev(crandn(rng, 64, S)), dev(crandn(rng, 64, S - L + 1)), dev(crandn(rng, L))
    plan = mx.ops.PulseCompressionPlan(S, L, n)
    a = host(plan.waveform_grad(gy, x, w))
    b = host(plan.waveform_grad(gy, x, w))
    assert np.array_equal(a, b)
    plan.close()


def radar_case(seed, C, P, S, L, nd, ntgt=6):
    """Unit complex noise plus ntgt echoes of the waveform exp(i phi) at
    seeded delays, each with a Doppler tone on bin kb of the nd-point
    transform; the target cells (kb, delay) are returned with the cube."""
    rng = np.random.default_rng(seed)
    phi = rng.uniform(0, 2 * np.pi, L).astype(np.float32)
    w = np.exp(1j * phi.astype(np.float64))
    X = crandn(rng, C, P, S)
    cells = []
    for _ in range(ntgt):
        c, d = int(rng.integers(C)), int(rng.integers(S - L + 1))
        kb = int(rng.integers(nd // 4, 3 * nd // 4))
        ramp = np.exp(2j * np.pi * kb * np.arange(P) / nd)
        X[c, :, d:d + L] += (0.5 * ramp[:, None] * w[None, :]).astype(np.complex64)
        cells.append((kb, d))
    return X, phi, cells


@pytest.mark.parametrize("C,P,S,L,nd,tpc", [
    (2, 32, 2048, 256, 32, "valid"), (3, 16, 3000, 700, 16, "same"),
    (2, 64, 9000, 1000, 64, "valid")])
def test_radar_gradient_vs_oracle(mx, C, P, S, L, nd, tpc):
    from matx_amd import autograd as ag
    from oracle import grad as og
    X, phi, cells = radar_case(C + P + S, C, P, S, L, nd)
    R = S - L + 1
    n = 1 << (S - 1).bit_length()
    wmap = og.contrast_weights(nd, R, cells, C)
    L64, g64 = og.radar_loss_and_grad(phi.astype(np.float64), X, wmap, n, nd, tpc)
    plan = mx.ops.PulseCompressionPlan(S, L, n)
    p = dev(phi, torch.float32).requires_grad_()
    w = torch.polar(torch.ones_like(p), p)
    pw = ag.radar_power(w, dev(X), plan, nd, tpc=tpc)
    loss = ag.weighted_power_sum(pw, dev(wmap, torch.float32))
    loss.backward()
    assert abs(float(loss.detach()) - L64) / abs(L64) < 1e-5
    assert rel(host(p.grad), g64) < GTOL
    plan.close()


def test_radar_gradient_sums_over_channels(mx):
    """Size-independent check: the gradient of the whole cube equals the
    sum of the per-channel gradients (the loss is a sum over channels)."""
    from matx_amd import autograd as ag
    C, P, S, L, nd = 4, 128, 8192, 1024, 128
    X, phi, cells = radar_case(77, C, P, S, L, nd)
    R = S - L + 1
    Xd = dev(X)
    wmap = ag.contrast_weights(nd, R, cells, C, device="cuda")
    plan = mx.ops.PulseCompressionPlan(S, L, 8192)

    def grad(x):
        p = dev(phi, torch.float32).requires_grad_()
        w = torch.polar(torch.ones_like(p), p)
        ag.weighted_power_sum(ag.radar_power(w, x, plan, nd), wmap).backward()
        return host(p.grad).astype(np.float64)

    g = grad(Xd)
    parts = sum(grad(Xd[c:c + 1]) for c in range(C))
    assert rel(g, parts) < 1e-5
    plan.close()
