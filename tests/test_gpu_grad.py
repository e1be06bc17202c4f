"""GPU parity of the differentiable path: per-row fft_conv and its adjoint,
the tapped Doppler stage and its adjoint (FP64 numpy formulas, rel L2 <=
1e-5), and the waveform/window design-loss gradients from torch.autograd
over the HIP kernels against the FP64 oracle gradient (rel <= 1e-4; the
oracle gradient itself is pinned to FP64 central differences in
tests/test_grad_oracle.py)."""
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
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)


@pytest.fixture(scope="module")
def mx():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import matx_amd
    matx_amd.lib()
    return matx_amd


def np_corr(x, w, n, conj=True):
    X = np.fft.fft(x.astype(np.complex128), n, axis=-1)
    W = np.fft.fft(w.astype(np.complex128), n, axis=-1)
    return np.fft.ifft(X * (np.conj(W) if conj else W), axis=-1)


# the last four transform at the rows' support (a smaller power of two than
# n) and write the other lags as zeros: convolution, a kernel row longer
# than the signal row, fewer kept lags than n
@pytest.mark.parametrize("B,S,L,n,nout,conj", [
    (5, 1000, 1000, 2048, 2048, True), (3, 700, 300, 1024, 1000, True),
    (4, 2048, 256, 4096, 4096, False), (2, 1024, 1024, 8192, 8192, True),
    (3, 500, 200, 4096, 4096, False), (4, 100, 900, 4096, 4096, True),
    (2, 600, 300, 8192, 5000, True), (2, 1500, 500, 8192, 8192, False)])
def test_fft_conv_rows(mx, B, S, L, n, nout, conj):
    rng = np.random.default_rng(n + S)
    x = crandn(rng, B, S)
    w = crandn(rng, B, L)
    out = torch.empty((B, nout), dtype=torch.complex64, device="cuda")
    mx.fft_conv(dev(x), dev(w), n, conj=conj, out=out)
    assert rel(host(out), np_corr(x, w, n, conj)[:, :nout]) < RTOL


def test_fft_conv_rows_autocorrelation(mx):
    rng = np.random.default_rng(1)
    s = crandn(rng, 6, 1024)
    sd = dev(s)
    got = host(mx.fft_conv(sd, sd, 8192))
    assert rel(got, np_corr(s, s, 8192)) < RTOL


@pytest.mark.parametrize("conj", [True, False])
def test_fft_conv_adjoint(mx, conj):
    rng = np.random.default_rng(2 + conj)
    B, S, L, n, nout = 4, 900, 333, 2048, 2048
    x, w, gy = crandn(rng, B, S), crandn(rng, B, L), crandn(rng, B, nout)
    gx, gw = mx.ops.fft_conv_adjoint(dev(gy), dev(x), dev(w), n, conj)
    X = np.fft.fft(x.astype(complex), n, axis=-1)
    W = np.fft.fft(w.astype(complex), n, axis=-1)
    GY = np.fft.fft(gy.astype(complex), n, axis=-1)
    if conj:
        rx, rw = np.fft.ifft(GY * W)[:, :S], np.fft.ifft(X * np.conj(GY))[:, :L]
    else:
        rx, rw = np.fft.ifft(GY * np.conj(W))[:, :S], np.fft.ifft(GY * np.conj(X))[:, :L]
    assert rel(host(gx), rx) < RTOL
    assert rel(host(gw), rw) < RTOL
    # adjoint identity <A x, gy> = <x, A^H gy> on the forward operator
    y = np_corr(x, w, n, conj)[:, :nout]
    lhs = np.vdot(gy.astype(complex), y)
    rhs = np.vdot(host(gx).astype(complex), x)
    assert abs(lhs - rhs) / abs(lhs) < 1e-4


@pytest.mark.parametrize("C,P,R,nd", [(2, 128, 300, 128), (3, 20, 77, 32), (1, 200, 50, 256)])
def test_doppler_taps_and_adjoint(mx, C, P, R, nd):
    rng = np.random.default_rng(C * P + R)
    x = crandn(rng, C, P, R)
    h = (np.hamming(P) + 0.1 * rng.standard_normal(P)).astype(np.float32)
    gp = rng.standard_normal((C, nd, R)).astype(np.float32)
    pw, rd = mx.ops.doppler_taps(dev(x), dev(h, torch.float32), nd, want_rd=True)
    D = np.fft.fft(h[None, :, None].astype(float) * x.astype(complex), nd, axis=1)
    assert rel(host(rd), D) < RTOL
    assert rel(host(pw), np.abs(D) ** 2) < RTOL
    gx, gt = mx.ops.doppler_taps_adjoint(dev(gp, torch.float32), dev(x), dev(h, torch.float32),
                                         nd, want_gx=True)
    gu = nd * np.fft.ifft(2 * gp * D, axis=1)[:, :P, :]
    assert rel(host(gx), h[None, :, None] * gu) < RTOL
    assert rel(host(gt), np.sum(np.real(np.conj(gu) * x), axis=(0, 2))) < RTOL


@pytest.mark.parametrize("size", ["small", "design"])
def test_design_loss_gradient(mx, size):
    from matx_amd import autograd as ag
    from oracle import grad as og
    if size == "small":
        phi, h, X, targets = ag.make_design_problem(5, B=4, Lw=256, C=2, P=32, R=100, T=4)
        n, nd = 1024, 32
    else:
        phi, h, X, targets = ag.make_design_problem(6)
        n, nd = 8192, 128
    L64, gphi64, gh64 = og.loss_and_grad(phi, h, X, targets, n, nd)
    p = dev(phi, torch.float32).requires_grad_()
    t = dev(h, torch.float32).requires_grad_()
    tg = torch.as_tensor(np.array(targets), dtype=torch.int64, device="cuda")
    L = ag.design_loss(p, t, dev(X), tg, n, nd)
    L.backward()
    assert abs(float(L.detach()) - L64) / abs(L64) < 1e-5
    assert rel(host(p.grad), gphi64) < GTOL
    assert rel(host(t.grad), gh64) < GTOL


def test_autocorr_adjoint_is_sum_of_both_terms(mx):
    rng = np.random.default_rng(9)
    B, S, n = 5, 1000, 4096
    s, gy = crandn(rng, B, S), crandn(rng, B, n)
    gs = mx.ops.autocorr_adjoint(dev(gy), dev(s), n)
    Sf = np.fft.fft(s.astype(complex), n, axis=-1)
    GY = np.fft.fft(gy.astype(complex), n, axis=-1)
    ref = (np.fft.ifft(GY * Sf) + np.fft.ifft(Sf * np.conj(GY)))[:, :S]
    assert rel(host(gs), ref) < RTOL


def _phase_isl_fp64(phi, n, g):
    """ISL of the n-point autocorrelation of exp(j phi) per row and its
    gradient w.r.t. phi for upstream g, in FP64 (closed forms: r = IFFT(|S|^2),
    dISL/dr = 2 r / P0 off lag 0 and -2 (S - P0) / P0^2 r at lag 0, the
    autocorrelation's adjoint IFFT(S .* 2Re(FFT(gr))), then the chain rule
    through s = exp(j phi))."""
    s = np.exp(1j * phi.astype(np.float64))
    S = np.fft.fft(s, n, axis=-1)
    r = np.fft.ifft(np.abs(S) ** 2, axis=-1)
    pr = np.abs(r) ** 2
    tot, p0 = pr.sum(-1), pr[:, 0]
    isl = (tot - p0) / p0
    gr = (2.0 * g / p0)[:, None] * r
    gr[:, 0] = -2.0 * g * (tot - p0) / p0 ** 2 * r[:, 0]
    gs = np.fft.ifft(S * 2.0 * np.real(np.fft.fft(gr, axis=-1)), axis=-1)[:, :phi.shape[1]]
    gphi = np.cos(phi) * gs.imag - np.sin(phi) * gs.real
    return isl, gphi


# the transform runs at the support (2048 for L = 1024 or 700) or at n
@pytest.mark.parametrize("B,L,n", [(4, 1024, 8192), (5, 700, 4096), (3, 300, 1024), (2, 3000, 8192)])
def test_phase_isl_head(mx, B, L, n):
    from matx_amd import autograd as ag
    rng = np.random.default_rng(L + n)
    phi = rng.uniform(0, 2 * np.pi, (B, L)).astype(np.float32)
    g = rng.standard_normal(B).astype(np.float32)
    isl64, gphi64 = _phase_isl_fp64(phi, n, g.astype(np.float64))
    pd = dev(phi, torch.float32)
    assert rel(host(mx.ops.phase_isl(pd, n)), isl64) < RTOL
    assert rel(host(mx.ops.phase_isl_adjoint(dev(g, torch.float32), pd, n)), gphi64) < GTOL
    # the same value and gradient as the composition it replaces
    p1 = pd.clone().requires_grad_()
    p2 = pd.clone().requires_grad_()
    gt = dev(g, torch.float32)
    (ag.PhaseIsl.apply(p1, n) * gt).sum().backward()
    s = torch.polar(torch.ones_like(p2), p2)
    (ag.Isl.apply(ag.autocorr(s, n)) * gt).sum().backward()
    assert rel(host(p1.grad), host(p2.grad)) < GTOL


def test_phase_isl_errors(mx):
    phi = torch.zeros((2, 100), device="cuda")
    with pytest.raises(Exception, match="power of two"):
        mx.ops.phase_isl(phi, 3000)
    with pytest.raises(Exception, match="L <= n"):
        mx.ops.phase_isl(torch.zeros((2, 2000), device="cuda"), 1024)


# shared columns (two targets on one column), a target outside the map
# (NaN, no gradient), fewer pulses than Doppler bins
@pytest.mark.parametrize("C,P,R,nd,T", [(2, 32, 50, 32, 5), (3, 128, 200, 128, 24), (1, 20, 30, 64, 3)])
def test_psl_taps_adjoint_matches_dense(mx, C, P, R, nd, T):
    from matx_amd import autograd as ag
    rng = np.random.default_rng(C * 100 + P + T)
    x = crandn(rng, C, P, R)
    h = (np.hamming(P) + 0.1 * rng.standard_normal(P)).astype(np.float32)
    tg = np.stack([rng.integers(0, C, T), rng.integers(0, R, T), rng.integers(0, nd, T)], 1).astype(np.int32)
    tg[-1, :2] = tg[0, :2]          # a column shared by two targets
    if T > 3:
        tg[1, 0] = C                # outside the map
    g = rng.standard_normal(T).astype(np.float32)
    xd, hd, td, gd = dev(x), dev(h, torch.float32), torch.as_tensor(tg).cuda(), dev(g, torch.float32)
    pw, _ = mx.ops.doppler_taps(xd, hd, nd)
    gt = host(mx.ops.psl_taps_adjoint(gd, pw, xd, hd, td, 2, 8))
    gp = mx.ops.psl_adjoint(gd, pw, td, 2, 8)
    _, gt_dense = mx.ops.doppler_taps_adjoint(gp, xd, hd, nd)
    assert np.all(np.isfinite(gt))
    assert rel(gt, host(gt_dense)) < GTOL
    # the autograd function: sparse backward (x without gradient) and the
    # dense fallback (x with gradient) against the unfused composition
    for x_grad in (False, True):
        x1 = xd.clone().requires_grad_(x_grad)
        x2 = xd.clone().requires_grad_(x_grad)
        h1 = hd.clone().requires_grad_()
        h2 = hd.clone().requires_grad_()
        valid = torch.as_tensor(tg[:, 0] < C).cuda()
        y1 = ag.DopplerPsl.apply(x1, h1, td, nd, 2, 8)
        y2 = ag.Psl.apply(ag.doppler_power(x2, h2, nd), td, 2, 8)
        assert rel(host(y1[valid]), host(y2[valid])) < RTOL
        (y1[valid] * gd[valid]).sum().backward()
        (y2[valid] * gd[valid]).sum().backward()
        assert rel(host(h1.grad), host(h2.grad)) < GTOL
        if x_grad:
            assert rel(host(x1.grad), host(x2.grad)) < GTOL


def test_isl_head_and_adjoint(mx):
    rng = np.random.default_rng(31)
    B, n = 5, 4096
    r = crandn(rng, B, n)
    g = rng.standard_normal(B).astype(np.float32)
    isl = host(mx.ops.isl(dev(r)))
    pr = np.abs(r.astype(complex)) ** 2
    ref = (pr.sum(-1) - pr[:, 0]) / pr[:, 0]
    assert rel(isl, ref) < RTOL
    gr = host(mx.ops.isl_adjoint(dev(g, torch.float32), dev(r)))
    gpr = np.broadcast_to(g[:, None] / pr[:, :1], pr.shape).copy()
    gpr[:, 0] = -g * (pr.sum(-1) - pr[:, 0]) / pr[:, 0] ** 2
    assert rel(gr, 2 * gpr * r) < RTOL


def test_psl_head_and_adjoint(mx):
    """Forward and adjoint against the FP64 formulas, with two targets on
    one column (their gradients add) and one outside the map (NaN, no
    gradient)."""
    rng = np.random.default_rng(32)
    C, nd, R, ml, q = 3, 64, 50, 2, 8
    pw = (rng.exponential(1.0, (C, nd, R)) + 0.1).astype(np.float32)
    tg = np.array([[0, 7, 5], [2, 49, 60], [0, 7, 30], [1, 0, 0], [3, 1, 1]], dtype=np.int32)
    g = rng.standard_normal(len(tg)).astype(np.float32)
    psl = host(mx.ops.psl(dev(pw, torch.float32), dev(tg, torch.int32), ml, q))
    k = np.arange(nd)
    ref, gref = [], np.zeros((C, nd, R))
    for t, (c, r, kd) in enumerate(tg[:-1]):
        col = pw[c, :, r].astype(float)
        d = np.minimum((k - kd) % nd, (kd - k) % nd)
        m = d > ml
        A = np.sum(col[m] ** q) ** (1 / q)
        ref.append(A / col[kd])
        gg = np.zeros(nd)
        gg[m] = A ** (1 - q) * col[m] ** (q - 1) / col[kd]
        gg[kd] = -A / col[kd] ** 2
        gref[c, :, r] += g[t] * gg
    assert rel(psl[:-1], ref) < RTOL and np.isnan(psl[-1])
    gp = host(mx.ops.psl_adjoint(dev(g, torch.float32), dev(pw, torch.float32),
                                 dev(tg, torch.int32), ml, q))
    assert rel(gp, gref) < RTOL
