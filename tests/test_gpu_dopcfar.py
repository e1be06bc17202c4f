"""GPU parity of the fused pulse-axis + CFAR entry (mxr_doppler_cfar) against
the FP64 oracle chain canceller -> window -> Doppler FFT -> |.|^2 -> CA-CFAR,
and against the unfused GPU entries (mxr_doppler then mxr_cfar).

Tolerances as everywhere: relative L2 <= 1e-5 for floating-point outputs;
detection masks identical except at cells whose oracle margin is below 1e-5.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

RTOL = 1e-5
MARGIN = 1e-5


def rel(a, b):
    a = np.asarray(a, dtype=np.complex128)
    b = np.asarray(b, dtype=np.complex128)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def dev(a, dtype=torch.complex64):
    return torch.as_tensor(np.ascontiguousarray(a)).to(dtype).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


@pytest.fixture(scope="module")
def mx():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import matx_amd
    matx_amd.lib()
    return matx_amd


def rect_mask(win, guard):
    m = np.ones(win)
    cy, cx, gy, gx = win[0] // 2, win[1] // 2, guard[0] // 2, guard[1] // 2
    m[cy - gy:cy + gy + 1, cx - gx:cx + gx + 1] = 0
    return m


def cube_with_tones(rng, C, P, R, nd, ntones=12):
    """Unit complex noise plus strong Doppler tones at random range bins."""
    y = (rng.standard_normal((C, P, R)) + 1j * rng.standard_normal((C, P, R))) / np.sqrt(2)
    p = np.arange(P)
    for _ in range(ntones):
        c, r, kd = rng.integers(C), rng.integers(R), rng.integers(nd)
        y[c, :, r] += rng.uniform(3, 300) * np.exp(2j * np.pi * kd * p / nd)
    return y.astype(np.complex64)


def oracle_chain(orc, y, mode, win_rows, nd, win, guard, pfa):
    t = y if mode == "none" else orc.tpc(y, mode)
    D = orc.doppler(t, win_rows, nd)
    pw = orc.abs2(D)
    return pw, orc.cfar(pw, rect_mask(win, guard), pfa)


def dets_bad(got, ref):
    near = np.abs(ref["margin"]) < MARGIN
    return int(((np.asarray(got) != ref["dets"]) & ~near).sum())


CASES = [
    # mode, nd, P, R, window, guard           (fused unless noted)
    ("valid", 128, 128, 1001, (5, 13), (3, 5)),
    ("valid", 128, 128, 777, (5, 13), (3, 3)),
    ("same", 64, 64, 500, (5, 13), (3, 3)),
    ("none", 32, 30, 203, (3, 5), (1, 3)),
    ("same", 16, 16, 130, (5, 9), (3, 3)),
    ("valid", 8, 10, 47, (5, 13), (3, 5)),
    ("valid", 256, 128, 300, (5, 13), (3, 5)),  # nd > 128: unfused path
    ("valid", 128, 128, 300, (7, 21), (3, 5)),  # no fused instance: unfused path
]


@pytest.mark.parametrize("mode,nd,P,R,win,guard", CASES)
def test_doppler_cfar_vs_oracle(mx, orc, mode, nd, P, R, win, guard):
    rng = np.random.default_rng(nd * 1000 + R)
    C, pfa = 2, 1e-5
    y = cube_with_tones(rng, C, P, R, nd)
    win_rows = P if mode == "none" else P - 2
    pw_ref, ref = oracle_chain(orc, y, mode, win_rows, nd, win, guard, pfa)
    assert ref["dets"].sum() > 0
    yd = dev(y)
    for want_ba, want_pow in ((True, True), (False, False), (False, True)):
        dets, ba, pw = mx.doppler_cfar(yd, nd, tpc=mode, win_rows=win_rows, win=win, guard=guard,
                                       pfa=pfa, want_ba=want_ba, want_pow=want_pow)
        assert dets_bad(host(dets), ref) == 0, (want_ba, want_pow)
        if want_ba:
            assert rel(host(ba), ref["ba"]) < RTOL
        if want_pow:
            assert rel(host(pw), pw_ref) < RTOL


def test_doppler_cfar_equals_unfused(mx, orc):
    rng = np.random.default_rng(11)
    C, P, R, nd = 3, 128, 2053, 128
    y = dev(cube_with_tones(rng, C, P, R, nd, 40))
    d1, b1, p1 = mx.doppler_cfar(y, nd, tpc="valid", guard=(3, 5), want_ba=True, want_pow=True)
    p2, _ = mx.doppler(y, nd, tpc="valid")
    d2, b2, _ = mx.cfar(p2, guard=(3, 5), want_ba=True)
    P1, P2 = host(p1), host(p2)
    # the same transform with the canceller and window applied in a different
    # order: agreement at rounding level (cells cancelled to near zero differ
    # relatively, not absolutely)
    assert rel(P1, P2) < 2e-6
    assert rel(host(b1), host(b2)) < 2e-6
    D1, D2 = host(d1), host(d2)
    assert D1.sum() > 0
    # a flip needs a cell within rounding of its threshold: rare
    assert int((D1 != D2).sum()) <= max(2, D1.size // 100000)


def test_doppler_cfar_strided_unaligned_outputs(mx, orc):
    rng = np.random.default_rng(12)
    C, P, R, nd = 2, 64, 301, 64
    y = cube_with_tones(rng, C, P, R, nd)
    pw_ref, ref = oracle_chain(orc, y, "valid", P - 2, nd, (5, 13), (3, 3), 1e-5)
    big_d = torch.full((C, nd, R + 3), -7, dtype=torch.int32, device="cuda")
    big_b = torch.full((C, nd, R + 5), -7.0, dtype=torch.float32, device="cuda")
    dets = big_d[:, :, 1:R + 1]      # 4-byte offset, odd row stride: scalar stores
    ba = big_b[:, :, 2:R + 2]
    mx.doppler_cfar(dev(y), nd, tpc="valid", dets=dets, ba=ba)
    assert dets_bad(host(dets), ref) == 0
    assert rel(host(ba), ref["ba"]) < RTOL
    assert np.all(host(big_d)[:, :, 0] == -7) and np.all(host(big_d)[:, :, R + 1:] == -7)
    assert np.all(host(big_b)[:, :, :2] == -7) and np.all(host(big_b)[:, :, R + 2:] == -7)


def test_fused_pipeline_full_range(mx, orc):
    """One channel of the 128 x 9000 pulse, L = 1000 workload end to end:
    pulse compression + fused pulse-axis/CFAR against the oracle's model
    chain (VALID canceller, 126-row window, 128-point Doppler)."""
    rng = np.random.default_rng(13)
    C, P, S, L, nd = 1, 128, 9000, 1000, 128
    x = ((rng.standard_normal((C, P, S)) + 1j * rng.standard_normal((C, P, S))) / np.sqrt(2))
    t = np.arange(L)
    w = np.exp(1j * np.pi * 0.4 * (t - L / 2) ** 2 / L)
    for _ in range(6):
        r, kd = rng.integers(S - L), rng.integers(nd)
        x[0, :, r:r + L] += rng.uniform(0.05, 0.5) * w[None, :] * np.exp(
            2j * np.pi * kd * np.arange(P) / nd)[:, None]
    x = x.astype(np.complex64)
    w = w.astype(np.complex64)
    st = orc.radar_pipeline(x, w, nd, mode="model", mask=rect_mask((5, 13), (3, 5)),
                            stages=True)
    y = mx.pulse_compress(dev(x), dev(w), n=16384)
    dets, ba, pw = mx.doppler_cfar(y, nd, tpc="valid", guard=(3, 5), want_ba=True, want_pow=True)
    assert rel(host(pw), st["pow"]) < RTOL
    assert rel(host(ba), st["ba"]) < RTOL
    assert dets_bad(host(dets), {"dets": st["dets"], "margin": st["margin"]}) == 0
    assert host(dets).sum() > 0


def test_pulse_compression_plan(mx, orc):
    """The plan (spectrum once, block kernel per cube) gives the per-call
    entry's result and the oracle's, across two waveforms and batch shapes."""
    rng = np.random.default_rng(14)
    S, L, n = 9000, 1000, 16384
    R = S - L + 1
    plan = mx.ops.PulseCompressionPlan(S, L, n, R)
    for shape in ((3, 5), (7,)):
        x = ((rng.standard_normal(shape + (S,)) + 1j * rng.standard_normal(shape + (S,)))
             .astype(np.complex64))
        w = (rng.standard_normal(L) + 1j * rng.standard_normal(L)).astype(np.complex64)
        xd, wd = dev(x), dev(w)
        plan.set_waveform(wd)
        y_plan = host(plan(xd))
        y_call = host(mx.pulse_compress(xd, wd, n=n))
        assert np.array_equal(y_plan, y_call)
        Wc = orc.waveform_spectrum(w, n, True, True)
        ref = orc.pulse_compress(x.reshape(-1, S), Wc, R).reshape(shape + (R,))
        assert rel(y_plan, ref) < RTOL
    with pytest.raises(mx.MatxError):
        plan(dev(np.zeros((2, S - 1), np.complex64)))   # wrong row length
    plan.close()
