"""GPU parity of the 64 x 512 x 32768 radar chain at its real row shapes on
one channel: S = 32768 samples per pulse, a 4096-sample LFM (plan n = 32768,
8192-point overlap-save blocks), the VALID three-pulse canceller (510 rows),
Hamming + 512-point Doppler FFT, |.|^2, CA-CFAR with a 21 x 7 window and a
5 x 3 hole (Pfa 1e-5), and the backward with respect to the waveform phases
(Doppler spectrum recomputed from the compressed cube).

Checked against the FP64 oracle (oracle/mxr_oracle.c, oracle/grad.py)
stage by stage, each stage within a stated FP32 error bound
(EPS = 2^-23, the FP32 machine epsilon):
  1. pulse compression: relative L2 <= 1e-5, and every lag within
        |y32 - y64| <= EPS * ||y64 row||_2
     (an FFT-based correlation's absolute error on one lag is a small
     multiple of eps times the norm of the block it came from);
  2. canceller + window + 512-point Doppler FFT + |.|^2, on the GPU's own
     compressed cube y32: every cell within
        |pw32 - pw64(y32)| <= log2(nd) * EPS * E_r,  E_r = sum_k pw64[k, r]
     (E_r is ||Doppler column r||^2: an FFT's error on one bin is at most
     about log2(nd) * eps * ||column||, and a power error is at most
     2 |D| |dD| <= 2 ||column|| |dD|);
  3. CA-CFAR on the GPU's own power map: the FP64 CFAR of pw32 equals the
     GPU mask outside a 1e-5 relative margin;
  4. end to end: power map relative L2 <= 1e-5 against the FP64 chain, and
     every detection cell that differs from the FP64 chain's mask is a cell
     whose FP64 margin pw64 - thr64 lies inside the band that the measured
     FP32 power errors (bounded by 1 and 2) can cross, plus the CFAR
     stage's own 1e-5 tie margin (3):
        |pw64 - thr64| <= |dpw| + (alpha / norm) * sum over the window |dpw|
                          + 1e-5 * thr64,
     with at most 1e-5 of the cells flipping;
  5. gradient of L = sum(wmap * pw) with respect to the phases: relative
     L2 <= 1e-4 against the FP64 analytic gradient (pinned to FP64 central
     differences in tests/test_grad_oracle.py).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

RTOL = 1e-5
GTOL = 1e-4
EPS = 2.0 ** -23
C, P, S, L, ND = 1, 512, 32768, 4096, 512
WIN, GUARD, PFA = (7, 21), (3, 5), 1e-5


def rel(a, b):
    a = np.asarray(a, dtype=np.complex128)
    b = np.asarray(b, dtype=np.complex128)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def rect_mask(win, guard):
    m = np.ones(win)
    cy, cx, gy, gx = win[0] // 2, win[1] // 2, guard[0] // 2, guard[1] // 2
    m[cy - gy:cy + gy + 1, cx - gx:cx + gx + 1] = 0
    return m


def window_sum(a, mask):
    """sum over the CFAR window (zero outside the map), FP64, same
    orientation as the oracle (mask symmetric)."""
    from scipy.signal import convolve2d
    return convolve2d(a, mask, mode="same")


@pytest.fixture(scope="module")
def mx():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import matx_amd
    matx_amd.lib()
    return matx_amd


@pytest.fixture(scope="module")
def case():
    """Unit complex noise (E|n|^2 = 1) plus 8 echoes of the 4096-sample LFM
    (0.5 fs) at seeded delays and Doppler bins, SNR uniform 0-25 dB."""
    rng = np.random.default_rng(5)
    n = np.arange(L, dtype=np.float64)
    phi = 2 * np.pi * (-0.25 * n + 0.25 / L * n * n)
    lfm = np.exp(1j * phi)
    x = ((rng.standard_normal((C, P, S)) + 1j * rng.standard_normal((C, P, S))) /
         np.sqrt(2)).astype(np.complex64)
    p = np.arange(P)
    cells = []
    for _ in range(8):
        d = int(rng.integers(0, S - L + 1))
        kd = int(rng.integers(0, ND))
        amp = 10 ** (rng.uniform(0, 25) / 20)
        ramp = amp * np.exp(2j * np.pi * kd * p / ND + 1j * rng.uniform(0, 2 * np.pi))
        x[0, :, d:d + L] += (ramp[:, None] * lfm[None, :]).astype(np.complex64)
        cells.append((kd, d))
    return x, phi.astype(np.float32), lfm.astype(np.complex64), cells


def test_large_chain_forward_vs_oracle(mx, orc, case):
    x, _, w, _ = case
    pipe = mx.FusedRadarPipeline(C, P, S, L, nd=ND, tpc="valid", win=WIN, guard=GUARD, pfa=PFA,
                                 want_pow=True)
    assert pipe.n == 32768
    pipe.set_waveform(torch.from_numpy(w).cuda())
    dets = pipe.step(torch.from_numpy(x).cuda())
    torch.cuda.synchronize()
    dets = dets.cpu().numpy()
    y32 = pipe.y.cpu().numpy().astype(np.complex128)
    pw = pipe.pw.cpu().numpy().astype(np.float64)
    mask = rect_mask(WIN, GUARD)
    st = orc.radar_pipeline(x, w, Nd=ND, mode="model", mask=mask, stages=True)

    # 1. pulse compression
    y64 = st["y"]
    assert rel(y32, y64) < RTOL
    row = np.linalg.norm(y64, axis=-1, keepdims=True)
    r1 = float((np.abs(y32 - y64) / (EPS * row)).max())
    assert r1 <= 1.0, f"compressed lags reach {r1:.3g} x eps*||row||"

    # 2. pulse-axis stages on the GPU's own compressed cube
    pw64y = orc.abs2(orc.doppler(orc.tpc(y32, "valid"), P - 2, ND))
    E = pw64y.sum(axis=1, keepdims=True)
    r2 = float((np.abs(pw - pw64y) / (np.log2(ND) * EPS * E)).max())
    assert r2 <= 1.0, f"power cells reach {r2:.3g} x log2(nd)*eps*||column||^2"

    # 3. the CFAR stage alone on the GPU's own power map
    cf = orc.cfar(pw, mask, PFA)
    near = np.abs(cf["margin"]) < 1e-5
    assert int(((dets != cf["dets"]) & ~near).sum()) == 0

    # 4. end to end against the FP64 chain
    pw64 = st["pow"]
    assert rel(pw, pw64) < RTOL
    norm = st["norm"]
    alpha = norm * (PFA ** (-1.0 / norm) - 1.0)
    thr = alpha * st["ba"]
    dpw = np.abs(pw - pw64)
    band = dpw + alpha / norm * window_sum(dpw[0], mask)[None] + 1e-5 * thr
    flips = dets != st["dets"]
    outside = flips & ~(np.abs(pw64 - thr) <= band)
    assert not outside.any(), f"{int(outside.sum())} detection cells differ beyond the FP32 errors"
    assert flips.mean() <= 1e-5, f"{int(flips.sum())} of {flips.size} cells flip"
    assert st["dets"].sum() > 0


def test_large_chain_gradient_vs_oracle(mx, case):
    from matx_amd import autograd as ag
    from oracle import grad as og
    x, phi, _, cells = case
    R = S - L + 1
    wmap = og.contrast_weights(ND, R, cells, C)
    L64, g64 = og.radar_loss_and_grad(phi.astype(np.float64), x, wmap, S, ND, "valid")
    plan = mx.ops.PulseCompressionPlan(S, L, S)
    p = torch.from_numpy(phi).cuda().requires_grad_()
    wv = torch.polar(torch.ones_like(p), p)
    pw = ag.radar_power(wv, torch.from_numpy(x).cuda(), plan, ND, tpc="valid")
    loss = ag.weighted_power_sum(pw, torch.from_numpy(wmap.astype(np.float32)).cuda())
    loss.backward()
    torch.cuda.synchronize()
    assert abs(float(loss.detach()) - L64) / abs(L64) < RTOL
    assert rel(p.grad.cpu().numpy(), g64) < GTOL
    plan.close()
