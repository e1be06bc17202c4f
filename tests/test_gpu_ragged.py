"""GPU parity of the ragged batch (mxr_ragged_*): every job's power map
against the per-job entries (mxr_pulse_compress then mxr_doppler, rel L2
<= 1e-5) and, for the smallest jobs, against the FP64 oracle; graph replays
against the eager result."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

RTOL = 1e-5


def rel(a, b):
    a = np.asarray(a, dtype=np.complex128)
    b = np.asarray(b, dtype=np.complex128)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


@pytest.fixture(scope="module")
def mx():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import matx_amd
    matx_amd.lib()
    return matx_amd


def make_jobs(seed, n=10):
    rng = np.random.default_rng(seed)
    jobs = []
    for _ in range(n):
        C = int(rng.integers(1, 4))
        P = int(rng.choice([16, 32, 64]))
        S = int(rng.integers(1024, 5000))
        L = int(rng.integers(64, min(1024, S)))
        x = (rng.standard_normal((C, P, S)) + 1j * rng.standard_normal((C, P, S))).astype(np.complex64)
        t = np.arange(L)
        w = np.exp(1j * np.pi * 0.5 * (t - L / 2) ** 2 / L).astype(np.complex64)
        x[:, :, 100:100 + L] += 3 * w
        jobs.append((torch.from_numpy(x).cuda(), torch.from_numpy(w).cuda()))
    return jobs


def plain(mx, x, w):
    C, P, S = x.shape
    n = mx.ragged.next_pow2(S)
    y = mx.pulse_compress(x, w, n)
    pw, _ = mx.doppler(y, mx.ragged.next_pow2(P), tpc="none", win_rows=P)
    return pw


@pytest.mark.parametrize("autotune,persistent", [(False, False), (True, False), (True, True)])
def test_ragged_matches_per_job_calls(mx, autotune, persistent):
    jobs = make_jobs(1)
    rb = mx.ragged.RaggedBatch(jobs, autotune=autotune, persistent=persistent)
    out = rb.execute()
    torch.cuda.synchronize()
    for i, (x, w) in enumerate(jobs):
        ref = plain(mx, x, w)
        assert rel(out[i].cpu().numpy(), ref.cpu().numpy()) < RTOL, (i, rb.block_size(i))
        assert rb.block_size(i) in (1024, 2048, 4096, 8192)
    rb.close()


def test_ragged_against_oracle(mx, orc):
    jobs = make_jobs(2, 3)
    rb = mx.ragged.RaggedBatch(jobs, autotune=True)
    out = rb.execute()
    torch.cuda.synchronize()
    for i, (x, w) in enumerate(jobs):
        xh, wh = x.cpu().numpy(), w.cpu().numpy()
        C, P, S = xh.shape
        n = mx.ragged.next_pow2(S)
        R = S - wh.shape[0] + 1
        Wc = orc.waveform_spectrum(wh, n, True, True)
        y = orc.pulse_compress(xh.reshape(C * P, S), Wc, R).reshape(C, P, R)
        D = orc.doppler(y, P, mx.ragged.next_pow2(P))
        assert rel(out[i].cpu().numpy(), orc.abs2(D)) < RTOL


def test_ragged_graph_replays(mx):
    jobs = make_jobs(3, 6)
    s = torch.cuda.Stream()
    rb = mx.ragged.RaggedBatch(jobs, autotune=False, stream=s)
    with torch.cuda.stream(s):
        rb.execute()
    s.synchronize()
    eager = [p.clone() for p in rb.pw]
    g = rb.graph(stream=s)
    for k in range(3):
        for p in rb.pw:
            p.fill_(-1.0)
        with torch.cuda.stream(s):
            plain(mx, *jobs[0])   # unrelated eager work between replays
            g.replay()
        s.synchronize()
        for p, e in zip(rb.pw, eager):
            assert torch.equal(p, e), k
    rb.close()
