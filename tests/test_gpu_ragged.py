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


@pytest.mark.parametrize("autotune", [False, True])
def test_ragged_edge_shapes(mx, orc, autotune):
    """Uneven corners of the job contract against the FP64 oracle: pulse
    counts that are not powers of two (P = 2, 3, 5, 17, 100, so nd = 2, 4,
    8, 32, 128 with zero rows), a single kept lag (S = L), a one-sample
    waveform, the shortest rows the batch takes (S = 513 .. 1024, one
    1024-point block) and a transform at 8192."""
    rng = np.random.default_rng(12)
    shapes = [(1, 2, 600, 600), (2, 3, 700, 1), (1, 5, 1023, 512), (1, 17, 2000, 7),
              (3, 100, 4097, 2048)]
    jobs = []
    for C, P, S, L in shapes:
        x = (rng.standard_normal((C, P, S)) + 1j * rng.standard_normal((C, P, S))).astype(np.complex64)
        w = (rng.standard_normal(L) + 1j * rng.standard_normal(L)).astype(np.complex64)
        jobs.append((torch.from_numpy(x).cuda(), torch.from_numpy(w).cuda()))
    rb = mx.ragged.RaggedBatch(jobs, autotune=autotune)
    out = rb.execute()
    torch.cuda.synchronize()
    for i, ((x, w), (C, P, S, L)) in enumerate(zip(jobs, shapes)):
        nd = mx.ragged.next_pow2(P)
        assert tuple(out[i].shape) == (C, nd, S - L + 1)
        xh, wh = x.cpu().numpy(), w.cpu().numpy()
        n = mx.ragged.next_pow2(S)
        Wc = orc.waveform_spectrum(wh, n, True, True)
        y = orc.pulse_compress(xh.reshape(C * P, S), Wc, S - L + 1).reshape(C, P, S - L + 1)
        ref = orc.abs2(orc.doppler(y, P, nd))
        assert rel(out[i].cpu().numpy(), ref) < RTOL, (i, shapes[i], rb.block_size(i))
    rb.close()


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


def test_ragged_capture_without_warmup(mx):
    """The plan builds its tables, spectra buffers and branch streams when it
    is created, so its very first execute can be captured into a graph with
    no eager run before it; every replay matches the per-job entries."""
    jobs = make_jobs(4, 5)
    s = torch.cuda.Stream()
    rb = mx.ragged.RaggedBatch(jobs, autotune=False, stream=s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        rb.execute()
    refs = [plain(mx, x, w) for x, w in jobs]
    torch.cuda.synchronize()
    for k in range(2):
        for p in rb.pw:
            p.fill_(-1.0)
        with torch.cuda.stream(s):
            g.replay()
        s.synchronize()
        for i, ref in enumerate(refs):
            r = float((rb.pw[i] - ref).norm() / ref.norm())
            assert r < RTOL, (k, i, r)
    rb.close()


def test_plans_survive_clear_caches(mx):
    """mxr_clear_caches frees the library's cached tables, spectra and
    workspaces but keeps what live plans point at: a pulse-compression plan
    and a ragged plan give the same result before and after the clear."""
    jobs = make_jobs(6, 3)
    x, w = jobs[0]
    C, P, S = x.shape
    plan = mx.ops.PulseCompressionPlan(S, w.shape[0], mx.ragged.next_pow2(S))
    plan.set_waveform(w)
    rb = mx.ragged.RaggedBatch(jobs, autotune=False)
    y0 = plan(x).clone()
    p0 = [p.clone() for p in rb.execute()]
    torch.cuda.synchronize()
    mx.lib().mxr_clear_caches()
    y1 = plan(x)
    p1 = rb.execute()
    torch.cuda.synchronize()
    assert torch.equal(y0, y1)
    for a, b in zip(p0, p1):
        assert torch.equal(a, b)
    plan.close()
    rb.close()


def config3_jobs(seed, n=256):
    """n jobs from the ragged workload's distribution: C uniform in {1..8},
    P in {16, 32, 64, 128, 256}, S uniform in [1024, 16384], L uniform in
    [64, min(2048, S)]; unit complex noise plus 5 echoes of the job's
    Hamming-windowed LFM (0.5 fs) at 10 dB.  The first five jobs are pinned
    to the corners of the distribution (P = 128 and 256, S = 16384,
    L = 2048, C = 8, the smallest job) so every kernel class is drawn."""
    rng = np.random.default_rng(seed)
    g = torch.Generator(device="cuda").manual_seed(seed)
    pinned = [(8, 256, 16384, 2048), (3, 128, 16384, 1500), (2, 256, 9000, 2048),
              (8, 128, 12000, 64), (1, 16, 1024, 64)]
    jobs, shapes = [], []
    for i in range(n):
        if i < len(pinned):
            C, P, S, L = pinned[i]
        else:
            C = int(rng.integers(1, 9))
            P = int(rng.choice([16, 32, 64, 128, 256]))
            S = int(rng.integers(1024, 16385))
            L = int(rng.integers(64, min(2048, S) + 1))
        t = np.arange(L)
        w = torch.tensor((np.exp(2j * np.pi * (-0.25 * t + 0.25 / L * t * t)) *
                          np.hamming(L)).astype(np.complex64), device="cuda")
        x = torch.randn((C, P, S), dtype=torch.complex64, device="cuda", generator=g)
        for d in rng.integers(0, S - L + 1, 5):
            x[:, :, int(d):int(d) + L] += np.sqrt(10.0) * w
        jobs.append((x, w))
        shapes.append((C, P, S, L))
    return jobs, shapes


@pytest.mark.parametrize("persistent", [False, True])
def test_ragged_config3_distribution(mx, orc, persistent):
    """The full 256-job batch: every job against the per-job entries on the
    GPU, and the pinned corner jobs plus three drawn jobs against the FP64
    oracle (rel L2 <= 1e-5), for the default and the persistent launch."""
    jobs, shapes = config3_jobs(31)
    rb = mx.ragged.RaggedBatch(jobs, autotune=True, persistent=persistent)
    out = rb.execute()
    torch.cuda.synchronize()
    ps = {s[1] for s in shapes}
    assert {128, 256} <= ps
    sizes = {rb.block_size(i) for i in range(len(jobs))}
    assert len(sizes) >= 2, sizes
    for i, (x, w) in enumerate(jobs):
        ref = plain(mx, x, w)
        r = float((out[i] - ref).norm() / ref.norm())
        assert r < RTOL, (i, shapes[i], rb.block_size(i), r)
    check = list(range(5)) + [17, 101, 222]
    for i in check:
        x, w = jobs[i]
        xh, wh = x.cpu().numpy(), w.cpu().numpy()
        C, P, S, L = shapes[i]
        n = mx.ragged.next_pow2(S)
        Wc = orc.waveform_spectrum(wh, n, True, True)
        y = orc.pulse_compress(xh.reshape(C * P, S), Wc, S - L + 1).reshape(C, P, S - L + 1)
        D = orc.doppler(y, P, mx.ragged.next_pow2(P))
        assert rel(out[i].cpu().numpy(), orc.abs2(D)) < RTOL, (i, shapes[i], rb.block_size(i))
    rb.close()
