"""GPU parity of the HIP path (through the C ABI) against the FP64 oracle.

Tolerances: relative L2 <= 1e-5 for every floating-point output (FP32 arithmetic
against an FP64 oracle); CFAR detection masks identical except at cells whose
oracle margin |pow - alpha*ba| / (alpha*ba) is below 1e-5 (float rounding may
legitimately flip those); integer outputs exact.
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


def crandn(rng, *shape):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)


@pytest.fixture(scope="module")
def mx():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import matx_amd
    matx_amd.lib()
    return matx_amd


def dets_match(got, ref):
    """Masks equal except where the oracle is within MARGIN of threshold."""
    got = np.asarray(got)
    near = np.abs(ref["margin"]) < MARGIN
    bad = (got != ref["dets"]) & ~near
    return int(bad.sum()), int(near.sum())


# ----------------------------------------------------------------- FFT
@pytest.mark.parametrize("n", [1, 2, 4, 8, 16, 32, 64, 128, 256, 512, 1024, 2048, 4096, 8192, 16384])
def test_fft_sizes(mx, orc, n):
    rng = np.random.default_rng(n)
    b = max(1, 8192 // n)
    x = crandn(rng, b, n)
    X = host(mx.fft(dev(x)))
    assert rel(X, orc.fft(x)) < RTOL
    xi = host(mx.ifft(dev(X)))
    assert rel(xi, x) < RTOL


def test_fft_pad_truncate_norms(mx, orc):
    rng = np.random.default_rng(1)
    x = crandn(rng, 3, 100)
    for n in (128, 64):
        for norm in ("backward", "forward", "ortho"):
            assert rel(host(mx.fft(dev(x), n, norm=norm)), orc.fft(x, n, norm=norm)) < RTOL
            assert rel(host(mx.ifft(dev(x), n, norm=norm)), orc.ifft(x, n, norm=norm)) < RTOL


def test_fft_strided_axis_and_inplace(mx, orc):
    rng = np.random.default_rng(2)
    x = crandn(rng, 4, 256, 300)
    xd = dev(x)
    # along a strided axis (the Doppler form, fft(tpcView.Permute({0,2,1})))
    got = host(mx.fft(xd, axis=1))
    assert rel(got, orc.fft(x, axis=1)) < RTOL
    # permuted view in, permuted view out, in place
    v = xd.permute(0, 2, 1)
    mx.fft(v, out=v)
    assert rel(host(xd), orc.fft(x, axis=1)) < RTOL


@pytest.mark.parametrize("n", [3, 5, 6, 7, 12, 100, 127, 243, 1000, 1023, 3000, 4097, 5000, 8191])
def test_fft_bluestein_sizes(mx, orc, n):
    """Sizes that are not powers of two (chirp-z through a power-of-two
    convolution of length >= 2n - 1) against the FP64 O(n^2) DFT."""
    rng = np.random.default_rng(n)
    b = max(1, 4096 // n)
    x = crandn(rng, b, n)
    X = host(mx.fft(dev(x)))
    assert rel(X, orc.fft(x)) < RTOL
    xi = host(mx.ifft(dev(X)))
    assert rel(xi, x) < RTOL


def test_fft_bluestein_pad_norms_strided_inplace(mx, orc):
    rng = np.random.default_rng(3)
    x = crandn(rng, 3, 100)
    for n in (90, 120, 1000):  # truncate, pad
        for norm in ("backward", "forward", "ortho"):
            assert rel(host(mx.fft(dev(x), n, norm=norm)), orc.fft(x, n, norm=norm)) < RTOL
            assert rel(host(mx.ifft(dev(x), n, norm=norm)), orc.ifft(x, n, norm=norm)) < RTOL
    y = crandn(rng, 2, 150, 64)
    yd = dev(y)
    assert rel(host(mx.fft(yd, axis=1)), orc.fft(y, axis=1)) < RTOL
    v = yd.permute(0, 2, 1)
    mx.fft(v, out=v)
    assert rel(host(yd), orc.fft(y, axis=1)) < RTOL


def test_fft_unsupported_size_raises(mx):
    x = torch.zeros((1, 3), dtype=torch.complex64, device="cuda")
    with pytest.raises(Exception, match="unsupported"):
        mx.fft(x, 1 << 25)


@pytest.mark.parametrize("n,b", [(32768, 3), (65536, 2), (1 << 18, 2), (1 << 20, 1), (1 << 22, 1)])
def test_fft_long_sizes(mx, orc, n, b):
    """Powers of two above one workgroup (two passes through the work
    matrix) against the FP64 oracle FFT, and the inverse round trip."""
    rng = np.random.default_rng(n)
    x = crandn(rng, b, n)
    X = host(mx.fft(dev(x)))
    assert rel(X, orc.fft(x)) < RTOL
    assert rel(host(mx.ifft(dev(X))), x) < RTOL


@pytest.mark.parametrize("n,b", [(8193, 3), (10000, 2), (50000, 2), (100003, 1), (1000000, 1)])
def test_fft_long_bluestein(mx, n, b):
    """Sizes that are not powers of two above the one-workgroup Bluestein
    range (two-pass convolution transforms) against numpy's FP64 FFT."""
    rng = np.random.default_rng(n)
    x = crandn(rng, b, n)
    X = host(mx.fft(dev(x)))
    assert rel(X, np.fft.fft(x.astype(np.complex128))) < RTOL
    assert rel(host(mx.ifft(dev(X))), x) < RTOL


def test_fft_long_pad_norms_strided_inplace_real(mx):
    """The size rules, norms, views and real forms on the two-pass paths."""
    rng = np.random.default_rng(4)
    x = crandn(rng, 2, 30000).astype(np.complex128)
    for n in (32768, 20000, 65536, 40000):  # pad / truncate, power of two or not
        for norm in ("backward", "forward", "ortho"):
            assert rel(host(mx.fft(dev(x), n, norm=norm)), np.fft.fft(x, n, norm=norm)) < RTOL
            assert rel(host(mx.ifft(dev(x), n, norm=norm)), np.fft.ifft(x, n, norm=norm)) < RTOL
    y = crandn(rng, 3, 32768, 4).astype(np.complex128)
    yd = dev(y)
    assert rel(host(mx.fft(yd, axis=1)), np.fft.fft(y, axis=1)) < RTOL
    v = yd.permute(0, 2, 1)
    mx.fft(v, out=v)
    assert rel(host(yd), np.fft.fft(y, axis=1)) < RTOL
    r = rng.standard_normal((2, 50000))
    for n in (65536, 50000):
        R = host(mx.rfft(dev(r, torch.float32), n))
        assert rel(R, np.fft.rfft(r, n)) < RTOL
        assert rel(host(mx.irfft(dev(R), n)), np.fft.irfft(np.fft.rfft(r, n), n)) < RTOL


@pytest.mark.parametrize("n", [1, 2, 8, 100, 1000, 1024, 4096, 4097, 5000, 16384])
def test_rfft_irfft_sizes(mx, n):
    """R2C / C2R (mxr_fft_r2c / mxr_fft_c2r, powers of two and Bluestein
    sizes) against numpy's FP64 rfft / irfft (numpy's FFT pins the oracle's
    in tests/test_oracle.py)."""
    rng = np.random.default_rng(n + 7)
    b = max(1, 8192 // n)
    x = rng.standard_normal((b, n)).astype(np.float32)
    X = host(mx.rfft(dev(x, torch.float32)))
    assert X.shape == (b, n // 2 + 1)
    assert rel(X, np.fft.rfft(x.astype(np.float64))) < RTOL
    xr = host(mx.irfft(dev(X), n))
    assert xr.dtype == np.float32 and xr.shape == (b, n)
    assert rel(xr, x) < RTOL
    # bins that are not a Hermitian half (imaginary DC / Nyquist parts):
    # dropped, as numpy and cuFFT drop them
    Z = crandn(rng, b, n // 2 + 1)
    assert rel(host(mx.irfft(dev(Z), n)), np.fft.irfft(Z.astype(np.complex128), n)) < RTOL


def test_rfft_irfft_pad_norms_axis(mx):
    rng = np.random.default_rng(11)
    x = rng.standard_normal((3, 100, 6)).astype(np.float32)
    xd = dev(x, torch.float32)
    for n in (64, 100, 128, 150):          # truncate, exact, pad (pow2 and not)
        for norm in ("backward", "forward", "ortho"):
            X = host(mx.rfft(xd, n, axis=1, norm=norm))
            ref = np.fft.rfft(x.astype(np.float64), n, axis=1, norm=norm)
            assert rel(X, ref) < RTOL, (n, norm)
            back = host(mx.irfft(dev(X), n, axis=1, norm=norm))
            assert rel(back, np.fft.irfft(ref, n, axis=1, norm=norm)) < RTOL, (n, norm)
    # irfft with fewer bins than n/2 + 1 (zero-padded) and with more (truncated)
    Z = crandn(rng, 4, 40)
    for n in (200, 40, 33):
        assert rel(host(mx.irfft(dev(Z), n)), np.fft.irfft(Z.astype(np.complex128), n)) < RTOL, n
    # default sizes: rfft keeps n, irfft makes 2 (N - 1)
    assert host(mx.irfft(dev(Z))).shape == (4, 78)


def test_fft2_family(mx, orc):
    rng = np.random.default_rng(12)
    x = crandn(rng, 2, 96, 128)
    X = host(mx.fft2(dev(x)))
    ref = orc.fft(orc.fft(x, axis=-1), axis=-2)
    assert rel(X, ref) < RTOL
    assert rel(host(mx.ifft2(dev(X))), x) < RTOL
    r = rng.standard_normal((2, 60, 100)).astype(np.float32)
    R = host(mx.rfft2(dev(r, torch.float32)))
    assert rel(R, np.fft.rfft2(r.astype(np.float64))) < RTOL
    assert rel(host(mx.irfft2(dev(R), 100)), r) < RTOL
    for norm in ("forward", "ortho"):
        assert rel(host(mx.fft2(dev(x), norm=norm)), np.fft.fft2(x.astype(np.complex128), norm=norm)) < RTOL


def test_dlpack_import_through_the_abi(mx):
    """A torch tensor and a permuted view handed over as DLPack capsules
    (mxr_tensor_from_dlpack, no copy) and transformed by mxr_fft_c2c."""
    from matx_amd import _lib as L
    rng = np.random.default_rng(41)
    x = crandn(rng, 3, 64, 40)
    xd = dev(x)
    for view, axis in ((xd, 2), (xd.permute(0, 2, 1), 2)):
        d, cap = L.desc_from_dlpack(view)
        assert d.data == view.data_ptr()
        assert list(d.shape[:3]) == list(view.shape) and list(d.stride[:3]) == list(view.stride())
        out = torch.empty(view.shape, dtype=torch.complex64, device="cuda")
        L.check(L.lib().mxr_fft_c2c(L.ref(L.desc(out)), L.ctypes.byref(d), 0, L.FORWARD,
                                    L.NORM_BACKWARD, axis, L.stream_handle(None)))
        ref = np.fft.fft(host(view).astype(np.complex128), axis=axis)
        assert rel(host(out), ref) < RTOL
        del cap


def test_dlpack_export_torch_round_trip(mx):
    """mxr_tensor_to_dlpack (DLPack export) consumed by torch: the torch
    tensor aliases the exported view (same memory, shape, strides), writes
    through either side are seen by the other, and the export is released
    when torch drops it."""
    from torch.utils import dlpack as tdl
    from matx_amd import _lib as L
    rng = np.random.default_rng(42)
    base = dev(crandn(rng, 4, 32, 50))
    for view in (base, base.permute(0, 2, 1), base[:, 3:20, ::2]):
        n0 = len(L._exports)
        y = tdl.from_dlpack(L.to_dlpack(view))
        assert y.data_ptr() == view.data_ptr()
        assert tuple(y.shape) == tuple(view.shape) and y.stride() == view.stride()
        assert y.dtype == torch.complex64 and y.device == view.device
        y.mul_(2)
        assert torch.equal(y, view)
        assert len(L._exports) == n0 + 1
        del y
        assert len(L._exports) == n0
    pw = torch.rand((3, 7), device="cuda")
    y = tdl.from_dlpack(L.to_dlpack(pw))
    assert torch.equal(y, pw) and y.dtype == torch.float32


# --------------------------------------------------------- fft_conv / PC
def lfm(L, bw, rng=None):
    n = np.arange(L)
    t = n / 1.0
    f0, f1 = -bw / 2, bw / 2
    ph = 2 * np.pi * (f0 * t + 0.5 * (f1 - f0) / L * t * t)
    return (np.exp(1j * ph) * np.hamming(L)).astype(np.complex64)


def test_fft_conv_4x32x2048_chirp256(mx, orc):
    """4 x 32 x 2048 cube, 256-sample LFM (0.5 fs), N = 4096, all lags."""
    rng = np.random.default_rng(3)
    x = crandn(rng, 4, 32, 2048)
    w = lfm(256, 0.5)
    got = host(mx.fft_conv(dev(x), dev(w), 4096))
    assert got.shape == (4, 32, 4096)
    assert rel(got, orc.fft_conv(x, w, 4096)) < RTOL
    conv = host(mx.fft_conv(dev(x), dev(w), 4096, conj=False))
    ref = np.fft.ifft(np.fft.fft(x.astype(complex), 4096) * np.fft.fft(w.astype(complex), 4096))
    assert rel(conv, ref) < RTOL


@pytest.mark.parametrize("S,L,n,nout", [
    (9000, 1000, 16384, 8001),    # radar config: overlap-save, M = 4096
    (2048, 256, 4096, 4096),      # full circular, single block
    (16000, 64, 16384, 16384),    # overlap-save blocks that wrap circularly
    (3000, 700, 4096, 2301),
    (1500, 1500, 2048, 1),        # waveform as long as the pulse
    (32768, 4096, 32768, 28673),  # M = 8192, 7 blocks of 4097 lags (not 64-aligned)
    (12000, 1500, 16384, 10501),  # M = 8192, 64-aligned blocks
    (16000, 3000, 16384, 16384),  # M = 8192 blocks that wrap circularly
    (5000, 300, 8192, 8192),      # one 8192-point block, full circular output
])
def test_fft_conv_shapes(mx, orc, S, L, n, nout):
    rng = np.random.default_rng(S + L)
    x = crandn(rng, 3, S)
    w = crandn(rng, L)
    got = host(mx.fft_conv(dev(x), dev(w), n, nout=nout))
    ref = orc.fft_conv(x, w, n)[:, :nout]
    assert rel(got, ref) < RTOL


def test_pulse_compress_matx_input(mx, orc, matx_inputs, golden):
    x, w = matx_inputs
    Wc = orc.waveform_spectrum(w, 16384)
    ref = orc.pulse_compress(x, Wc, 8001)
    got = host(mx.pulse_compress(dev(x), dev(w), 16384))
    assert rel(got, ref) < RTOL
    yi = golden["model_y_idx"]
    assert rel(got[yi[:, 0], yi[:, 1]], golden["model_y_val"]) < RTOL


def test_pulse_compress_inplace_writeback(mx, orc, matx_inputs):
    x, w = matx_inputs
    C, N, S, R = 2, 16384, 9000, 8001
    buf = torch.zeros((C, 128, N), dtype=torch.complex64, device="cuda")
    buf[:, :, :S] = dev(x)
    wv = torch.zeros(N, dtype=torch.complex64, device="cuda")
    wv[:1000] = dev(w)
    mx.pulse_compress(buf[:, :, :S], wv[:1000], N, nout=R, writeback=True, out=buf[:, :, :R])
    Wc, w2 = orc.waveform_spectrum(w, N, return_prepared=True)
    ref = orc.pulse_compress(x, Wc, R)
    got = host(buf)
    for c in range(C):
        assert rel(got[c, :, :R], ref) < RTOL
    assert np.array_equal(got[:, :, R:S], np.broadcast_to(x[:, R:S], (C, 128, S - R)))
    assert rel(host(wv)[:1000], w2) < RTOL
    assert np.all(host(wv)[1000:] == 0)


# ------------------------------------------------------------ pulse axis
@pytest.mark.parametrize("mode,nd", [("same", 256), ("valid", 128), ("valid", 256), ("same", 512)])
def test_doppler_modes(mx, orc, mode, nd):
    rng = np.random.default_rng(nd)
    C, P, R = 2, 128, 333
    y = crandn(rng, C, P, R)
    pw, rd = mx.doppler(dev(y), nd, tpc=mode, want_rd=True)
    t = orc.tpc(y, mode)
    D = orc.doppler(t, P - 2, nd)
    assert rel(host(rd), D) < RTOL
    assert rel(host(pw), np.abs(D) ** 2) < RTOL


def test_doppler_inplace_stage_form(mx, orc):
    rng = np.random.default_rng(5)
    C, P, Pr, R = 2, 128, 256, 200
    t = np.zeros((C, Pr, R), np.complex64)
    t[:, :P] = crandn(rng, C, P, R)
    td = dev(t)
    mx.doppler(td, Pr, tpc="none", win_rows=P - 2, rd=td, want_pow=False)
    D = orc.doppler(t[:, :P], P - 2, Pr)
    assert rel(host(td), D) < RTOL


# ------------------------------------------------------------------ CFAR
def test_cfar_random_maps(mx, orc):
    rng = np.random.default_rng(6)
    pw = rng.exponential(1.0, (3, 128, 1500)).astype(np.float32)
    for _ in range(40):
        pw[rng.integers(3), rng.integers(128), rng.integers(1500)] *= rng.uniform(5, 1e6)
    ref = orc.cfar(pw)
    dets, ba, _ = mx.cfar(dev(pw, torch.float32), want_ba=True)
    assert rel(host(ba), ref["ba"]) < RTOL
    bad, near = dets_match(host(dets), ref)
    assert bad == 0, (bad, near)
    assert host(dets).sum() > 0


def test_cfar_complex_input_and_edges(mx, orc):
    rng = np.random.default_rng(7)
    D = crandn(rng, 1, 9, 20)     # smaller than the window in rows
    D[0, 4, 10] *= 30
    pwr = np.abs(D.astype(complex)) ** 2
    ref = orc.cfar(pwr)
    pw_out = torch.empty((1, 9, 20), dtype=torch.float32, device="cuda")
    dets, ba, _ = mx.cfar(dev(D), pw=pw_out, want_ba=True)
    assert rel(host(pw_out), pwr) < RTOL
    assert rel(host(ba), ref["ba"]) < RTOL
    bad, _ = dets_match(host(dets), ref)
    assert bad == 0


@pytest.mark.parametrize("C,Ny,Nx,win,guard,aligned", [
    (2, 512, 3001, (7, 21), (3, 5), True),    # the large cube's window, ragged last tile
    (3, 40, 777, (5, 13), (3, 5), True),      # fewer rows than one tile band
    (1, 130, 1024, (5, 13), (3, 3), True),    # MatX mask
    (2, 70, 513, (7, 21), (3, 5), False),     # unaligned rows: the register-only kernel
])
def test_cfar_lds_tiles(mx, orc, C, Ny, Nx, win, guard, aligned):
    """The LDS-staged CFAR (f32 map, 16-byte aligned rows) and the
    register-only one (unaligned rows) against the FP64 oracle, with the
    background averages and detections near strong cells and map edges."""
    rng = np.random.default_rng(Ny + Nx)
    pw = rng.exponential(1.0, (C, Ny, Nx)).astype(np.float32)
    for _ in range(60):
        pw[rng.integers(C), rng.integers(Ny), rng.integers(Nx)] *= rng.uniform(5, 1e5)
    pw[0, 0, 0] *= 1e4
    pw[-1, -1, -1] *= 1e4
    mask = np.ones(win)
    cy, cx, gy, gx = win[0] // 2, win[1] // 2, guard[0] // 2, guard[1] // 2
    mask[cy - gy:cy + gy + 1, cx - gx:cx + gx + 1] = 0
    ref = orc.cfar(pw.astype(np.float64), mask)
    ld = (Nx + 63) // 64 * 64 if aligned else Nx + 1
    buf = torch.zeros((C, Ny, ld), dtype=torch.float32, device="cuda")
    buf[:, :, :Nx] = dev(pw, torch.float32)
    view = buf[:, :, :Nx]
    dets, ba, _ = mx.cfar(view, win=win, guard=guard, want_ba=True)
    assert rel(host(ba), ref["ba"]) < RTOL
    bad, _ = dets_match(host(dets), ref)
    assert bad == 0
    d2, _, _ = mx.cfar(view, win=win, guard=guard)   # detection-only form (tabulated alpha/norm)
    bad, _ = dets_match(host(d2), ref)
    assert bad == 0
    assert host(dets).sum() > 0


# ------------------------------------------------------------------ conv
def test_conv_kats_gpu(mx):
    import json
    import os
    k = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "kat_convcorr.json")))
    x = dev(np.array(k["conv1d"]["in"], np.float32), torch.float32)
    f = dev(np.array(k["conv1d"]["filter"], np.float32), torch.float32)
    for m in ("full", "same", "valid"):
        assert np.allclose(host(mx.conv1d(x, f, m)), k["conv1d"][m], atol=1e-5)
    x2 = dev(np.array(k["conv2d"]["in"], np.float32), torch.float32)
    f2 = dev(np.array(k["conv2d"]["filter"], np.float32), torch.float32)
    for m in ("full", "same", "valid"):
        assert np.allclose(host(mx.conv2d(x2, f2, m)), k["conv2d"][m], atol=1e-4)


def test_conv_random_strided(mx, orc):
    rng = np.random.default_rng(8)
    x = crandn(rng, 3, 50, 40)
    f = rng.standard_normal(7).astype(np.float32)
    xd = dev(x)
    v = xd.permute(0, 2, 1)        # conv along the middle axis via a permuted view
    for m in ("full", "same", "valid"):
        got = host(mx.conv1d(v, dev(f, torch.float32), m))
        ref = orc.conv1d(np.transpose(x, (0, 2, 1)), f, m)
        assert rel(got, ref) < RTOL
    f2 = crandn(rng, 5, 13)
    for m in ("full", "same", "valid"):
        assert rel(host(mx.conv2d(xd, dev(f2), m)), orc.conv2d(x, f2, m)) < RTOL


def test_conv2d_real_tiles(mx, orc):
    """Real maps and real filters (the CA-CFAR window sums) through the
    staged-tile kernel: several 32 x 64 tiles per map, batch dims, a strided
    output view, the 5 x 13 CFAR mask, other filter heights (one kernel
    instance per height), the largest staged window (16 x 64) and one beyond
    it (the per-output kernel), all three modes."""
    rng = np.random.default_rng(9)
    x = rng.random((2, 3, 70, 150)).astype(np.float32)
    mask = np.ones((5, 13), np.float32)
    mask[1:4, 5:8] = 0
    for f in (mask, rng.standard_normal((16, 64)).astype(np.float32),
              rng.standard_normal((17, 3)).astype(np.float32), np.ones((1, 1), np.float32),
              rng.standard_normal((3, 7)).astype(np.float32), rng.standard_normal((9, 2)).astype(np.float32)):
        for m in ("full", "same", "valid"):
            ref = orc.conv2d(x.reshape(6, 70, 150), f, m).reshape(2, 3, *orc.conv2d(x[0, 0], f, m).shape[-2:])
            got = host(mx.conv2d(dev(x, torch.float32), dev(f, torch.float32), m))
            assert rel(got, ref) < RTOL, (f.shape, m)
    # output written through a transposed view of a larger buffer
    ref = orc.conv2d(x[0, 0], mask, "same")
    buf = torch.full((160, 80), -1.0, device="cuda")
    outv = buf[:150, :70].t()
    mx.conv2d(dev(x[0, 0], torch.float32), dev(mask, torch.float32), "same", out=outv)
    assert rel(host(outv), ref) < RTOL
    assert float(buf[150:].min()) == -1.0 and float(buf[:, 70:].min()) == -1.0


def test_conv1d_strided_axis_into_permuted_view(mx, orc):
    """The canceller form: conv1d along the pulse axis of permuted views, in
    and out (conv axis strided, range axis unit stride), through the strip
    kernel; complex input, real and complex taps, every mode."""
    rng = np.random.default_rng(10)
    C, P, R = 3, 130, 300
    y = crandn(rng, C, P, R)
    yd = dev(y)
    # register-window kernel up to 16 taps (bounds 4, 8, 16), strip kernel
    # beyond; real and complex taps
    filters = (np.array([1, -2, 1], np.float32), crandn(rng, 5), rng.standard_normal(8).astype(np.float32),
               crandn(rng, 12), rng.standard_normal(16).astype(np.float32), rng.standard_normal(40).astype(np.float32))
    for f in filters:
        fd = dev(f, torch.complex64 if np.iscomplexobj(f) else torch.float32)
        for m in ("full", "same", "valid"):
            n = {"full": P + len(f) - 1, "same": P, "valid": P - len(f) + 1}[m]
            outbuf = torch.zeros((C, n, R), dtype=torch.complex64, device="cuda")
            mx.conv1d(yd.permute(0, 2, 1), fd, m, out=outbuf.permute(0, 2, 1))
            ref = orc.conv1d(np.transpose(y, (0, 2, 1)), f, m)
            assert rel(host(outbuf.permute(0, 2, 1)), ref) < RTOL, (len(f), m)
    # real input (a real output view): every window-kernel bound, short
    # conv axis (fewer samples than one strip and than the filter)
    yr = rng.standard_normal((C, 7, R)).astype(np.float32)
    for f in (np.array([1, -2, 1], np.float32), rng.standard_normal(6).astype(np.float32),
              crandn(rng, 9), rng.standard_normal(16).astype(np.float32)):
        fd = dev(f, torch.complex64 if np.iscomplexobj(f) else torch.float32)
        odt = torch.complex64 if np.iscomplexobj(f) else torch.float32
        for m in ("full", "same"):
            n = {"full": 7 + len(f) - 1, "same": 7}[m]
            outbuf = torch.zeros((C, n, R), dtype=odt, device="cuda")
            mx.conv1d(dev(yr, torch.float32).permute(0, 2, 1), fd, m, out=outbuf.permute(0, 2, 1))
            ref = orc.conv1d(np.transpose(yr, (0, 2, 1)).astype(np.complex64), f, m)
            got = host(outbuf.permute(0, 2, 1))
            assert rel(got, ref if np.iscomplexobj(f) else ref.real) < RTOL, (len(f), m)


# ------------------------------------------------- end to end vs reference
def test_model_chain_17_detections(mx, orc, matx_inputs, golden):
    """numpy-model dataflow on MatX's test input: 17 detections per channel
    (test/01_radar/MultiChannelRadarPipeline.cu:187-197), 16 channels."""
    x, w = matx_inputs
    C = 16
    xd = dev(x).unsqueeze(0).expand(C, -1, -1).contiguous()
    pipe = mx.FusedRadarPipeline(C, 128, 9000, 1000, nd=256, tpc="valid", want_pow=True)
    dets = host(pipe.step(xd, dev(w)))
    assert dets.sum() == 17 * C
    for c in range(C):
        assert np.array_equal(np.argwhere(dets[c] == 1), golden["model_det_idx"])
    pw = host(pipe.pw)[0]
    st = orc.radar_pipeline(x[None], w, Nd=256, mode="model", stages=True)
    assert rel(pw, st["pow"][0]) < RTOL


def test_matx_stage_api(mx, orc, matx_inputs, golden):
    """The four stage tests of MultiChannelRadarPipeline.cu, run end to end on
    the MatX-compatible RadarPipeline (SAME canceller, Pr = 256)."""
    x, w = matx_inputs
    C, P, S, L = 4, 128, 9000, 1000
    p = mx.RadarPipeline(P, S, L, C)
    p.GetInputView()[:, :, :S] = dev(x)
    p.GetwaveformView()[:L] = dev(w)
    st = orc.radar_pipeline(x[None], w, mode="matx", stages=True)
    p.PulseCompression()
    got = host(p.GetInputView())[:, :, :8001]
    for c in range(C):
        assert rel(got[c], st["y"][0]) < RTOL
    p.ThreePulseCanceller()
    tp = host(p.GetTPCView())
    for c in range(C):
        assert rel(tp[c, :P], st["tpc"][0]) < RTOL
        assert np.all(tp[c, P:] == 0)
    p.DopplerProcessing()
    dop = host(p.GetTPCView())
    for c in range(C):
        assert rel(dop[c], st["D"][0]) < RTOL
    p.CFARDetections()
    dets = host(p.GetDetections())
    ba = host(p.GetBackgroundAverages())[:, 2:258, 6:8007]
    ref = {"dets": st["dets"][0], "margin": st["margin"][0]}
    for c in range(C):
        assert rel(ba[c], st["ba"][0]) < RTOL
        bad, _ = dets_match(dets[c], ref)
        assert bad == 0
        assert np.array_equal(np.argwhere(dets[c] == 1), golden["matx_det_idx"])
    nt = host(p.GetnormT())[0, 2:258, 6:8007]
    assert np.array_equal(nt, st["norm"][0])


# ------------------------------------------------------- elementwise / sum
def test_ew_expressions(mx):
    from matx_amd import expr as E
    rng = np.random.default_rng(20)
    a = crandn(rng, 4, 126, 300)
    b = rng.standard_normal(300).astype(np.float32)
    ad, bd = dev(a), dev(b, torch.float32)
    # xc = xc * hamming<1>  (simple_radar_pipeline.h:353-355), in place
    out = ad.clone()
    E.run(out, E.wrap(out) * E.hamming(1, 126))
    assert rel(host(out), a * np.hamming(126)[None, :, None]) < RTOL
    # broadcast real vector, conj, division by a scalar constant
    o2 = torch.empty_like(ad)
    E.run(o2, E.conj(ad) * bd / 3.0 + 1.5)
    assert rel(host(o2), np.conj(a) * b / 3.0 + 1.5) < RTOL
    # abs2 into a real output, comparisons into int32
    p = torch.empty(ad.shape, dtype=torch.float32, device="cuda")
    E.run(p, E.abs2(ad))
    assert rel(host(p), np.abs(a) ** 2) < RTOL
    m = torch.empty(ad.shape, dtype=torch.int32, device="cuda")
    E.run(m, E.abs2(ad) > 2.0)
    assert np.array_equal(host(m), (np.abs(a.astype(complex)) ** 2 > 2.0).astype(np.int32))
    # calcDets body: xpow > norm*(pfa^(-1/norm)-1)*ba  (simple_radar_pipeline.h:86-98)
    nrm = dev(rng.integers(20, 57, (4, 126, 300)).astype(np.float32), torch.float32)
    ba = dev(rng.exponential(1.0, (4, 126, 300)).astype(np.float32), torch.float32)
    d2 = torch.empty(ad.shape, dtype=torch.int32, device="cuda")
    E.run(d2, E.abs2(ad) > E.wrap(nrm) * (E.pow(1e-5, -1.0 / E.wrap(nrm)) - 1.0) * ba)
    n_ = host(nrm).astype(np.float64)
    thr = n_ * (np.power(1e-5, -1.0 / n_) - 1) * host(ba)
    pw = np.abs(a.astype(complex)) ** 2
    near = np.abs(pw - thr) / thr < 1e-5
    assert ((host(d2) != (pw > thr)) & ~near).sum() == 0


def test_sum(mx):
    from matx_amd import expr as E
    rng = np.random.default_rng(21)
    w = crandn(rng, 1000)
    s = host(E.sum(dev(w)))
    assert abs(s - w.astype(complex).sum()) / abs(w.astype(complex).sum()) < 1e-5
    r = rng.standard_normal((7, 333)).astype(np.float32)
    sr = host(E.sum(dev(r, torch.float32)[:, ::2]))
    assert abs(sr - r[:, ::2].astype(np.float64).sum()) < 1e-4 * np.abs(r).sum()
