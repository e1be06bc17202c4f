"""GPU parity of conv1d / corr by the FFT method (mxr_conv1d_fft; MatX
MATX_C_METHOD_FFT, include/matx/transforms/conv.h:51-172, corr.h:60-67)
against the FP64 direct convolution of the oracle (pinned to MatX's ConvCorr
KATs and numpy.convolve in tests/test_oracle.py).  corr(x, f) is checked as
conv1d(x, reverse(conj(f))).  Relative L2 <= 1e-5.

Covers the fused path (n + L - 1 <= 16384: real and complex, even and odd
filters, filters longer than the signal, strided batch views) and the long
path (overlap-save correlation plus the fused head), including the size of
MatX's CorrelationConvolutionLargeTest (n = 8*122880 + 2*32768, L = 209).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

RTOL = 1e-5


def rel(a, b):
    a = np.asarray(a, dtype=np.complex128)
    b = np.asarray(b, dtype=np.complex128)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def sig(rng, shape, cplx):
    x = rng.standard_normal(shape)
    if cplx:
        x = x + 1j * rng.standard_normal(shape)
        return x.astype(np.complex64)
    return x.astype(np.float32)


def dev(a):
    return torch.as_tensor(np.ascontiguousarray(a)).cuda()


@pytest.fixture(scope="module")
def mx():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import matx_amd
    matx_amd.lib()
    return matx_amd


def modes(N, L):
    return ("full", "same", "valid") if N >= L else ("full", "same")


@pytest.mark.parametrize("rows,N,L,xc,fc", [
    ((3,), 256, 16, True, True),       # ConvCorr.cu sizes (even and odd filters)
    ((3,), 256, 15, True, False),
    ((2, 2), 128, 8, False, False),    # real in, real filter: real output
    ((4,), 128, 7, False, True),
    ((2,), 16, 256, True, True),       # filter longer than the signal
    ((5,), 1000, 1, True, True),
    ((2,), 5000, 3000, True, True),
    ((1,), 10000, 6385, False, False),  # n + L - 1 = 16384
])
def test_conv1d_corr_fft_fused(mx, orc, rows, N, L, xc, fc):
    rng = np.random.default_rng(N + L)
    x = sig(rng, rows + (N,), xc)
    f = sig(rng, (L,), fc)
    fr = np.conj(f[::-1])
    for mode in modes(N, L):
        got = host(mx.conv1d(dev(x), dev(f), mode, method="fft"))
        ref = orc.conv1d(x, f, mode)
        if not (xc or fc):
            assert got.dtype == np.float32
            ref = ref.real
        assert rel(got, ref) < RTOL, mode
        got = host(mx.corr(dev(x), dev(f), mode, method="fft"))
        ref = orc.conv1d(x, fr, mode)
        assert rel(got, ref.real if not (xc or fc) else ref) < RTOL, mode
        if L <= 4096:  # the direct method's filter limit
            got = host(mx.corr(dev(x), dev(f), mode, method="direct"))
            assert rel(got, ref.real if not (xc or fc) else ref) < RTOL, mode


def test_conv1d_fft_strided_views(mx, orc):
    """A permuted input view (the transform axis strided) and an output view
    into a wider buffer."""
    rng = np.random.default_rng(5)
    x = sig(rng, (3, 300, 4), True)
    f = sig(rng, (33,), True)
    xd = dev(x).permute(0, 2, 1)              # [3, 4, 300], axis stride 4
    buf = torch.zeros((3, 4, 400), dtype=torch.complex64, device="cuda")
    out = buf[:, :, 10:10 + 300]
    mx.conv1d(xd, dev(f), "same", method="fft", out=out)
    ref = orc.conv1d(np.moveaxis(x, 1, 2), f, "same")
    assert rel(host(out), ref) < RTOL
    assert float(buf[:, :, :10].abs().max()) == 0.0


@pytest.mark.parametrize("N,L", [(8 * 122880 + 2 * 32768, 209), (100000, 4096), (40000, 3)])
def test_conv1d_corr_fft_long(mx, orc, N, L):
    """Rows past the fused kernel: overlap-save correlation blocks for the
    outputs at FULL index >= L-1, the fused kernel on x[0, L-1) before."""
    rng = np.random.default_rng(N % 1000 + L)
    x = sig(rng, (N,), True)
    f = sig(rng, (L,), True)
    fr = np.conj(f[::-1])
    xd, fd = dev(x), dev(f)
    for mode in ("full", "same", "valid"):
        assert rel(host(mx.conv1d(xd, fd, mode, method="fft")), orc.conv1d(x, f, mode)) < RTOL, mode
        assert rel(host(mx.corr(xd, fd, mode, method="fft")), orc.conv1d(x, fr, mode)) < RTOL, mode


def test_conv1d_fft_long_batched(mx, orc):
    rng = np.random.default_rng(9)
    x = sig(rng, (3, 20000), True)
    f = sig(rng, (500,), True)
    got = host(mx.conv1d(dev(x), dev(f), "same", method="fft"))
    assert rel(got, orc.conv1d(x, f, "same")) < RTOL


def test_conv1d_fft_errors(mx):
    x = torch.zeros((20000,), dtype=torch.float32, device="cuda")
    f = torch.zeros((9,), dtype=torch.float32, device="cuda")
    with pytest.raises(Exception, match="c64"):
        mx.conv1d(x, f, "full", method="fft")   # long rows need complex operands
    with pytest.raises(Exception, match="VALID"):
        mx.conv1d(f, x[:100], "valid", method="fft",
                  out=torch.empty((0,), dtype=torch.float32, device="cuda"))
    with pytest.raises(ValueError):
        mx.conv1d(f, f, "full", method="winograd")
