"""mxr_sum_noalloc (the entry point that allocates nothing: partials in a
file-scope device array) against mxr_sum through the C ABI: bit-identical
results on two calls in a row, the second with a smaller input, for complex
and real inputs and a strided view; calls on a stream being captured into a
graph replay to the same result."""
import ctypes

import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def L():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from matx_amd import _lib
    _lib.lib()
    return _lib


def both(L, x):
    """(mxr_sum result, mxr_sum_noalloc result) for x on the current stream."""
    a = torch.full((1,), -7, dtype=x.dtype, device="cuda")
    b = torch.full((1,), -9, dtype=x.dtype, device="cuda")
    s = L.stream_handle(None)
    xd = L.desc(x)
    L.check(L.lib().mxr_sum(ctypes.byref(L.desc(a)), ctypes.byref(xd), s))
    L.check(L.lib().mxr_sum_noalloc(ctypes.byref(L.desc(b)), ctypes.byref(xd), s))
    torch.cuda.synchronize()
    return a, b


def test_noalloc_sum_matches_sum_twice(L):
    g = torch.Generator(device="cuda").manual_seed(5)
    big = torch.randn((3, 1000, 700), dtype=torch.complex64, device="cuda", generator=g)
    small = torch.randn((37,), dtype=torch.complex64, device="cuda", generator=g)
    for x in (big, small):               # 512 partials, then 1
        a, b = both(L, x)
        assert torch.equal(a, b)
        ref = x.to(torch.complex128).sum()
        assert abs(complex(a.item()) - complex(ref.item())) <= 1e-5 * float(x.abs().sum())
    r_big = torch.rand((4096, 300), device="cuda", generator=g)
    r_small = torch.rand((5, 9), device="cuda", generator=g)
    for x in (r_big, r_small.t()):       # real, then a transposed (strided) view
        a, b = both(L, x)
        assert torch.equal(a, b)


def test_noalloc_sum_rejects_bad_calls_before_launch(L):
    x = torch.ones((8,), dtype=torch.complex64, device="cuda")
    bad = torch.ones((2,), dtype=torch.complex64, device="cuda")
    r = L.lib().mxr_sum_noalloc(ctypes.byref(L.desc(bad)), ctypes.byref(L.desc(x)),
                                L.stream_handle(None))
    assert r == 4  # the output must hold one element


def test_noalloc_sum_under_graph_capture(L):
    x = torch.randn((64, 513), dtype=torch.complex64, device="cuda")
    out = torch.zeros((1,), dtype=torch.complex64, device="cuda")
    s = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        L.check(L.lib().mxr_sum_noalloc(ctypes.byref(L.desc(out)), ctypes.byref(L.desc(x)),
                                        L.stream_handle(s)))
    ref, _ = both(L, x)
    for k in range(3):
        out.fill_(0)
        both(L, x[:5])                   # eager calls in between use the same scratch
        with torch.cuda.stream(s):
            g.replay()
        s.synchronize()
        assert torch.equal(out, ref), k
