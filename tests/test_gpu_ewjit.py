"""GPU parity of the expression compiler (mxr_ew_eval's default path: one
generated HIP kernel per expression shape, compiled with hipRTC; MatX's
fused elementwise kernels, include/matx/executors/kernel.h, and its NVRTC
string generator, executors/jit_cuda.h:325-420).

Every opcode on complex, real and broadcast operands is checked against
numpy in FP64 (relative L2 <= 1e-5; comparisons exact away from ties), the
profiler confirms that the generated kernel ran, and the interpreter kernel
(MXR_EW_JIT=0, in a child process) gives the same values."""
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

RTOL = 1e-5
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def rel(a, b):
    a = np.asarray(a, dtype=np.complex128)
    b = np.asarray(b, dtype=np.complex128)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def inputs():
    rng = np.random.default_rng(31)
    a = ((rng.standard_normal((3, 50, 70)) + 1j * rng.standard_normal((3, 50, 70))) / 2).astype(np.complex64)
    b = ((rng.standard_normal(70) + 1j * rng.standard_normal(70)) / 2).astype(np.complex64)
    r = rng.uniform(0.2, 3.0, (3, 50, 70)).astype(np.float32)
    return a, b, r


def cases():
    """(name, build(E, a, b, r), numpy(a, b, r), output dtype)."""
    N = lambda E, op, *xs: E.Node(op, [E.wrap(x) for x in xs])  # noqa: E731
    ham = 0.54 - 0.46 * np.cos(2 * np.pi * np.arange(70) / 69)
    return [
        ("add_cc", lambda E, a, b, r: E.wrap(a) + b, lambda a, b, r: a + b, torch.complex64),
        ("sub_cr", lambda E, a, b, r: E.wrap(a) - r, lambda a, b, r: a - r, torch.complex64),
        ("mul_cc", lambda E, a, b, r: E.wrap(a) * b, lambda a, b, r: a * b, torch.complex64),
        ("mul_rc", lambda E, a, b, r: E.wrap(r) * a, lambda a, b, r: r * a, torch.complex64),
        ("div_cc", lambda E, a, b, r: E.wrap(a) / b, lambda a, b, r: a / b, torch.complex64),
        ("div_cr", lambda E, a, b, r: E.wrap(a) / r, lambda a, b, r: a / r, torch.complex64),
        ("div_rr", lambda E, a, b, r: E.wrap(r) / 2.5, lambda a, b, r: r / 2.5, torch.float32),
        ("pow", lambda E, a, b, r: E.pow(r, 1.5), lambda a, b, r: r ** 1.5, torch.float32),
        ("gt", lambda E, a, b, r: E.abs2(a) > r, lambda a, b, r: np.abs(a) ** 2 > r, torch.int32),
        ("lt", lambda E, a, b, r: E.abs2(a) < r, lambda a, b, r: np.abs(a) ** 2 < r, torch.int32),
        ("ge", lambda E, a, b, r: E.wrap(r) >= 1.0, lambda a, b, r: r >= 1.0, torch.int32),
        ("le", lambda E, a, b, r: E.wrap(r) <= 1.0, lambda a, b, r: r <= 1.0, torch.int32),
        ("eq", lambda E, a, b, r: N(E, "==", a, a), lambda a, b, r: np.ones(a.shape), torch.int32),
        ("max", lambda E, a, b, r: N(E, "max", r, 1.0), lambda a, b, r: np.maximum(r, 1.0), torch.float32),
        ("min", lambda E, a, b, r: N(E, "min", r, 1.0), lambda a, b, r: np.minimum(r, 1.0), torch.float32),
        ("conj", lambda E, a, b, r: E.conj(a) * b, lambda a, b, r: np.conj(a) * b, torch.complex64),
        ("abs2", lambda E, a, b, r: E.abs2(a), lambda a, b, r: np.abs(a) ** 2, torch.float32),
        ("abs", lambda E, a, b, r: E.abs(a), lambda a, b, r: np.abs(a), torch.float32),
        ("abs_r", lambda E, a, b, r: E.abs(-E.wrap(r)), lambda a, b, r: r, torch.float32),
        ("sqrt", lambda E, a, b, r: E.sqrt(r), lambda a, b, r: np.sqrt(r), torch.float32),
        ("neg", lambda E, a, b, r: -E.wrap(a), lambda a, b, r: -a, torch.complex64),
        ("real", lambda E, a, b, r: E.real(a), lambda a, b, r: a.real, torch.float32),
        ("imag", lambda E, a, b, r: E.imag(a), lambda a, b, r: a.imag, torch.float32),
        ("imag_r", lambda E, a, b, r: E.imag(r) + 1.0, lambda a, b, r: np.ones(r.shape), torch.float32),
        ("exp_c", lambda E, a, b, r: E.exp(a), lambda a, b, r: np.exp(a.astype(np.complex128)), torch.complex64),
        ("exp_r", lambda E, a, b, r: E.exp(r), lambda a, b, r: np.exp(r), torch.float32),
        ("log_c", lambda E, a, b, r: N(E, "log", a), lambda a, b, r: np.log(a.astype(np.complex128)), torch.complex64),
        # log of a negative real: log|x| + i pi, the principal value at
        # (x, +0) (r * -1 keeps the imaginary part +0 on both paths; a
        # complex negation would give -0, the other side of the cut)
        ("log_neg", lambda E, a, b, r: N(E, "log", E.wrap(r) * -1.0),
         lambda a, b, r: np.log((-r).astype(np.complex128)), torch.complex64),
        ("cos", lambda E, a, b, r: N(E, "cos", r), lambda a, b, r: np.cos(r), torch.float32),
        ("sin", lambda E, a, b, r: N(E, "sin", r), lambda a, b, r: np.sin(r), torch.float32),
        ("select", lambda E, a, b, r: E.select(E.wrap(r) > 1.0, a, b),
         lambda a, b, r: np.where(r > 1.0, a, np.broadcast_to(b, a.shape)), torch.complex64),
        ("hamming", lambda E, a, b, r: E.wrap(a) * E.hamming(2, 70), lambda a, b, r: a * ham, torch.complex64),
        ("index", lambda E, a, b, r: E.wrap(r) * E.index(1), lambda a, b, r: r * np.arange(50)[None, :, None],
         torch.float32),
        ("cconst", lambda E, a, b, r: E.wrap(r) * (0.5 - 2j) + a, lambda a, b, r: r * (0.5 - 2j) + a,
         torch.complex64),
        # calcDets (simple_radar_pipeline.h:86-98) with norms n = 8 r + 20 and
        # averages r / 20: thresholds in the range of |a|^2, so both outcomes
        # occur; cells within 1e-5 of the threshold are not compared
        ("calcdets", lambda E, a, b, r: E.abs2(a) > (E.wrap(r) * 8.0 + 20.0) *
         (E.pow(1e-5, -1.0 / (E.wrap(r) * 8.0 + 20.0)) - 1.0) * (E.wrap(r) / 20.0),
         lambda a, b, r: (np.abs(a) ** 2, (8 * r + 20) * (np.power(1e-5, -1.0 / (8 * r + 20)) - 1) * r / 20),
         torch.int32),
    ]


def evaluate():
    """Runs every case on the GPU; returns {name: host array}."""
    from matx_amd import expr as E
    a, b, r = inputs()
    ad, bd, rd = (torch.as_tensor(x).cuda() for x in (a, b, r))
    out = {}
    for name, build, _, dt in cases():
        o = torch.empty(a.shape, dtype=dt, device="cuda")
        E.run(o, build(E, ad, bd, rd))
        torch.cuda.synchronize()
        out[name] = o.cpu().numpy()
    return out


@pytest.fixture(scope="module")
def mx():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import matx_amd
    matx_amd.lib()
    return matx_amd


def test_pow_constant_base_accuracy(mx):
    """pow with a constant base (the calcDets form pow(pfa, -1/norm)) takes
    exp2(y log2(base)) with log2 in double: every element within 4 float
    ulps of the FP64 value across exponents up to the float range, and the
    special cases (results outside the normal range, NaN, base 1, zero and
    negative bases) as powf gives them."""
    from matx_amd import expr as E
    rng = np.random.default_rng(7)
    y = np.concatenate([-1.0 / rng.uniform(1.0, 65.0, 4000),          # the CFAR exponents
                        rng.uniform(-10.0, 10.0, 4000),
                        rng.uniform(-125.0 / 16.6, 125.0 / 16.6, 4000)]).astype(np.float32)
    yd = torch.as_tensor(y).cuda()
    for base in (1e-5, 2.5, 0.75, 3.0e4):
        o = torch.empty_like(yd)
        E.run(o, E.pow(float(base), yd))
        got = o.cpu().numpy().astype(np.float64)
        ref = np.power(np.float64(np.float32(base)), y.astype(np.float64))
        ok = (ref > 1.2e-38) & (ref < 3.0e38)
        ulp = np.spacing(ref[ok].astype(np.float32)).astype(np.float64)
        assert np.max(np.abs(got[ok] - ref[ok]) / ulp) <= 4.0, base
    sp = np.array([np.nan, 0.0, 1.0, -1.0, 200.0, -200.0, 1e30, -1e30], np.float32)
    spd = torch.as_tensor(sp).cuda()
    for base in (1e-5, 2.5, 1.0, 0.0, -2.0):
        o = torch.empty_like(spd)
        E.run(o, E.pow(float(base), spd))
        ref = torch.pow(torch.tensor(float(base), dtype=torch.float32), torch.as_tensor(sp)).numpy()
        got = o.cpu().numpy()
        assert np.array_equal(np.isnan(got), np.isnan(ref)), base
        fin = ~np.isnan(ref)
        assert np.allclose(got[fin], ref[fin], rtol=1e-6, atol=0.0), (base, got, ref)


def test_every_opcode_vs_fp64(mx):
    from matx_amd import _lib
    lib = _lib.lib()
    _lib.check(lib.mxr_profile_reset())
    _lib.check(lib.mxr_profile_enable(1))
    got = evaluate()
    _lib.check(lib.mxr_profile_enable(0))
    cnt, _ = _lib.profile_query("ew_jit_kernel")
    assert cnt >= len(cases()), "the generated kernels did not run"
    a, b, r = inputs()
    for name, _, ref, dt in cases():
        want = ref(a.astype(np.complex128) if np.iscomplexobj(a) else a, b.astype(np.complex128),
                   r.astype(np.float64))
        if isinstance(want, tuple):  # (value, threshold): value > threshold
            v, t = want
            near = np.abs(v - t) <= 1e-5 * np.abs(t)
            assert ((got[name] != (v > t)) & ~near).sum() == 0, name
            assert 0 < got[name].sum() < got[name].size, name
        elif dt == torch.int32:
            assert np.array_equal(got[name], np.asarray(want).astype(np.int32)), name
        else:
            assert rel(got[name], want) < RTOL, name


def test_matches_the_interpreter(mx, tmp_path):
    got = evaluate()
    path = tmp_path / "interp.npz"
    code = ("import sys, numpy as np; sys.path.insert(0, %r); "
            "import tests.test_gpu_ewjit as t; np.savez(%r, **t.evaluate())" % (ROOT, str(path)))
    env = dict(os.environ, MXR_EW_JIT="0")
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True,
                       timeout=120, cwd=ROOT)
    assert r.returncode == 0, r.stderr[-2000:]
    ref = np.load(path)
    for name, _, _, dt in cases():
        if dt == torch.int32:
            assert np.array_equal(got[name], ref[name]), name
        else:
            assert rel(got[name], ref[name]) < 1e-6, name


def test_new_constant_reuses_the_kernel(mx):
    """Constants are kernel arguments: a second value gives the right result
    through the cached kernel."""
    from matx_amd import expr as E
    a, b, r = inputs()
    rd = torch.as_tensor(r).cuda()
    o = torch.empty_like(rd)
    for c in (2.0, -3.5):
        E.run(o, E.wrap(rd) * c + 1.0)
        assert rel(o.cpu().numpy(), r.astype(np.float64) * c + 1.0) < RTOL
