"""CPU-side checks of the native library: it loads, exports every entry point
include/mxr.h declares, and rejects bad calls on the host before any launch
(the MatX size/shape checks that throw matxException, e.g.
test/00_transform/FFT.cu:661-719).  No call here reaches a GPU: every case
fails validation, which runs before any HIP call."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "mxr.h")).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(mxr_\w+)\s*\(", text, re.M)))


@pytest.fixture(scope="module")
def L():
    from matx_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libmxr.so not built (run python -m matx_amd.build)")
    return _lib


def fake(L, shape, dtype, addr=0x10000, strides=None):
    d = L.mxr_tensor()
    d.data = addr
    d.dtype = dtype
    d.rank = len(shape)
    if strides is None:
        strides, s = [], 1
        for n in reversed(shape):
            strides.insert(0, s)
            s *= n
    for i, (n, st) in enumerate(zip(shape, strides)):
        d.shape[i] = n
        d.stride[i] = st
    return d


def test_exports_every_declared_symbol(L):
    syms = declared_symbols()
    assert len(syms) >= 10
    assert set(syms) == set(L.EXPORTS)
    lib = L.lib()
    for s in syms:
        assert hasattr(lib, s), s


def test_version_and_last_error(L):
    assert L.lib().mxr_version() >= 100
    assert isinstance(L.lib().mxr_last_error(), bytes)


def status(L, fn, *args):
    r = fn(*args)
    msg = L.lib().mxr_last_error().decode()
    return r, msg


def test_fft_size_checks(L):
    lib = L.lib()
    tc = fake(L, [16], L.C64)
    # C2C output larger than input without an explicit size: no implicit pad
    t2 = fake(L, [32], L.C64, addr=0x90000)
    r, msg = status(L, lib.mxr_fft_c2c, ctypes.byref(t2), ctypes.byref(tc), 0, 0, 0, 0, None)
    assert r == 4 and "does not match" in msg
    # output smaller than the transform
    t3 = fake(L, [9], L.C64, addr=0x90000)
    r, _ = status(L, lib.mxr_fft_c2c, ctypes.byref(t3), ctypes.byref(tc), 0, 0, 0, 0, None)
    assert r == 4
    # sizes past every path (powers of two to 2^24, others to 2^23) are
    # reported as unsupported before any device work
    n5 = (1 << 23) + 1
    t5 = fake(L, [n5], L.C64, addr=0x90000)
    r, msg = status(L, lib.mxr_fft_c2c, ctypes.byref(t5), ctypes.byref(tc), n5, 0, 0, 0, None)
    assert r == 1 and "unsupported" in msg
    t6 = fake(L, [1 << 25], L.C64, addr=0x90000)
    r, _ = status(L, lib.mxr_fft_c2c, ctypes.byref(t6), ctypes.byref(tc), 1 << 25, 0, 0, 0, None)
    assert r == 1
    # wrong dtype
    tf = fake(L, [16], L.F32, addr=0x90000)
    r, _ = status(L, lib.mxr_fft_c2c, ctypes.byref(tf), ctypes.byref(tc), 0, 0, 0, 0, None)
    assert r == 6
    # partial overlap is an alias error
    ta = fake(L, [16], L.C64, addr=0x10000 + 8)
    r, _ = status(L, lib.mxr_fft_c2c, ctypes.byref(ta), ctypes.byref(tc), 0, 0, 0, 0, None)
    assert r == 8


def test_conv_size_checks(L):
    lib = L.lib()
    x = fake(L, [4], L.F32)
    f = fake(L, [3], L.F32, addr=0x20000)
    bad = fake(L, [5], L.F32, addr=0x30000)
    r, _ = status(L, lib.mxr_conv1d, ctypes.byref(bad), ctypes.byref(x), ctypes.byref(f), 0, None)
    assert r == 4                                # FULL needs 6
    cplx_out = fake(L, [6], L.F32, addr=0x30000)
    fc = fake(L, [3], L.C64, addr=0x20000)
    r, _ = status(L, lib.mxr_conv1d, ctypes.byref(cplx_out), ctypes.byref(x), ctypes.byref(fc), 0,
                  None)
    assert r == 6                                # complex filter needs complex output
    x2 = fake(L, [3, 3], L.F32)
    f2 = fake(L, [2, 2], L.F32, addr=0x20000)
    o2 = fake(L, [3, 3], L.F32, addr=0x30000)
    r, _ = status(L, lib.mxr_conv2d, ctypes.byref(o2), ctypes.byref(x2), ctypes.byref(f2), 0, None)
    assert r == 4                                # FULL needs 4x4


def test_pipeline_stage_checks(L):
    lib = L.lib()
    y = fake(L, [2, 8, 100], L.C64)
    pw = fake(L, [2, 4, 100], L.F32, addr=0x900000)
    # a 4-point Doppler FFT cannot hold 8 canceller rows
    r, msg = status(L, lib.mxr_doppler, ctypes.byref(pw), None, ctypes.byref(y), 1, 6, 4, None)
    assert r == 4 and "shorter" in msg
    r, _ = status(L, lib.mxr_doppler, None, None, ctypes.byref(y), 1, 6, 8, None)
    assert r == 2                                # no output
    inp = fake(L, [2, 8, 100], L.F32)
    dets = fake(L, [2, 8, 100], L.I32, addr=0x900000)
    p = L.mxr_cfar_params(5, 13, 4, 3, 1e-5)     # even guard
    r, _ = status(L, lib.mxr_cfar, ctypes.byref(dets), None, None, ctypes.byref(inp),
                  ctypes.byref(p), None)
    assert r == 2
    p = L.mxr_cfar_params(5, 13, 3, 3, 1.5)      # pfa out of range
    r, _ = status(L, lib.mxr_cfar, ctypes.byref(dets), None, None, ctypes.byref(inp),
                  ctypes.byref(p), None)
    assert r == 2
    x = fake(L, [4, 9000], L.C64)
    w = fake(L, [1000], L.C64, addr=0x9000000)
    out = fake(L, [4, 20000], L.C64, addr=0xA000000)
    r, msg = status(L, lib.mxr_pulse_compress, ctypes.byref(out), ctypes.byref(x), ctypes.byref(w),
                    16384, 3, None)
    assert r == 4 and "exceed" in msg
    r, _ = status(L, lib.mxr_pulse_compress, ctypes.byref(out), ctypes.byref(x), ctypes.byref(w),
                  10000, 3, None)
    assert r == 1                                # n not a power of two


def test_doppler_cfar_checks(L):
    lib = L.lib()
    y = fake(L, [2, 8, 100], L.C64)
    dets = fake(L, [2, 8, 100], L.I32, addr=0x900000)
    p = L.mxr_cfar_params(5, 13, 3, 3, 1e-5)
    # a 4-point Doppler FFT cannot hold 8 canceller rows (dets shaped for nd=4)
    d4 = fake(L, [2, 4, 100], L.I32, addr=0x900000)
    r, msg = status(L, lib.mxr_doppler_cfar, ctypes.byref(d4), None, None, ctypes.byref(y), 1, 6,
                    4, ctypes.byref(p), None)
    assert r == 4 and "shorter" in msg
    r, _ = status(L, lib.mxr_doppler_cfar, ctypes.byref(dets), None, None, ctypes.byref(y), 1, 6,
                  8, None, None)
    assert r == 2                                # null params
    bad = L.mxr_cfar_params(5, 13, 3, 4, 1e-5)   # even guard
    r, _ = status(L, lib.mxr_doppler_cfar, ctypes.byref(dets), None, None, ctypes.byref(y), 1, 6,
                  8, ctypes.byref(bad), None)
    assert r == 2
    d_wrong = fake(L, [2, 8, 99], L.I32, addr=0x900000)
    r, _ = status(L, lib.mxr_doppler_cfar, ctypes.byref(d_wrong), None, None, ctypes.byref(y), 1,
                  6, 8, ctypes.byref(p), None)
    assert r == 4                                # output shape
    d_alias = fake(L, [2, 8, 100], L.I32)        # same address as y
    r, _ = status(L, lib.mxr_doppler_cfar, ctypes.byref(d_alias), None, None, ctypes.byref(y), 1,
                  6, 8, ctypes.byref(p), None)
    assert r == 8


def test_ragged_plan_checks(L):
    import ctypes
    lib = L.lib()
    jobs = (L.mxr_ragged_job * 1)()
    j = jobs[0]
    j.x, j.C, j.P, j.S, j.x_cs, j.x_ps = 0x10000, 2, 16, 1000, 16000, 1000
    j.w, j.L = 0x20000, 1200                      # waveform longer than the pulse
    j.pw, j.pw_cs, j.pw_ks = 0x30000, 16 * 1000, 1000
    h = ctypes.c_void_p()
    r, msg = status(L, lib.mxr_ragged_plan_create, ctypes.byref(h), jobs, 1, 0, None)
    assert r == 4 and "L <= S" in msg
    j.L = 100
    j.pw = 0                                      # no output
    r, _ = status(L, lib.mxr_ragged_plan_create, ctypes.byref(h), jobs, 1, 0, None)
    assert r == 2
    j.pw, j.x_ps = 0x30000, 500                   # pulses overlap
    r, _ = status(L, lib.mxr_ragged_plan_create, ctypes.byref(h), jobs, 1, 0, None)
    assert r == 2
    r, _ = status(L, lib.mxr_ragged_plan_create, ctypes.byref(h), jobs, 1, 8, None)
    assert r == 2                                 # unknown flag
    j.x_ps, j.P, j.pw_cs = 1000, 1, 1000          # one pulse: no Doppler transform
    r, msg = status(L, lib.mxr_ragged_plan_create, ctypes.byref(h), jobs, 1, 0, None)
    assert r == 1 and "P >= 2" in msg
    j.P, j.pw_cs, j.S, j.x_ps, j.x_cs, j.pw_ks = 2, 2 * 100, 100, 100, 200, 100
    r, msg = status(L, lib.mxr_ragged_plan_create, ctypes.byref(h), jobs, 1, 0, None)
    assert r == 1 and "S >= 513" in msg           # rows shorter than every block size


def test_header_matches_python_constants(L):
    text = open(os.path.join(ROOT, "include", "mxr.h")).read()
    assert "MXR_TPC_NONE 3" in text and L.TPC_NONE == 3
    assert "MXR_PC_WRITEBACK 4" in text and L.PC_WRITEBACK == 4


def test_radar_gradient_checks(L):
    lib = L.lib()
    y = fake(L, [2, 8, 100], L.C64)
    gy = fake(L, [2, 8, 100], L.C64, addr=0x900000)
    gp4 = fake(L, [2, 4, 100], L.F32, addr=0xA00000)
    r, msg = status(L, lib.mxr_doppler_adjoint, ctypes.byref(gy), ctypes.byref(gp4),
                    ctypes.byref(y), 1, 6, 4, None)
    assert r == 4 and "shorter" in msg              # 4-point FFT, 8 canceller rows
    gp = fake(L, [2, 8, 100], L.F32, addr=0xA00000)
    gy_bad = fake(L, [2, 8, 99], L.C64, addr=0x900000)
    r, _ = status(L, lib.mxr_doppler_adjoint, ctypes.byref(gy_bad), ctypes.byref(gp),
                  ctypes.byref(y), 1, 6, 8, None)
    assert r == 4                                   # gy must have y's shape
    gy_alias = fake(L, [2, 8, 100], L.C64)
    r, _ = status(L, lib.mxr_doppler_adjoint, ctypes.byref(gy_alias), ctypes.byref(gp),
                  ctypes.byref(y), 1, 6, 8, None)
    assert r == 8
    r, _ = status(L, lib.mxr_doppler_adjoint, ctypes.byref(gy), ctypes.byref(gp),
                  ctypes.byref(y), 7, 6, 8, None)
    assert r == 2                                   # unknown canceller mode
    gw = fake(L, [10], L.C64, addr=0xB00000)
    r, _ = status(L, lib.mxr_pc_plan_waveform_grad, None, ctypes.byref(gw), ctypes.byref(gy),
                  ctypes.byref(y), ctypes.byref(gw), None)
    assert r == 2                                   # null plan


def test_loss_head_checks(L):
    lib = L.lib()
    r = fake(L, [4, 100], L.C64)
    isl_bad = fake(L, [3], L.F32, addr=0x900000)
    r_, _ = status(L, lib.mxr_isl, ctypes.byref(isl_bad), ctypes.byref(r), None)
    assert r_ == 4                                  # one value per row
    r_strided = fake(L, [4, 100], L.C64, strides=[200, 2])
    isl = fake(L, [4], L.F32, addr=0x900000)
    r_, _ = status(L, lib.mxr_isl, ctypes.byref(isl), ctypes.byref(r_strided), None)
    assert r_ == 2                                  # rows need unit stride
    pw = fake(L, [2, 8, 10], L.F32)
    tg = fake(L, [3, 2], L.I32, addr=0x900000)
    out = fake(L, [3], L.F32, addr=0xA00000)
    r_, _ = status(L, lib.mxr_psl, ctypes.byref(out), ctypes.byref(pw), ctypes.byref(tg), 2, 8, None)
    assert r_ == 4                                  # targets must be [T, 3]
    tg = fake(L, [3, 3], L.I32, addr=0x900000)
    r_, _ = status(L, lib.mxr_psl, ctypes.byref(out), ctypes.byref(pw), ctypes.byref(tg), 2, 0, None)
    assert r_ == 2                                  # q >= 1
    gp = fake(L, [2, 8, 10], L.F32, addr=0xB00000, strides=[160, 20, 2])
    g = fake(L, [3], L.F32, addr=0xC00000)
    r_, _ = status(L, lib.mxr_psl_adjoint, ctypes.byref(gp), ctypes.byref(g), ctypes.byref(pw),
                   ctypes.byref(tg), 2, 8, None)
    assert r_ == 2                                  # gpow must be contiguous


def test_dlpack_descriptor(L):
    """mxr_tensor_from_dlpack on hand-built DLTensor structs (DLPack's
    layout): offsets, compact and explicit strides, dtype codes, and the
    refusals (host memory, unsupported dtypes, lanes).  No device calls."""
    lib = L.lib()
    shape = (ctypes.c_int64 * 3)(2, 3, 5)

    def dl(device=10, code=5, bits=64, lanes=1, strides=None, offset=0):
        t = L.mxr_dl_tensor()
        t.data = 0x100000
        t.device_type, t.device_id, t.ndim = device, 0, 3
        t.dtype_code, t.dtype_bits, t.dtype_lanes = code, bits, lanes
        t.shape = ctypes.cast(shape, ctypes.POINTER(ctypes.c_int64))
        if strides is not None:
            arr = (ctypes.c_int64 * 3)(*strides)
            t._keep = arr
            t.strides = ctypes.cast(arr, ctypes.POINTER(ctypes.c_int64))
        t.byte_offset = offset
        return t

    d = L.mxr_tensor()
    assert lib.mxr_tensor_from_dlpack(ctypes.byref(d), ctypes.byref(dl(offset=16))) == 0
    assert d.data == 0x100000 + 16 and d.dtype == L.C64 and d.rank == 3
    assert list(d.shape[:3]) == [2, 3, 5] and list(d.stride[:3]) == [15, 5, 1]
    assert lib.mxr_tensor_from_dlpack(ctypes.byref(d), ctypes.byref(dl(code=2, bits=32, strides=(1, 2, 6)))) == 0
    assert d.dtype == L.F32 and list(d.stride[:3]) == [1, 2, 6]
    assert lib.mxr_tensor_from_dlpack(ctypes.byref(d), ctypes.byref(dl(code=0, bits=32))) == 0
    assert d.dtype == L.I32
    r, msg = status(L, lib.mxr_tensor_from_dlpack, ctypes.byref(d), ctypes.byref(dl(device=1)))
    assert r == 1 and "ROCm" in msg
    r, _ = status(L, lib.mxr_tensor_from_dlpack, ctypes.byref(d), ctypes.byref(dl(code=2, bits=64)))
    assert r == 6
    r, _ = status(L, lib.mxr_tensor_from_dlpack, ctypes.byref(d), ctypes.byref(dl(lanes=2)))
    assert r == 6
    r, _ = status(L, lib.mxr_tensor_from_dlpack, ctypes.byref(d), ctypes.byref(dl(offset=4)))
    assert r == 2


def test_expression_rejects_complex_operands_of_real_ops(L):
    """sqrt / cos / pow / ordering of a complex value read only the real part
    in the kernels, so the host rejects such programs (MXR_INVALID_TYPE)
    instead of returning a silently wrong result; the same ops on a real
    input pass validation (they are not launched here: the output pointer is
    fake, so only the rejecting calls are made)."""
    lib = L.lib()
    out = fake(L, [4, 8], L.C64)
    xc = fake(L, [4, 8], L.C64, addr=0x90000)
    ins = (ctypes.POINTER(L.mxr_tensor) * 1)(ctypes.pointer(xc))
    I = ctypes.c_int32 * 3
    LOAD, CONST, SQRT, COS, POW, GT, ABS2, SELECT = 1, 2, 33, 39, 14, 15, 31, 50
    def prog(*ops):
        arr = (I * len(ops))(*[I(*o) for o in ops])
        return arr, len(ops)
    consts = (ctypes.c_float * 4)(2.0, 0.0, 1.0, 0.5)   # const 0 real, const 1 complex
    cases = [
        [(LOAD, 0, 0), (SQRT, 0, 0)],
        [(LOAD, 0, 0), (COS, 0, 0)],
        [(LOAD, 0, 0), (CONST, 0, 0), (POW, 0, 0)],
        [(CONST, 1, 0), (CONST, 0, 0), (GT, 0, 0)],
        [(LOAD, 0, 0), (LOAD, 0, 0), (LOAD, 0, 0), (SELECT, 0, 0)],
    ]
    for ops in cases:
        p, n = prog(*ops)
        r, msg = status(L, lib.mxr_ew_eval, ctypes.byref(out), ins, 1, p, n, consts, 2, None)
        assert r == 6 and "real operand" in msg, (ops, r, msg)


def test_dlpack_export_descriptor(L):
    """mxr_tensor_to_dlpack (tensor_t::ToDlPack, include/matx/core/
    tensor.h:1471-1555) on a host-built descriptor: the DLManagedTensor
    carries the view's pointer, shape, element strides, dtype and device;
    importing it back gives the same descriptor; the deleter runs the
    producer's release callback exactly once."""
    lib = L.lib()
    released = []

    @L.RELEASE_FN
    def rel(ctx):
        released.append(ctx)

    for dt, code, bits in ((L.C64, 5, 64), (L.F32, 2, 32), (L.I32, 0, 32)):
        t = fake(L, [3, 5, 7], dt, addr=0x200000, strides=[70, 1, 10])   # a permuted view
        out = ctypes.POINTER(L.mxr_dl_managed)()
        assert lib.mxr_tensor_to_dlpack(ctypes.byref(t), 10, 2, rel, ctypes.c_void_p(77 + dt),
                                        ctypes.byref(out)) == 0
        m = out.contents
        d = m.dl_tensor
        assert d.data == 0x200000 and d.device_type == 10 and d.device_id == 2 and d.ndim == 3
        assert (d.dtype_code, d.dtype_bits, d.dtype_lanes) == (code, bits, 1)
        assert [d.shape[i] for i in range(3)] == [3, 5, 7]
        assert [d.strides[i] for i in range(3)] == [70, 1, 10]
        assert d.byte_offset == 0
        back = L.mxr_tensor()
        assert lib.mxr_tensor_from_dlpack(ctypes.byref(back), ctypes.byref(d)) == 0
        assert back.data == t.data and back.dtype == dt and back.rank == 3
        assert list(back.shape[:3]) == [3, 5, 7] and list(back.stride[:3]) == [70, 1, 10]
        assert not released
        m.deleter(out)
        assert released == [77 + dt]
        released.clear()
    # a scalar (rank 0) exports with no shape entries
    s = fake(L, [], L.F32, addr=0x300000)
    out = ctypes.POINTER(L.mxr_dl_managed)()
    assert lib.mxr_tensor_to_dlpack(ctypes.byref(s), 11, 0, rel, None, ctypes.byref(out)) == 0
    assert out.contents.dl_tensor.ndim == 0 and out.contents.dl_tensor.device_type == 11
    out.contents.deleter(out)
    assert released == [None]
    # rejected: a device type that is not ROCm memory, a bad dtype, null output
    r, msg = status(L, lib.mxr_tensor_to_dlpack, ctypes.byref(s), 2, 0, rel, None, ctypes.byref(out))
    assert r == 1 and "device type" in msg
    bad = fake(L, [4], 9)
    r, _ = status(L, lib.mxr_tensor_to_dlpack, ctypes.byref(bad), 10, 0, rel, None, ctypes.byref(out))
    assert r == 6
    r, _ = status(L, lib.mxr_tensor_to_dlpack, ctypes.byref(s), 10, 0, rel, None, None)
    assert r == 2
