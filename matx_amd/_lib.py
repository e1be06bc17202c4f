"""ctypes binding of libmxr.so (the C ABI declared in include/mxr.h).

The product path has no fallback: if the HIP library is not built, or a call
fails, this module raises.  Tensors are torch device tensors; descriptors carry
their data pointer, shape and element strides (the mxr_tensor struct).
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmxr.so")

MXR_MAX_RANK = 8
C64, F32, I32 = 1, 2, 3
FULL, SAME, VALID = 0, 1, 2
TPC_NONE = 3
NORM_BACKWARD, NORM_FORWARD, NORM_ORTHO = 0, 1, 2
FORWARD, INVERSE = 0, 1
PC_WINDOW, PC_NORMALIZE, PC_WRITEBACK = 1, 2, 4

# Every entry point include/mxr.h declares.
EXPORTS = ("mxr_last_error", "mxr_version", "mxr_clear_caches", "mxr_profile_enable",
           "mxr_profile_reset", "mxr_profile_query", "mxr_tensor_from_dlpack", "mxr_tensor_to_dlpack",
           "mxr_fft_c2c", "mxr_fft_r2c", "mxr_fft_c2r",
           "mxr_fft_conv", "mxr_pulse_compress", "mxr_doppler", "mxr_cfar", "mxr_doppler_cfar",
           "mxr_conv1d", "mxr_conv1d_fft", "mxr_conv2d", "mxr_ew_eval", "mxr_sum", "mxr_sum_noalloc",
           "mxr_pc_plan_create",
           "mxr_pc_plan_set_waveform", "mxr_pc_plan_execute", "mxr_pc_plan_destroy",
           "mxr_fft_conv_adjoint", "mxr_autocorr_adjoint", "mxr_doppler_taps", "mxr_doppler_taps_adjoint",
           "mxr_ragged_plan_create", "mxr_ragged_execute", "mxr_ragged_plan_block_size",
           "mxr_ragged_plan_destroy", "mxr_doppler_adjoint", "mxr_pc_plan_waveform_grad",
           "mxr_isl", "mxr_isl_adjoint", "mxr_psl", "mxr_psl_adjoint", "mxr_phase_isl",
           "mxr_phase_isl_adjoint", "mxr_psl_taps_adjoint",
           "mxr_log_set_level", "mxr_log_get_level", "mxr_log_set_dest", "mxr_log_write",
           "mxr_trace_set_level", "mxr_trace_get_level", "mxr_trace_available", "mxr_trace_push",
           "mxr_trace_pop", "mxr_trace_range_start", "mxr_trace_range_end")

STATUS_NAMES = {0: "SUCCESS", 1: "NOT_SUPPORTED", 2: "INVALID_PARAMETER",
                3: "INVALID_DIM", 4: "INVALID_SIZE", 5: "OUT_OF_MEMORY",
                6: "INVALID_TYPE", 7: "HIP_ERROR", 8: "ALIAS_ERROR"}


class MatxError(RuntimeError):
    """Raised for a non-zero mxr_status (mirrors matxException)."""

    def __init__(self, status: int, msg: str):
        self.status = status
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {msg}")


class mxr_tensor(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("dtype", ctypes.c_int32),
                ("rank", ctypes.c_int32),
                ("shape", ctypes.c_int64 * MXR_MAX_RANK),
                ("stride", ctypes.c_int64 * MXR_MAX_RANK)]


class mxr_dl_tensor(ctypes.Structure):
    """DLPack's DLTensor layout (mxr.h mxr_dl_tensor)."""
    _fields_ = [("data", ctypes.c_void_p), ("device_type", ctypes.c_int32),
                ("device_id", ctypes.c_int32), ("ndim", ctypes.c_int32),
                ("dtype_code", ctypes.c_uint8), ("dtype_bits", ctypes.c_uint8),
                ("dtype_lanes", ctypes.c_uint16), ("shape", ctypes.POINTER(ctypes.c_int64)),
                ("strides", ctypes.POINTER(ctypes.c_int64)), ("byte_offset", ctypes.c_uint64)]


RELEASE_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p)


class mxr_dl_managed(ctypes.Structure):
    pass


mxr_dl_managed._fields_ = [("dl_tensor", mxr_dl_tensor), ("manager_ctx", ctypes.c_void_p),
                           ("deleter", ctypes.CFUNCTYPE(None, ctypes.POINTER(mxr_dl_managed)))]


class mxr_cfar_params(ctypes.Structure):
    _fields_ = [("win_y", ctypes.c_int32), ("win_x", ctypes.c_int32),
                ("guard_y", ctypes.c_int32), ("guard_x", ctypes.c_int32),
                ("pfa", ctypes.c_float)]


_lib = None


class mxr_ragged_job(ctypes.Structure):
    _fields_ = [("x", ctypes.c_void_p), ("C", ctypes.c_int64), ("P", ctypes.c_int64),
                ("S", ctypes.c_int64), ("x_cs", ctypes.c_int64), ("x_ps", ctypes.c_int64),
                ("w", ctypes.c_void_p), ("L", ctypes.c_int64), ("pw", ctypes.c_void_p),
                ("pw_cs", ctypes.c_int64), ("pw_ks", ctypes.c_int64)]


RAGGED_AUTOTUNE = 1
RAGGED_PERSISTENT = 2


def lib():
    """Loads libmxr.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build the HIP library with "
            "`python -m matx_amd.build` (there is no CPU fallback)")
    L = ctypes.CDLL(LIB_PATH)
    T = ctypes.POINTER(mxr_tensor)
    V = ctypes.c_void_p
    I64 = ctypes.c_int64
    I = ctypes.c_int
    L.mxr_last_error.restype = ctypes.c_char_p
    L.mxr_last_error.argtypes = []
    L.mxr_version.restype = I
    L.mxr_clear_caches.restype = I
    L.mxr_profile_enable.argtypes = [I]
    L.mxr_profile_enable.restype = I
    L.mxr_profile_reset.argtypes = []
    L.mxr_profile_reset.restype = I
    L.mxr_profile_query.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_int64),
                                    ctypes.POINTER(ctypes.c_double)]
    L.mxr_profile_query.restype = I
    L.mxr_log_set_level.argtypes = [I]
    L.mxr_log_set_level.restype = I
    L.mxr_log_get_level.argtypes = []
    L.mxr_log_get_level.restype = I
    L.mxr_log_set_dest.argtypes = [ctypes.c_char_p]
    L.mxr_log_set_dest.restype = I
    L.mxr_log_write.argtypes = [I, ctypes.c_char_p]
    L.mxr_log_write.restype = None
    L.mxr_trace_set_level.argtypes = [I]
    L.mxr_trace_set_level.restype = I
    L.mxr_trace_get_level.argtypes = []
    L.mxr_trace_get_level.restype = I
    L.mxr_trace_available.argtypes = []
    L.mxr_trace_available.restype = I
    L.mxr_trace_push.argtypes = [ctypes.c_char_p, I]
    L.mxr_trace_push.restype = I
    L.mxr_trace_pop.argtypes = []
    L.mxr_trace_pop.restype = None
    L.mxr_trace_range_start.argtypes = [ctypes.c_char_p, I, ctypes.POINTER(I)]
    L.mxr_trace_range_start.restype = I
    L.mxr_trace_range_end.argtypes = [I]
    L.mxr_trace_range_end.restype = I
    sigs = {
        "mxr_tensor_from_dlpack": [T, V],
        "mxr_tensor_to_dlpack": [T, ctypes.c_int32, ctypes.c_int32, RELEASE_FN, V,
                                 ctypes.POINTER(ctypes.POINTER(mxr_dl_managed))],
        "mxr_fft_c2c": [T, T, I64, I, I, I, V],
        "mxr_fft_r2c": [T, T, I64, I, I, V],
        "mxr_fft_c2r": [T, T, I64, I, I, V],
        "mxr_fft_conv": [T, T, T, I64, I, V],
        "mxr_pulse_compress": [T, T, T, I64, I, V],
        "mxr_doppler": [T, T, T, I, I64, I64, V],
        "mxr_cfar": [T, T, T, T, ctypes.POINTER(mxr_cfar_params), V],
        "mxr_doppler_cfar": [T, T, T, T, I, I64, I64, ctypes.POINTER(mxr_cfar_params), V],
        "mxr_conv1d": [T, T, T, I, V],
        "mxr_conv1d_fft": [T, T, T, I, I, V],
        "mxr_conv2d": [T, T, T, I, V],
        "mxr_ew_eval": [T, ctypes.c_void_p, I, ctypes.c_void_p, I, ctypes.c_void_p, I, V],
        "mxr_sum": [T, T, V],
        "mxr_sum_noalloc": [T, T, V],
        "mxr_pc_plan_create": [ctypes.POINTER(V), I64, I64, I64, I64, I],
        "mxr_pc_plan_set_waveform": [V, T, V],
        "mxr_pc_plan_execute": [V, T, T, V],
        "mxr_pc_plan_destroy": [V],
        "mxr_fft_conv_adjoint": [T, T, T, T, T, I64, I, V],
        "mxr_autocorr_adjoint": [T, T, T, I64, V],
        "mxr_doppler_taps": [T, T, T, T, I64, V],
        "mxr_doppler_taps_adjoint": [T, T, T, T, T, I64, V],
        "mxr_ragged_plan_create": [ctypes.POINTER(V), ctypes.POINTER(mxr_ragged_job), I, I, V],
        "mxr_ragged_execute": [V, V],
        "mxr_ragged_plan_block_size": [V, I, ctypes.POINTER(I)],
        "mxr_ragged_plan_destroy": [V],
        "mxr_doppler_adjoint": [T, T, T, I, I64, I64, V],
        "mxr_pc_plan_waveform_grad": [V, T, T, T, T, V],
        "mxr_isl": [T, T, V],
        "mxr_isl_adjoint": [T, T, T, V],
        "mxr_phase_isl": [T, T, I64, V],
        "mxr_phase_isl_adjoint": [T, T, T, I64, V],
        "mxr_psl": [T, T, T, I, I, V],
        "mxr_psl_adjoint": [T, T, T, T, I, I, V],
        "mxr_psl_taps_adjoint": [T, T, T, T, T, T, I, I, V],
    }
    for name, argt in sigs.items():
        fn = getattr(L, name)
        fn.argtypes = argt
        fn.restype = I
    _lib = L
    return L


def check(status: int) -> None:
    if status != 0:
        msg = lib().mxr_last_error().decode(errors="replace")
        raise MatxError(status, msg)


def desc(t):
    """mxr_tensor for a torch tensor (or None -> NULL pointer)."""
    if t is None:
        return None
    import torch
    dt = {torch.complex64: C64, torch.float32: F32, torch.int32: I32}.get(t.dtype)
    if dt is None:
        raise MatxError(6, f"unsupported dtype {t.dtype}")
    if t.dim() > MXR_MAX_RANK:
        raise MatxError(3, f"rank {t.dim()} > {MXR_MAX_RANK}")
    d = mxr_tensor()
    d.data = t.data_ptr()
    d.dtype = dt
    d.rank = t.dim()
    for i, (s, st) in enumerate(zip(t.shape, t.stride())):
        d.shape[i] = s
        d.stride[i] = st
    return d


def ref(d):
    return ctypes.byref(d) if d is not None else None


def profile_query(name: str):
    """(launch count, total ms) of recorded launches whose kernel name
    contains `name`."""
    n = ctypes.c_int64()
    t = ctypes.c_double()
    check(lib().mxr_profile_query(name.encode(), ctypes.byref(n), ctypes.byref(t)))
    return n.value, t.value


def desc_from_dlpack(obj):
    """(mxr_tensor, capsule) for a DLPack producer's memory, no copy: obj has
    __dlpack__ (a torch, CuPy ... array) or is a "dltensor" capsule.  The
    descriptor is valid while the capsule (and its producer) are alive."""
    cap = obj.__dlpack__() if hasattr(obj, "__dlpack__") else obj
    get = ctypes.pythonapi.PyCapsule_GetPointer
    get.restype = ctypes.c_void_p
    get.argtypes = [ctypes.py_object, ctypes.c_char_p]
    ptr = get(cap, b"dltensor")
    d = mxr_tensor()
    check(lib().mxr_tensor_from_dlpack(ctypes.byref(d), ptr))
    return d, cap


# Exports whose consumer has not called the deleter yet: export id -> (the
# producer object kept alive, the release callback object).
_exports = {}
_export_seq = [0]


@RELEASE_FN
def _release_export(ctx):
    _exports.pop(ctx, None)


def to_dlpack(t):
    """A "dltensor" capsule (DLPack export through mxr_tensor_to_dlpack, no
    copy) of a device tensor; the tensor is kept alive until the consumer
    calls the capsule's deleter."""
    d = desc(t)
    _export_seq[0] += 1
    key = _export_seq[0]
    _exports[key] = t
    out = ctypes.POINTER(mxr_dl_managed)()
    try:
        check(lib().mxr_tensor_to_dlpack(ctypes.byref(d), 10, t.device.index or 0, _release_export,
                                         ctypes.c_void_p(key), ctypes.byref(out)))
    except Exception:
        _exports.pop(key, None)
        raise
    new = ctypes.pythonapi.PyCapsule_New
    new.restype = ctypes.py_object
    new.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_void_p]
    return new(ctypes.cast(out, ctypes.c_void_p), b"dltensor", None)


def stream_handle(stream=None):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)
