"""Library log and trace ranges over the C ABI (include/mxr.h).

The log mirrors MatX's logger (include/matx/core/log.h:110-401): levels
TRACE 0 .. FATAL 5, OFF 6, start-up values from MXR_LOG_LEVEL / MXR_LOG_DEST.
Trace ranges mirror MatX's NVTX ranges (include/matx/core/nvtx.h:48-168)
through roctx; rocprofv3 --marker-trace records them beside the kernels.
"""
import contextlib
import ctypes

from ._lib import check, lib

TRACE, DEBUG, INFO, WARN, ERROR, FATAL, OFF = range(7)
LEVEL_NONE, LEVEL_USER, LEVEL_API, LEVEL_INTERNAL, LEVEL_ALL = 0, 1, 2, 3, 0xFFFFFF


def set_log_level(level: int) -> None:
    """Messages at `level` and above are written (OFF = 6 silences the log)."""
    check(lib().mxr_log_set_level(int(level)))


def log_level() -> int:
    return int(lib().mxr_log_get_level())


def set_log_dest(dest: str) -> None:
    """"stderr", "stdout" or a file path (appended)."""
    check(lib().mxr_log_set_dest(dest.encode()))


def log(level: int, message: str) -> None:
    lib().mxr_log_write(int(level), message.encode())


def set_trace_level(level: int) -> None:
    """Ranges at `level` or below are emitted (MATX_NVTX_SET_LOG_LEVEL)."""
    check(lib().mxr_trace_set_level(int(level)))


def trace_level() -> int:
    return int(lib().mxr_trace_get_level())


def trace_available() -> bool:
    """True when roctx could be loaded (ranges are no-ops otherwise)."""
    return bool(lib().mxr_trace_available())


@contextlib.contextmanager
def trace_range(name: str, level: int = LEVEL_USER):
    """A nested range around the body (MATX_NVTX_START)."""
    opened = lib().mxr_trace_push(name.encode(), int(level))
    try:
        yield
    finally:
        if opened:
            lib().mxr_trace_pop()


def range_start(name: str, level: int = LEVEL_USER, rid: int = -1) -> int:
    """Opens a range ended by id (MATX_NVTX_START_RANGE); rid < 0 picks a
    fresh id.  Returns the id."""
    i = ctypes.c_int(rid)
    check(lib().mxr_trace_range_start(name.encode(), int(level), ctypes.byref(i)))
    return i.value


def range_end(rid: int) -> None:
    """Ends a range opened by range_start (MATX_NVTX_END_RANGE)."""
    check(lib().mxr_trace_range_end(int(rid)))
