"""Library log and trace ranges (MatX's logger, include/matx/core/log.h:
110-401, and NVTX ranges, include/matx/core/nvtx.h:48-168), on the CPU:
environment start-up values, levels, destinations, the failed-call ERROR
line, range filtering by level and range ids."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture
def T():
    from matx_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libmxr.so not built")
    import matx_amd.trace as t
    lvl, tl = t.log_level(), t.trace_level()
    yield t
    t.set_log_level(lvl)
    t.set_log_dest("stderr")
    t.set_trace_level(tl)


def run_child(code, **env):
    e = dict(os.environ, **env)
    return subprocess.run([sys.executable, "-c", code], cwd=ROOT, env=e, capture_output=True,
                          text=True, timeout=120)


# a call that fails validation before touching the GPU: a null descriptor
FAIL = ("import ctypes, matx_amd._lib as L; "
        "L.lib().mxr_fft_c2c(None, None, 16, 0, 0, 0, None)")


def test_env_level_and_file_dest(tmp_path):
    log = tmp_path / "mxr.log"
    r = run_child(FAIL + "; L.lib().mxr_log_write(2, b'caller info'); L.lib().mxr_log_write(0, b'hidden')",
                  MXR_LOG_LEVEL="INFO", MXR_LOG_DEST=str(log))
    assert r.returncode == 0, r.stderr
    text = log.read_text()
    assert "[mxr ERROR]" in text and "null descriptor" in text
    assert "[mxr INFO] caller info" in text
    assert "hidden" not in text           # TRACE is below INFO
    assert r.stdout == ""                 # nothing on stdout


def test_numeric_level_and_default_off(tmp_path):
    log = tmp_path / "mxr.log"
    r = run_child(FAIL + "; L.lib().mxr_log_write(3, b'a warning')", MXR_LOG_LEVEL="4",
                  MXR_LOG_DEST=str(log))
    assert r.returncode == 0, r.stderr
    text = log.read_text()
    assert "[mxr ERROR]" in text and "a warning" not in text
    # no MXR_LOG_LEVEL: the log is off, nothing on stderr either
    env = {k: v for k, v in os.environ.items() if k != "MXR_LOG_LEVEL"}
    r = subprocess.run([sys.executable, "-c", FAIL], cwd=ROOT, env=env, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0 and "[mxr" not in r.stderr and r.stdout == ""


def test_api_level_and_dest(T, tmp_path):
    from matx_amd import MatxError
    log = tmp_path / "api.log"
    T.set_log_dest(str(log))
    T.set_log_level(T.WARN)
    assert T.log_level() == T.WARN
    T.log(T.INFO, "below the level")
    T.log(T.WARN, "at the level")
    text = log.read_text()
    assert "at the level" in text and "below the level" not in text
    with pytest.raises(MatxError):
        T.set_log_level(7)
    with pytest.raises(MatxError):
        T.set_log_dest(str(tmp_path / "no" / "such" / "dir" / "x.log"))


def test_trace_levels_filter_ranges(T):
    from matx_amd import lib
    T.set_trace_level(T.LEVEL_NONE)
    assert lib().mxr_trace_push(b"off", T.LEVEL_USER) == 0
    T.set_trace_level(T.LEVEL_USER)
    assert lib().mxr_trace_push(b"internal", T.LEVEL_INTERNAL) == 0
    if T.trace_available():
        assert lib().mxr_trace_push(b"user", T.LEVEL_USER) == 1
        lib().mxr_trace_pop()
    with T.trace_range("scoped"):
        pass


def test_trace_range_ids(T):
    from matx_amd import MatxError
    T.set_trace_level(T.LEVEL_ALL)
    rid = T.range_start("by id", T.LEVEL_USER, 41)
    assert rid == 41
    if T.trace_available():
        with pytest.raises(MatxError):
            T.range_start("again", T.LEVEL_USER, 41)
    T.range_end(41)
    T.range_end(41)                       # ending a closed id does nothing
    a = T.range_start("auto")
    b = T.range_start("auto")
    assert a >= 1 << 24 and b == a + 1
    T.range_end(a)
    T.range_end(b)
    with pytest.raises(MatxError):
        T.set_trace_level(-1)


def test_env_trace_level():
    r = run_child("import matx_amd.trace as t; print(t.trace_level())", MXR_TRACE_LEVEL="INTERNAL")
    assert r.returncode == 0, r.stderr
    assert r.stdout.strip() == "3"
    r = run_child("import matx_amd.trace as t; print(t.trace_level())")
    assert r.stdout.strip() == "2"        # API, MatX's default
