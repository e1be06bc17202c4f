"""FP64 CPU oracle for the radar hot path -- TEST INFRASTRUCTURE ONLY.

Thin numpy/ctypes wrapper around ``oracle/liborc.so`` (built from
``oracle/mxr_oracle.c`` by ``oracle/Makefile``).  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg import this
module, and only as the checker; the product package ``matx_amd`` never does.

Pipeline semantics follow MatX (paths relative to the MatX repository root):

* ``mode="matx"``  -- MatX's own end-to-end dataflow,
  ``examples/simple_radar_pipeline.h:273-436``: SAME three-pulse canceller over
  all P pulses, Hamming(P-2) on rows 0..P-3 only, rows P-2 and P-1 unwindowed,
  Doppler FFT of ``Nd`` points (MatX: next power of two strictly above P,
  ``:185-188``), CFAR over all ``Nd`` rows.
* ``mode="model"`` -- the numpy golden model,
  ``test/test_vectors/generators/01_radar.py:24-125``: VALID canceller (rows
  1..P-2), Hamming(P-2) on all of them, Doppler FFT of ``Nd`` points.

Every function's pin lives in ``tests/test_oracle.py``.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liborc.so")
_lib = None

FULL, SAME, VALID = 0, 1, 2
_MODES = {"full": FULL, "same": SAME, "valid": VALID}

# MatX CFAR mask, examples/simple_radar_pipeline.h:237-241 (5 Doppler x 13 range,
# 3x3 zero hole around the cell under test).
MATX_CFAR_MASK = np.ones((5, 13), dtype=np.float64)
MATX_CFAR_MASK[1:4, 5:8] = 0.0
MATX_PFA = 1e-5


def build(force: bool = False) -> str:
    """Compile liborc.so with the committed Makefile (gcc)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.c_void_p
        I = ctypes.c_int64
        C = ctypes.c_int
        D = ctypes.c_double
        sigs = {
            "orc_dft": [P, P, I, C],
            "orc_fft": [P, I, C],
            "orc_fft_rows": [P, I, I, I, C],
            "orc_waveform_spectrum": [P, I, I, C, C, P, P],
            "orc_pulse_compress": [P, I, I, I, P, I, I, P, I],
            "orc_fft_conv": [P, I, I, I, P, I, I, P, I],
            "orc_conv1d": [P, I, I, I, P, I, C, P, I],
            "orc_conv2d": [P, I, I, I, P, I, I, C, P],
            "orc_tpc": [P, I, I, I, I, C, P],
            "orc_doppler": [P, I, I, I, I, I, P],
            "orc_cfar": [P, I, I, I, P, I, I, D, P, P, P, P],
            "orc_abs2": [P, I, P],
        }
        for name, argt in sigs.items():
            fn = getattr(L, name)
            fn.argtypes = argt
            fn.restype = None
        L.orc_hamming.argtypes = [I, I]
        L.orc_hamming.restype = D
        _lib = L
    return _lib


def _c128(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.complex128)


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


# ----------------------------------------------------------------- transforms
def dft(x) -> np.ndarray:
    """O(n^2) forward DFT (no scaling)."""
    x = _c128(x)
    y = np.empty_like(x)
    lib().orc_dft(_ptr(x), _ptr(y), x.size, 0)
    return y


def idft_unscaled(x) -> np.ndarray:
    x = _c128(x)
    y = np.empty_like(x)
    lib().orc_dft(_ptr(x), _ptr(y), x.size, 1)
    return y


def fft(x, n: int | None = None, axis: int = -1, inverse: bool = False,
        norm: str = "backward") -> np.ndarray:
    """FFT along ``axis`` with MatX's size rule: ``n`` pads with zeros or
    truncates (fft_common.h:39-140).  norm follows FFTNorm
    (operator_options.h:45-49): "backward" = unscaled forward / 1/n inverse,
    "forward" = 1/n forward / unscaled inverse, "ortho" = 1/sqrt(n) both."""
    x = np.moveaxis(_c128(x), axis, -1)
    m = x.shape[-1]
    n = m if n is None else int(n)
    buf = np.zeros(x.shape[:-1] + (n,), dtype=np.complex128)
    k = min(m, n)
    buf[..., :k] = x[..., :k]
    buf = np.ascontiguousarray(buf)
    rows = int(np.prod(buf.shape[:-1])) if buf.ndim > 1 else 1
    lib().orc_fft_rows(_ptr(buf), rows, n, n, 1 if inverse else 0)
    if norm == "ortho":
        buf /= np.sqrt(n)
    elif (norm == "backward") == inverse:
        buf /= n
    return np.moveaxis(buf, -1, axis)


def ifft(x, n: int | None = None, axis: int = -1, norm: str = "backward"):
    return fft(x, n, axis, inverse=True, norm=norm)


def hamming(n: int) -> np.ndarray:
    L = lib()
    return np.array([L.orc_hamming(i, n) for i in range(n)], dtype=np.float64)


def waveform_spectrum(w, N: int, window: bool = True, normalize: bool = True,
                      return_prepared: bool = False):
    w = _c128(w).ravel()
    Wc = np.empty(N, dtype=np.complex128)
    w2 = np.empty(min(w.size, N), dtype=np.complex128)
    lib().orc_waveform_spectrum(_ptr(w), w.size, N, int(window), int(normalize),
                                _ptr(Wc), _ptr(w2))
    return (Wc, w2) if return_prepared else Wc


def pulse_compress(x, Wc, nout: int) -> np.ndarray:
    """x: (..., S) -> (..., nout); Wc the length-N conjugated spectrum."""
    x = _c128(x)
    Wc = _c128(Wc).ravel()
    S = x.shape[-1]
    rows = int(np.prod(x.shape[:-1])) if x.ndim > 1 else 1
    y = np.empty(x.shape[:-1] + (nout,), dtype=np.complex128)
    lib().orc_pulse_compress(_ptr(x), rows, S, S, _ptr(Wc), Wc.size, nout,
                             _ptr(y), nout)
    return y


def fft_conv(x, w, N: int) -> np.ndarray:
    """out = ifft(fft(x, N) * conj(fft(w, N))) along the last axis."""
    x = _c128(x)
    w = _c128(w).ravel()
    S = x.shape[-1]
    rows = int(np.prod(x.shape[:-1])) if x.ndim > 1 else 1
    y = np.empty(x.shape[:-1] + (N,), dtype=np.complex128)
    lib().orc_fft_conv(_ptr(x), rows, S, S, _ptr(w), w.size, N, _ptr(y), N)
    return y


def conv1d(x, f, mode: str = "full") -> np.ndarray:
    """Direct convolution along the last axis; f is 1-D (broadcast)."""
    x = _c128(x)
    f = _c128(f).ravel()
    m = _MODES[mode]
    n = x.shape[-1]
    olen = {FULL: n + f.size - 1, SAME: n, VALID: n - f.size + 1}[m]
    rows = int(np.prod(x.shape[:-1])) if x.ndim > 1 else 1
    out = np.empty(x.shape[:-1] + (olen,), dtype=np.complex128)
    lib().orc_conv1d(_ptr(x), rows, n, n, _ptr(f), f.size, m, _ptr(out), olen)
    return out


def conv2d(x, f, mode: str = "full") -> np.ndarray:
    """Direct 2-D convolution over the last two axes; f is 2-D."""
    x = _c128(x)
    f = _c128(f)
    m = _MODES[mode]
    H, W = x.shape[-2:]
    fh, fw = f.shape
    oh, ow = {FULL: (H + fh - 1, W + fw - 1), SAME: (H, W),
              VALID: (H - fh + 1, W - fw + 1)}[m]
    batch = int(np.prod(x.shape[:-2])) if x.ndim > 2 else 1
    out = np.empty(x.shape[:-2] + (oh, ow), dtype=np.complex128)
    lib().orc_conv2d(_ptr(x), batch, H, W, _ptr(f), fh, fw, m, _ptr(out))
    return out


# ------------------------------------------------------------------ stages
def tpc(y, mode: str = "same") -> np.ndarray:
    """Three-pulse canceller.  y: (C, P, R) -> (C, P, R) SAME or (C, P-2, R)."""
    y = _c128(y)
    C, P, R = y.shape
    m = _MODES[mode]
    rows = P - 2 if m == VALID else P
    out = np.empty((C, rows, R), dtype=np.complex128)
    lib().orc_tpc(_ptr(y), C, P, R, R, m, _ptr(out))
    return out


def doppler(t, win_rows: int, Nd: int) -> np.ndarray:
    t = _c128(t)
    C, rows, R = t.shape
    D = np.empty((C, Nd, R), dtype=np.complex128)
    lib().orc_doppler(_ptr(t), C, rows, R, win_rows, Nd, _ptr(D))
    return D


def abs2(D) -> np.ndarray:
    D = _c128(D)
    pw = np.empty(D.shape, dtype=np.float64)
    lib().orc_abs2(_ptr(D), D.size, _ptr(pw))
    return pw


def cfar(pw, mask=MATX_CFAR_MASK, pfa: float = MATX_PFA):
    """Returns dict(dets int32, ba, norm, margin) for pw (C, Ny, Nx)."""
    pw = np.ascontiguousarray(pw, dtype=np.float64)
    if pw.ndim == 2:
        pw = pw[None]
    mask = np.ascontiguousarray(mask, dtype=np.float64)
    C, Ny, Nx = pw.shape
    my, mx = mask.shape
    ba = np.empty_like(pw)
    nm = np.empty_like(pw)
    mg = np.empty_like(pw)
    dets = np.empty(pw.shape, dtype=np.int32)
    lib().orc_cfar(_ptr(pw), C, Ny, Nx, _ptr(mask), my, mx, float(pfa),
                   _ptr(ba), _ptr(nm), _ptr(mg), _ptr(dets))
    return {"dets": dets, "ba": ba, "norm": nm, "margin": mg}


def next_pow2(n: int, strict: bool = False) -> int:
    """MatX size rules: non-strict for samples (simple_radar_pipeline.h:180-183),
    strict for pulses (:185-188)."""
    r = 1
    while (r <= n) if strict else (r < n):
        r *= 2
    return r


def radar_pipeline(x, w, Nd: int | None = None, mode: str = "matx",
                   mask=MATX_CFAR_MASK, pfa: float = MATX_PFA,
                   N: int | None = None, stages: bool = False):
    """Full chain on x (C, P, S) with waveform w (L,).

    Returns dets (C, Nd, R) int32, or a dict of every stage when stages=True.
    """
    x = _c128(x)
    if x.ndim == 2:
        x = x[None]
    C, P, S = x.shape
    w = _c128(w).ravel()
    L = w.size
    R = S - L + 1
    N = next_pow2(S) if N is None else N
    if Nd is None:
        Nd = next_pow2(P, strict=True) if mode == "matx" else 256
    Wc, w2 = waveform_spectrum(w, N, True, True, return_prepared=True)
    y = pulse_compress(x.reshape(C * P, S), Wc, R).reshape(C, P, R)
    t = tpc(y, "same" if mode == "matx" else "valid")
    D = doppler(t, P - 2, Nd)
    pw = abs2(D)
    cf = cfar(pw, mask, pfa)
    if not stages:
        return cf["dets"]
    return {"Wc": Wc, "w2": w2, "y": y, "tpc": t, "D": D, "pow": pw, **cf}
