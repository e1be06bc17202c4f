"""MatX operator semantics over torch device tensors, executed by libmxr.so.

Each function mirrors the MatX operator of the same name (paths relative to
the MatX repository root) and calls the HIP library through its C ABI.  Output
tensors may be passed in (`out=`) to write into existing views -- MatX's
``(lhs = op).run(exec)`` -- otherwise they are allocated on the input's device.
Everything is enqueued on the current torch stream (or ``stream=``).
"""
from __future__ import annotations

import torch

from . import _lib as L

L_ = L  # alias for scopes where L is a waveform length

_MODES = {"full": L.FULL, "same": L.SAME, "valid": L.VALID}
_NORMS = {"backward": L.NORM_BACKWARD, "forward": L.NORM_FORWARD, "ortho": L.NORM_ORTHO}


def _mode(m):
    return _MODES[m] if isinstance(m, str) else int(m)


def empty_rows(shape, dtype, device, align: int = 64):
    """Uninitialised tensor of `shape` whose rows (last dimension) start every
    roundup(shape[-1], align) elements: a view into a padded allocation, so
    that each row begins on a 512-byte (c64) or 256-byte (f32) boundary.  The
    pulse-axis kernels read and write 16-range-bin column strips; on rows of
    odd length every strip straddles two cache lines."""
    shape = tuple(int(s) for s in shape)
    n = shape[-1]
    ld = (n + align - 1) // align * align
    return torch.empty(shape[:-1] + (ld,), dtype=dtype, device=device)[..., :n]


def _fft(x, n, axis, norm, out, inverse, stream):
    ax = axis % x.dim()
    n = int(n) if n else x.shape[ax]
    if out is None:
        shape = list(x.shape)
        shape[ax] = n
        out = torch.empty(shape, dtype=torch.complex64, device=x.device)
    L.check(L.lib().mxr_fft_c2c(L.ref(L.desc(out)), L.ref(L.desc(x)), n,
                                L.INVERSE if inverse else L.FORWARD,
                                _NORMS[norm], ax, L.stream_handle(stream)))
    return out


def fft(x, n: int = 0, axis: int = -1, norm: str = "backward", out=None, stream=None):
    """``fft(x, n)`` along ``axis`` (include/matx/operators/fft.h:538-577):
    zero-pads or truncates to n; forward unscaled under the default norm."""
    return _fft(x, n, axis, norm, out, False, stream)


def ifft(x, n: int = 0, axis: int = -1, norm: str = "backward", out=None, stream=None):
    """``ifft`` (include/matx/operators/fft.h:649-680), 1/n under "backward"."""
    return _fft(x, n, axis, norm, out, True, stream)


def rfft(x, n: int = 0, axis: int = -1, norm: str = "backward", out=None, stream=None):
    """Real-to-complex ``rfft(x, n)`` (include/matx/operators/fft.h:580-626):
    f32 in, n/2 + 1 complex bins out along ``axis`` (mxr_fft_r2c)."""
    ax = axis % x.dim()
    n = int(n) if n else x.shape[ax]
    if out is None:
        shape = list(x.shape)
        shape[ax] = n // 2 + 1
        out = torch.empty(shape, dtype=torch.complex64, device=x.device)
    L.check(L.lib().mxr_fft_r2c(L.ref(L.desc(out)), L.ref(L.desc(x)), n, _NORMS[norm], ax,
                                L.stream_handle(stream)))
    return out


def irfft(x, n: int = 0, axis: int = -1, norm: str = "backward", out=None, stream=None):
    """Complex-to-real ``irfft(x, n)`` (include/matx/operators/fft.h:686-740):
    bins 0..n/2 in, n real samples out (n defaults to 2 (N - 1)) (mxr_fft_c2r)."""
    ax = axis % x.dim()
    n = int(n) if n else 2 * (x.shape[ax] - 1)
    if out is None:
        shape = list(x.shape)
        shape[ax] = n
        out = torch.empty(shape, dtype=torch.float32, device=x.device)
    L.check(L.lib().mxr_fft_c2r(L.ref(L.desc(out)), L.ref(L.desc(x)), n, _NORMS[norm], ax,
                                L.stream_handle(stream)))
    return out


def fft2(x, norm: str = "backward", out=None, stream=None):
    """2-D FFT over the last two axes (fft2, include/matx/operators/fft.h:
    743-1130): the transform along the rows, then along the columns of the
    result, in place (the column pass reads its strided axis lane-parallel
    over the contiguous row index)."""
    out = _fft(x, 0, -1, norm, out, False, stream)
    return _fft(out, 0, -2, norm, out, False, stream)


def ifft2(x, norm: str = "backward", out=None, stream=None):
    """2-D inverse FFT over the last two axes (ifft2)."""
    out = _fft(x, 0, -1, norm, out, True, stream)
    return _fft(out, 0, -2, norm, out, True, stream)


def rfft2(x, norm: str = "backward", out=None, stream=None):
    """2-D real-to-complex FFT (rfft2, include/matx/operators/fft.h:1236-1272):
    rfft along the last axis (n/2 + 1 bins), then the complex transform
    along the second-to-last."""
    out = rfft(x, 0, -1, norm, out, stream)
    return _fft(out, 0, -2, norm, out, False, stream)


def irfft2(x, n: int = 0, norm: str = "backward", out=None, stream=None):
    """2-D complex-to-real inverse FFT (irfft2, include/matx/operators/
    fft.h:1275-1310): the complex inverse along the second-to-last axis into
    a temporary, then irfft along the last (n samples, default 2 (N - 1))."""
    t = _fft(x, 0, -2, norm, None, True, stream)
    return irfft(t, n, -1, norm, out, stream)


def fft_conv(x, w, n: int, conj: bool = True, nout: int | None = None, out=None, stream=None):
    """out = ifft(fft(x, n) * conj(fft(w, n))) along the last axis (lags
    0..nout-1, default all n); conj=False gives circular convolution."""
    nout = n if nout is None else nout
    if out is None:
        out = torch.empty(x.shape[:-1] + (nout,), dtype=torch.complex64, device=x.device)
    L.check(L.lib().mxr_fft_conv(L.ref(L.desc(out)), L.ref(L.desc(x)), L.ref(L.desc(w)),
                                 int(n), 1 if conj else 0, L.stream_handle(stream)))
    return out


def pulse_compress(x, w, n: int, nout: int | None = None, window: bool = True,
                   normalize: bool = True, writeback: bool = False, out=None, stream=None):
    """PulseCompression stage (examples/simple_radar_pipeline.h:273-301):
    matched filter of every row of x against hamming-windowed, unit-energy w,
    keeping lags 0..nout-1 (default S-L+1).  ``out=x`` works in place."""
    S = x.shape[-1]
    nout = S - w.shape[0] + 1 if nout is None else nout
    if out is None:
        out = torch.empty(x.shape[:-1] + (nout,), dtype=torch.complex64, device=x.device)
    flags = (L.PC_WINDOW if window else 0) | (L.PC_NORMALIZE if normalize else 0) | \
            (L.PC_WRITEBACK if writeback else 0)
    L.check(L.lib().mxr_pulse_compress(L.ref(L.desc(out)), L.ref(L.desc(x)), L.ref(L.desc(w)),
                                       int(n), flags, L.stream_handle(stream)))
    return out


class PulseCompressionPlan:
    """Pulse compression with the waveform prepared once (mxr_pc_plan_*):
    rows of S samples, waveforms of L samples, transform size n, nout lags
    kept.  set_waveform(w) computes the windowed, unit-energy conj spectrum
    (examples/simple_radar_pipeline.h:276-296); calling the plan filters a
    cube with it, giving mxr_pulse_compress's result."""

    def __init__(self, S: int, L: int, n: int, nout: int | None = None, window: bool = True,
                 normalize: bool = True):
        self.S, self.L, self.n = S, L, n
        self.nout = S - L + 1 if nout is None else nout
        flags = (L_.PC_WINDOW if window else 0) | (L_.PC_NORMALIZE if normalize else 0)
        h = L_.ctypes.c_void_p()
        L_.check(L_.lib().mxr_pc_plan_create(L_.ctypes.byref(h), S, L, n, self.nout, flags))
        self._h = h

    def set_waveform(self, w, stream=None):
        L_.check(L_.lib().mxr_pc_plan_set_waveform(self._h, L_.ref(L_.desc(w)),
                                                   L_.stream_handle(stream)))
        self._w = w  # keep the waveform alive until the spectrum is computed

    def __call__(self, x, out=None, stream=None):
        if out is None:
            out = torch.empty(x.shape[:-1] + (self.nout,), dtype=torch.complex64, device=x.device)
        L_.check(L_.lib().mxr_pc_plan_execute(self._h, L_.ref(L_.desc(out)), L_.ref(L_.desc(x)),
                                              L_.stream_handle(stream)))
        return out

    def waveform_grad(self, gy, x, w, out=None, stream=None):
        """dL/dw of a real loss through y = self(x) with waveform w, summed
        over every row (mxr_pc_plan_waveform_grad); gy = dL/dy."""
        if out is None:
            out = torch.empty((self.L,), dtype=torch.complex64, device=x.device)
        L_.check(L_.lib().mxr_pc_plan_waveform_grad(self._h, L_.ref(L_.desc(out)),
                                                    L_.ref(L_.desc(gy)), L_.ref(L_.desc(x)),
                                                    L_.ref(L_.desc(w)), L_.stream_handle(stream)))
        return out

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            L_.check(L_.lib().mxr_pc_plan_destroy(self._h))
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def doppler(y, nd: int, tpc: str = "same", win_rows: int | None = None, rd=None, pw=None,
            want_rd: bool = False, want_pow: bool = True, stream=None):
    """Three-pulse canceller + Hamming window + nd-point Doppler FFT (+ |.|^2).
    tpc: "same" (MatX stage dataflow), "valid" (numpy model) or "none"
    (input already cancelled).  Returns (pow, rd)."""
    C, P, R = y.shape
    mode = {"same": L.SAME, "valid": L.VALID, "none": L.TPC_NONE}[tpc]
    rows = P - 2 if mode == L.VALID else P
    win_rows = (P - 2 if mode != L.TPC_NONE else rows) if win_rows is None else win_rows
    if pw is None and want_pow:
        pw = torch.empty((C, nd, R), dtype=torch.float32, device=y.device)
    if rd is None and want_rd:
        rd = torch.empty((C, nd, R), dtype=torch.complex64, device=y.device)
    L.check(L.lib().mxr_doppler(L.ref(L.desc(pw)), L.ref(L.desc(rd)), L.ref(L.desc(y)), mode,
                                int(win_rows), int(nd), L.stream_handle(stream)))
    return pw, rd


def doppler_adjoint(gpow, y, nd: int, tpc: str = "same", win_rows: int | None = None, gy=None,
                    stream=None):
    """dL/dy of a real loss through pow = doppler(y, nd, tpc, win_rows)[0]
    (mxr_doppler_adjoint); gpow = dL/dpow may be any strided view (an
    expanded one included).  The Doppler spectrum is recomputed from y."""
    C, P, R = y.shape
    mode = {"same": L.SAME, "valid": L.VALID, "none": L.TPC_NONE}[tpc]
    rows = P - 2 if mode == L.VALID else P
    win_rows = (P - 2 if mode != L.TPC_NONE else rows) if win_rows is None else win_rows
    if gy is None:
        gy = torch.empty_like(y)
    L.check(L.lib().mxr_doppler_adjoint(L.ref(L.desc(gy)), L.ref(L.desc(gpow)), L.ref(L.desc(y)),
                                        mode, int(win_rows), int(nd), L.stream_handle(stream)))
    return gy


def cfar(inp, win=(5, 13), guard=(3, 3), pfa: float = 1e-5, dets=None, ba=None, pw=None,
         want_ba: bool = False, stream=None):
    """CA-CFAR (examples/simple_radar_pipeline.h:394-436): inp is f32 power or
    c64 (|.|^2 on load).  Returns (dets, ba, pow)."""
    if inp.dim() == 2:
        inp = inp.unsqueeze(0)
    if dets is None:
        dets = torch.empty(inp.shape, dtype=torch.int32, device=inp.device)
    if ba is None and want_ba:
        ba = torch.empty(inp.shape, dtype=torch.float32, device=inp.device)
    p = L.mxr_cfar_params(win[0], win[1], guard[0], guard[1], float(pfa))
    L.check(L.lib().mxr_cfar(L.ref(L.desc(dets)), L.ref(L.desc(ba)), L.ref(L.desc(pw)),
                             L.ref(L.desc(inp)), L.ctypes.byref(p), L.stream_handle(stream)))
    return dets, ba, pw


def doppler_cfar(y, nd: int, tpc: str = "same", win_rows: int | None = None, win=(5, 13),
                 guard=(3, 3), pfa: float = 1e-5, dets=None, ba=None, pw=None,
                 want_ba: bool = False, want_pow: bool = False, stream=None):
    """Canceller + window + nd-point Doppler FFT + |.|^2 + CA-CFAR in one
    pass (ThreePulseCanceller + DopplerProcessing + CFARDetections,
    examples/simple_radar_pipeline.h:321-357,394-436).  Returns (dets, ba,
    pow); the power map is written only when requested."""
    C, P, R = y.shape
    mode = {"same": L.SAME, "valid": L.VALID, "none": L.TPC_NONE}[tpc]
    rows = P - 2 if mode == L.VALID else P
    win_rows = (P - 2 if mode != L.TPC_NONE else rows) if win_rows is None else win_rows
    if dets is None:
        dets = torch.empty((C, nd, R), dtype=torch.int32, device=y.device)
    if ba is None and want_ba:
        ba = torch.empty((C, nd, R), dtype=torch.float32, device=y.device)
    if pw is None and want_pow:
        pw = torch.empty((C, nd, R), dtype=torch.float32, device=y.device)
    p = L.mxr_cfar_params(win[0], win[1], guard[0], guard[1], float(pfa))
    L.check(L.lib().mxr_doppler_cfar(L.ref(L.desc(dets)), L.ref(L.desc(ba)), L.ref(L.desc(pw)),
                                     L.ref(L.desc(y)), mode, int(win_rows), int(nd),
                                     L.ctypes.byref(p), L.stream_handle(stream)))
    return dets, ba, pw


def fft_conv_adjoint(gy, x, w, n: int, conj: bool = True, want_gx: bool = True,
                     want_gw: bool = True, stream=None):
    """Gradients of a real loss through y = fft_conv(x, w, n) with one
    waveform per row: returns (gx, gw) for gy = dL/dRe(y) + i dL/dIm(y)."""
    gx = torch.empty_like(x) if want_gx else None
    gw = torch.empty_like(w) if want_gw else None
    L.check(L.lib().mxr_fft_conv_adjoint(L.ref(L.desc(gx)), L.ref(L.desc(gw)), L.ref(L.desc(gy)),
                                         L.ref(L.desc(x)), L.ref(L.desc(w)), int(n),
                                         1 if conj else 0, L.stream_handle(stream)))
    return gx, gw


def autocorr_adjoint(gy, s, n: int, stream=None):
    """Gradient w.r.t. s of r = fft_conv(s, s, n) (rows' matched filters to
    themselves) for gy = dL/dRe(r) + i dL/dIm(r)."""
    gs = torch.empty_like(s)
    L.check(L.lib().mxr_autocorr_adjoint(L.ref(L.desc(gs)), L.ref(L.desc(gy)), L.ref(L.desc(s)),
                                         int(n), L.stream_handle(stream)))
    return gs


def doppler_taps(x, taps, nd: int, want_rd: bool = False, want_pow: bool = True, stream=None):
    """Doppler stage with an arbitrary real window: pow = |FFT_nd(taps .* x)|^2
    along pulses (x: [C, P, R]).  Returns (pow, rd)."""
    C, P, R = x.shape
    pw = torch.empty((C, nd, R), dtype=torch.float32, device=x.device) if want_pow else None
    rd = torch.empty((C, nd, R), dtype=torch.complex64, device=x.device) if want_rd else None
    L.check(L.lib().mxr_doppler_taps(L.ref(L.desc(pw)), L.ref(L.desc(rd)), L.ref(L.desc(x)),
                                     L.ref(L.desc(taps)), int(nd), L.stream_handle(stream)))
    return pw, rd


def doppler_taps_adjoint(gpow, x, taps, nd: int, want_gx: bool = False, want_gtaps: bool = True,
                         stream=None):
    """Gradients of a real loss through pow = doppler_taps(x, taps, nd):
    returns (gx, gtaps) for gpow = dL/dpow."""
    gx = torch.empty_like(x) if want_gx else None
    gt = torch.empty_like(taps) if want_gtaps else None
    L.check(L.lib().mxr_doppler_taps_adjoint(L.ref(L.desc(gt)), L.ref(L.desc(gx)),
                                             L.ref(L.desc(gpow)), L.ref(L.desc(x)),
                                             L.ref(L.desc(taps)), int(nd),
                                             L.stream_handle(stream)))
    return gx, gt


def isl(r, stream=None):
    """Integrated sidelobe level of each row of r ([B, n] c64): [B] f32."""
    out = torch.empty((r.shape[0],), dtype=torch.float32, device=r.device)
    L.check(L.lib().mxr_isl(L.ref(L.desc(out)), L.ref(L.desc(r)), L.stream_handle(stream)))
    return out


def phase_isl(phi, n: int, out=None, stream=None):
    """ISL of the n-point autocorrelation of exp(j phi) per row (phi: [B, L]
    f32), one kernel: equal to isl(fft_conv(s, s, n)) with s = polar(1, phi)."""
    if out is None:
        out = torch.empty((phi.shape[0],), dtype=torch.float32, device=phi.device)
    L.check(L.lib().mxr_phase_isl(L.ref(L.desc(out)), L.ref(L.desc(phi)), int(n),
                                  L.stream_handle(stream)))
    return out


def phase_isl_adjoint(gisl, phi, n: int, stream=None):
    gphi = torch.empty_like(phi)
    L.check(L.lib().mxr_phase_isl_adjoint(L.ref(L.desc(gphi)), L.ref(L.desc(gisl)), L.ref(L.desc(phi)),
                                          int(n), L.stream_handle(stream)))
    return gphi


def isl_adjoint(gisl, r, stream=None):
    gr = torch.empty_like(r)
    L.check(L.lib().mxr_isl_adjoint(L.ref(L.desc(gr)), L.ref(L.desc(gisl)), L.ref(L.desc(r)),
                                    L.stream_handle(stream)))
    return gr


def psl(pw, targets, ml: int = 2, q: int = 8, out=None, stream=None):
    """q-norm peak-sidelobe ratio of the Doppler columns pw[c, :, r] at
    targets ([T, 3] int32: channel, range bin, Doppler bin): [T] f32."""
    if out is None:
        out = torch.empty((targets.shape[0],), dtype=torch.float32, device=pw.device)
    L.check(L.lib().mxr_psl(L.ref(L.desc(out)), L.ref(L.desc(pw)), L.ref(L.desc(targets)),
                            int(ml), int(q), L.stream_handle(stream)))
    return out


def psl_adjoint(gpsl, pw, targets, ml: int = 2, q: int = 8, stream=None):
    gp = torch.empty_like(pw, memory_format=torch.contiguous_format)
    L.check(L.lib().mxr_psl_adjoint(L.ref(L.desc(gp)), L.ref(L.desc(gpsl)), L.ref(L.desc(pw)),
                                    L.ref(L.desc(targets)), int(ml), int(q),
                                    L.stream_handle(stream)))
    return gp


def psl_taps_adjoint(gpsl, pw, x, taps, targets, ml: int = 2, q: int = 8, stream=None):
    """d(sum_t gpsl_t psl_t)/dtaps for pw = doppler_taps(x, taps)'s power,
    from the target columns alone: [P] f32."""
    gt = torch.empty_like(taps)
    L.check(L.lib().mxr_psl_taps_adjoint(L.ref(L.desc(gt)), L.ref(L.desc(gpsl)), L.ref(L.desc(pw)),
                                         L.ref(L.desc(x)), L.ref(L.desc(taps)), L.ref(L.desc(targets)),
                                         int(ml), int(q), L.stream_handle(stream)))
    return gt


def conv1d(x, f, mode="full", method="direct", out=None, stream=None):
    """conv1d along the last axis (include/matx/operators/conv.h:244-266):
    method "direct" (mxr_conv1d) or "fft" (mxr_conv1d_fft, MatX's
    MATX_C_METHOD_FFT, include/matx/transforms/conv.h:51-172)."""
    return _conv1d(x, f, mode, method, False, out, stream)


def corr(x, f, mode="full", method="direct", out=None, stream=None):
    """corr(x, f) = conv1d(x, reverse(conj(f))) (include/matx/transforms/
    corr.h:60-67).  The direct method reverses the (short) filter first; the
    FFT method transforms the reversed filter in its spectrum kernel."""
    if method == "direct":
        fr = torch.flip(f, [-1])
        fr = fr.conj().resolve_conj() if fr.is_complex() else fr.contiguous()
        return _conv1d(x, fr, mode, method, False, out, stream)
    return _conv1d(x, f, mode, method, True, out, stream)


def _conv1d(x, f, mode, method, is_corr, out, stream):
    m = _mode(m=mode)
    n, k = x.shape[-1], f.shape[-1]
    olen = {L.FULL: n + k - 1, L.SAME: n, L.VALID: n - k + 1}[m]
    cplx = x.is_complex() or f.is_complex()
    if out is None:
        out = torch.empty(x.shape[:-1] + (olen,), device=x.device,
                          dtype=torch.complex64 if cplx else torch.float32)
    if method == "direct":
        L.check(L.lib().mxr_conv1d(L.ref(L.desc(out)), L.ref(L.desc(x)), L.ref(L.desc(f)), m,
                                   L.stream_handle(stream)))
    elif method == "fft":
        L.check(L.lib().mxr_conv1d_fft(L.ref(L.desc(out)), L.ref(L.desc(x)), L.ref(L.desc(f)), m,
                                       1 if is_corr else 0, L.stream_handle(stream)))
    else:
        raise ValueError(f"conv1d: method must be 'direct' or 'fft', got {method!r}")
    return out


def conv2d(x, f, mode="full", out=None, stream=None):
    """Direct conv2d over the last two axes (include/matx/operators/conv.h:463-490)."""
    m = _mode(m=mode)
    H, W = x.shape[-2:]
    fh, fw = f.shape
    oh, ow = {L.FULL: (H + fh - 1, W + fw - 1), L.SAME: (H, W),
              L.VALID: (H - fh + 1, W - fw + 1)}[m]
    cplx = x.is_complex() or f.is_complex()
    if out is None:
        out = torch.empty(x.shape[:-2] + (oh, ow), device=x.device,
                          dtype=torch.complex64 if cplx else torch.float32)
    L.check(L.lib().mxr_conv2d(L.ref(L.desc(out)), L.ref(L.desc(x)), L.ref(L.desc(f)), m,
                               L.stream_handle(stream)))
    return out
