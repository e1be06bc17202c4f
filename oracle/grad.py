"""FP64 oracle for the differentiable waveform/window design loss.

Test infrastructure only (imported by tests/ and bench.py's CPU-baseline leg,
never by the product path).  numpy float64/complex128 throughout.

The loss (the differentiable form of the pulse-compression + Doppler path,
built from fft_conv, the Doppler window, the Doppler FFT and |.|^2):

  s_b   = exp(i phi_b)                                    b < B, length Lw
  r_b   = fft_conv(s_b, s_b, n) = IFFT_n(|FFT_n(s_b)|^2)  (autocorrelation,
          the matched filter of a waveform to itself, examples/fft_conv.cu:90-108)
  ISL_b = sum_{l != 0} |r_b[l]|^2 / |r_b[0]|^2
  Pw    = |FFT_nd(h .* X)|^2 along pulses                 X: [C, P, R] c64
  PSL_t = (sum_{k not in ML_t} Pw[c_t, k, r_t]^q)^(1/q) / Pw[c_t, kd_t, r_t]
          ML_t = Doppler bins within `ml` of kd_t (circular), q = 8
  L     = mean_b ISL_b + lam * mean_t PSL_t

`loss_and_grad` returns L and the analytic gradients dL/dphi, dL/dh written
with the same adjoint formulas the GPU kernels implement
(gx = IFFT(GY .* W), gw = IFFT(X .* conj(GY)) for a correlation; gu = F^H gD,
gh = sum Re(conj(gu) x) for the window), in the convention
g = dL/dRe + i dL/dIm for complex intermediates.  `loss` alone is what the
central finite differences differentiate.
"""
from __future__ import annotations

import numpy as np


def waveforms(phi):
    return np.exp(1j * np.asarray(phi, dtype=np.float64))


def autocorr(s, n):
    S = np.fft.fft(s, n, axis=-1)
    return np.fft.ifft(S * np.conj(S), axis=-1)


def isl_terms(phi, n):
    r = autocorr(waveforms(phi), n)
    pr = np.abs(r) ** 2
    return (pr.sum(-1) - pr[:, 0]) / pr[:, 0], r, pr


def doppler_power(X, h, nd):
    u = np.asarray(h, dtype=np.float64)[None, :, None] * np.asarray(X, dtype=np.complex128)
    D = np.fft.fft(u, nd, axis=1)
    return np.abs(D) ** 2, D


def sidelobe_mask(nd, kd, ml):
    k = np.arange(nd)
    d = np.minimum((k - kd) % nd, (kd - k) % nd)
    return d > ml


def psl_terms(Pw, targets, ml, q):
    out = []
    for (c, r, kd) in targets:
        col = Pw[c, :, r]
        m = sidelobe_mask(Pw.shape[1], kd, ml)
        A = np.sum(col[m] ** q) ** (1.0 / q)
        out.append(A / col[kd])
    return np.array(out)


def loss(phi, h, X, targets, n, nd, lam=1.0, ml=2, q=8):
    isl, _, _ = isl_terms(phi, n)
    Pw, _ = doppler_power(X, h, nd)
    return float(isl.mean() + lam * psl_terms(Pw, targets, ml, q).mean())


def loss_and_grad(phi, h, X, targets, n, nd, lam=1.0, ml=2, q=8):
    phi = np.asarray(phi, dtype=np.float64)
    B, Lw = phi.shape
    s = waveforms(phi)
    isl, r, pr = isl_terms(phi, n)
    # dL/dpr: l != 0 -> 1/(B pr0); l = 0 -> -(sum_{l!=0} pr)/(B pr0^2)
    pr0 = pr[:, :1]
    g_pr = np.broadcast_to(1.0 / (B * pr0), pr.shape).copy()
    g_pr[:, 0] = -(pr.sum(-1) - pr[:, 0]) / (B * pr[:, 0] ** 2)
    g_r = 2.0 * g_pr * r                                   # pr = |r|^2
    # autocorrelation r = corr(s, s): gs = gx + gw
    S = np.fft.fft(s, n, axis=-1)
    GR = np.fft.fft(g_r, axis=-1)
    gx = np.fft.ifft(GR * S, axis=-1)[:, :Lw]
    gw = np.fft.ifft(S * np.conj(GR), axis=-1)[:, :Lw]
    g_s = gx + gw
    g_phi = np.real(np.conj(g_s) * 1j * s)                 # ds = i s dphi

    Pw, D = doppler_power(X, h, nd)
    T = len(targets)
    g_Pw = np.zeros_like(Pw)
    for (c, rr, kd) in targets:
        col = Pw[c, :, rr]
        m = sidelobe_mask(nd, kd, ml)
        A = np.sum(col[m] ** q) ** (1.0 / q)
        M = col[kd]
        g = np.zeros(nd)
        g[m] = (A ** (1 - q)) * col[m] ** (q - 1) / M
        g[kd] += -A / M ** 2
        g_Pw[c, :, rr] += lam * g / T
    g_D = 2.0 * g_Pw * D
    g_u = nd * np.fft.ifft(g_D, axis=1)[:, :X.shape[1], :]   # F^H, unnormalised
    g_h = np.sum(np.real(np.conj(g_u) * X), axis=(0, 2))
    L = float(isl.mean() + lam * psl_terms(Pw, targets, ml, q).mean())
    return L, g_phi, g_h


def central_fd(f, x, idx, h=1e-3):
    """Central differences of scalar f at flat indices idx of array x."""
    x = np.array(x, dtype=np.float64)
    out = []
    for i in idx:
        xp = x.copy()
        xm = x.copy()
        xp.flat[i] += h
        xm.flat[i] -= h
        out.append((f(xp) - f(xm)) / (2 * h))
    return np.array(out)


# ------------------------------------------------- radar chain (whole cube)
# Waveform design over a cube: the loss is a fixed weighting of the Doppler
# power map of the full chain (examples/simple_radar_pipeline.h:273-357)
#
#   w2   = h_L w / ||h_L w||                     (PulseCompression :288-294)
#   y    = corr(x, w2), lags m < R = S - L + 1   (:295-300)
#   t    = canceller(y) (SAME: P rows, VALID: P-2 rows, or none)
#   D    = FFT_nd(hw .* t) along pulses, hw_i = hamming(win_rows)_i for
#          i < win_rows, 1 after (:346-356)
#   Pw   = |D|^2,  L = sum_{c,k,r} wmap[k, r] Pw[c, k, r]
#
# The forward stages are the C oracle's (oracle/oracle.py); the gradient
# below is written from the adjoint of each stage, in the convention
# g = dL/dRe + i dL/dIm, and pinned by central differences in
# tests/test_grad_oracle.py.  `mistake` plants one wrong adjoint term (tests
# only): "no_norm" drops the normalisation term, "tpc_shift" misplaces the
# VALID canceller's transpose, "conj_corr" conjugates the wrong operand of
# the correlation adjoint, "window_all" windows every canceller row.

def _tpc_rows(P, tpc):
    return P - 2 if tpc == "valid" else P


def _doppler_taps(P, tpc, win_rows, rows_window=None):
    from . import oracle as orc
    rows = _tpc_rows(P, tpc)
    wr = rows_window if rows_window is not None else win_rows
    hw = np.ones(rows)
    hw[:wr] = orc.hamming(wr)
    return hw


def radar_power_w(w, X, n, nd, tpc="valid", win_rows=None):
    """Pw [C, nd, R] of the chain for a complex waveform w [L]."""
    from . import oracle as orc
    X = np.asarray(X, dtype=np.complex128)
    C, P, S = X.shape
    L = len(w)
    R = S - L + 1
    rows = _tpc_rows(P, tpc)
    win_rows = (P - 2 if tpc != "none" else rows) if win_rows is None else win_rows
    Wc = orc.waveform_spectrum(w, n, True, True)
    y = orc.pulse_compress(X.reshape(C * P, S), Wc, R).reshape(C, P, R)
    t = y if tpc == "none" else orc.tpc(y, tpc)
    return orc.abs2(orc.doppler(t, win_rows, nd)), y


def radar_loss_w(w, X, wmap, n, nd, tpc="valid", win_rows=None):
    Pw, _ = radar_power_w(w, X, n, nd, tpc, win_rows)
    return float(np.sum(np.asarray(wmap, dtype=np.float64)[None] * Pw))


def radar_loss(phi, X, wmap, n, nd, tpc="valid", win_rows=None):
    return radar_loss_w(waveforms(phi), X, wmap, n, nd, tpc, win_rows)


def radar_grad_w(w, X, wmap, n, nd, tpc="valid", win_rows=None, mistake=None):
    """(L, dL/dw) for a complex waveform w [L]."""
    from . import oracle as orc
    w = np.asarray(w, dtype=np.complex128)
    X = np.asarray(X, dtype=np.complex128)
    C, P, S = X.shape
    L = len(w)
    R = S - L + 1
    rows = _tpc_rows(P, tpc)
    win_rows = (P - 2 if tpc != "none" else rows) if win_rows is None else win_rows
    Pw, y = radar_power_w(w, X, n, nd, tpc, win_rows)
    wmap = np.asarray(wmap, dtype=np.float64)
    loss_v = float(np.sum(wmap[None] * Pw))
    # recompute the Doppler spectrum (numpy, not the stored stage)
    hw = _doppler_taps(P, tpc, win_rows, rows if mistake == "window_all" else None)
    if tpc == "none":
        t = y
    elif tpc == "same":
        yp = np.pad(y, ((0, 0), (1, 1), (0, 0)))
        t = yp[:, :-2] - 2 * yp[:, 1:-1] + yp[:, 2:]
    else:
        t = y[:, :-2] - 2 * y[:, 1:-1] + y[:, 2:]
    D = np.fft.fft(hw[None, :, None] * t, nd, axis=1)
    gD = 2.0 * wmap[None] * D
    gu = nd * np.fft.ifft(gD, axis=1)[:, :rows]                # F^H, unnormalised
    s = hw[None, :, None] * gu
    if tpc == "none":
        gy = s
    else:
        o = 1 if tpc == "valid" and mistake != "tpc_shift" else 0
        # gy[p] = s[p-o+1] - 2 s[p-o] + s[p-o-1], s = 0 outside [0, rows)
        def s_at(i):
            out = np.zeros((C, P, R), dtype=np.complex128)
            p = np.arange(P)
            ok = (p - o + i >= 0) & (p - o + i < rows)
            out[:, ok] = s[:, (p - o + i)[ok]]
            return out
        gy = s_at(1) - 2 * s_at(0) + s_at(-1)
    # correlation adjoint: g2[l] = sum_rows sum_m x[m + l] conj(gy[m])
    xs = X.reshape(C * P, S)
    gys = gy.reshape(C * P, R)
    if mistake == "conj_corr":
        G = np.conj(np.fft.fft(xs, n, axis=-1)) * np.fft.fft(gys, n, axis=-1)
    else:
        G = np.fft.fft(xs, n, axis=-1) * np.conj(np.fft.fft(gys, n, axis=-1))
    g2 = np.fft.ifft(G.sum(0))[:L]
    hL = orc.hamming(L)
    w1 = hL * w
    nu = np.linalg.norm(w1)
    g1 = g2 / nu
    if mistake != "no_norm":
        g1 = g1 - w1 * np.real(np.vdot(g2, w1)) / nu ** 3
    return loss_v, hL * g1


def radar_loss_and_grad(phi, X, wmap, n, nd, tpc="valid", win_rows=None, mistake=None):
    """(L, dL/dphi) for phase-only waveforms w = exp(i phi)."""
    w = waveforms(phi)
    loss_v, gw = radar_grad_w(w, X, wmap, n, nd, tpc, win_rows, mistake)
    return loss_v, np.real(np.conj(gw) * 1j * w)


def contrast_weights(nd, R, cells, C=1):
    """wmap [nd, R]: 1/(C n_clutter) on clutter cells and -1/(C n_target) on
    the target cells (k, r), so L = mean clutter power - mean target power
    (means over channels and cells)."""
    wmap = np.full((nd, R), 1.0)
    tgt = np.zeros((nd, R), dtype=bool)
    for k, r in cells:
        tgt[k, r] = True
    nt = int(tgt.sum())
    nc = nd * R - nt
    wmap[:] = 1.0 / (C * nc)
    wmap[tgt] = -1.0 / (C * nt)
    return wmap
