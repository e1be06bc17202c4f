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
