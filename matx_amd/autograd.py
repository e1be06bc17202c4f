"""torch.autograd.Functions over the HIP kernels, for gradient-based
waveform and Doppler-window design.

FFTConv:      y = fft_conv(x, w, n) with one waveform per row; backward runs
              mxr_fft_conv_adjoint (gx = IFFT(GY .* W), gw = IFFT(X .* conj(GY))
              for a correlation).  With x and w the same tensor (a waveform's
              matched filter to itself) torch adds both gradients.
DopplerPower: pow = |FFT_nd(taps .* x)|^2 along pulses; backward runs
              mxr_doppler_taps_adjoint, which recomputes the spectrum from x
              instead of keeping it (the Doppler-stage recomputation that keeps
              the saved state at the size of the input).

RadarPower:   pow = doppler(pulse_compress(x, w), nd) for a whole cube and
              one waveform (the PulseCompression -> ThreePulseCanceller ->
              DopplerProcessing chain); backward runs mxr_doppler_adjoint
              (Doppler spectrum recomputed from the compressed cube, the only
              activation kept) and mxr_pc_plan_waveform_grad (the waveform
              gradient summed over every pulse).
WeightedPowerSum: L = sum_{c,k,r} wmap[k, r] pow[c, k, r]; its backward hands
              on wmap expanded over channels (a stride-0 view, never
              materialised), which the pulse-axis adjoint reads as is.

DopplerPsl:   psl = Psl(DopplerPower(x, taps, nd)); the backward takes the taps
              gradient from the target columns (mxr_psl_taps_adjoint) when x
              needs none.
PhaseIsl:     isl = ISL(fft_conv(s, s, n)) of s = exp(j phi) per row, one kernel
              each way (mxr_phase_isl / mxr_phase_isl_adjoint); the same value
              and gradient as Isl(autocorr(polar(1, phi))).

`design_loss` is the waveform/window design objective: integrated sidelobe
level of every waveform's matched-filter output plus a peak-sidelobe penalty
on the Doppler spectra of a target cube.
"""
from __future__ import annotations

import torch

from . import ops


class FFTConv(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, w, n: int, conj: bool = True):
        x = x.contiguous()
        w = w.contiguous()
        ctx.save_for_backward(x, w)
        ctx.n, ctx.conj = n, conj
        return ops.fft_conv(x, w, n, conj=conj)

    @staticmethod
    def backward(ctx, gy):
        x, w = ctx.saved_tensors
        gx, gw = ops.fft_conv_adjoint(gy.contiguous(), x, w, ctx.n, ctx.conj,
                                      want_gx=ctx.needs_input_grad[0],
                                      want_gw=ctx.needs_input_grad[1])
        return gx, gw, None, None


class AutoCorr(torch.autograd.Function):
    """r = fft_conv(s, s, n): each row's matched filter to itself.  The two
    adjoint terms of FFTConv add up to one pass (mxr_autocorr_adjoint)."""

    @staticmethod
    def forward(ctx, s, n: int):
        s = s.contiguous()
        ctx.save_for_backward(s)
        ctx.n = n
        return ops.fft_conv(s, s, n)

    @staticmethod
    def backward(ctx, gr):
        (s,) = ctx.saved_tensors
        return ops.autocorr_adjoint(gr.contiguous(), s, ctx.n), None


class DopplerPower(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, taps, nd: int):
        x = x.contiguous()
        taps = taps.contiguous()
        ctx.save_for_backward(x, taps)
        ctx.nd = nd
        pw, _ = ops.doppler_taps(x, taps, nd)
        return pw

    @staticmethod
    def backward(ctx, gpow):
        x, taps = ctx.saved_tensors
        gx, gt = ops.doppler_taps_adjoint(gpow.contiguous(), x, taps, ctx.nd,
                                          want_gx=ctx.needs_input_grad[0],
                                          want_gtaps=ctx.needs_input_grad[1])
        return gx, gt, None


class RadarPower(torch.autograd.Function):
    @staticmethod
    def forward(ctx, w, x, plan, nd: int, tpc: str, win_rows):
        w = w.contiguous()
        plan.set_waveform(w)
        # compressed cube and power map on row-padded storage (aligned strips
        # for the pulse-axis kernels, forward and adjoint)
        y = plan(x, out=ops.empty_rows(x.shape[:-1] + (plan.nout,), torch.complex64, x.device))
        C, _, R = y.shape
        pw, _ = ops.doppler(y, nd, tpc=tpc, win_rows=win_rows,
                            pw=ops.empty_rows((C, nd, R), torch.float32, y.device))
        ctx.save_for_backward(w, x, y)
        ctx.plan, ctx.nd, ctx.tpc, ctx.win_rows = plan, nd, tpc, win_rows
        return pw

    @staticmethod
    def backward(ctx, gpow):
        w, x, y = ctx.saved_tensors
        gy = ops.doppler_adjoint(gpow, y, ctx.nd, tpc=ctx.tpc, win_rows=ctx.win_rows,
                                 gy=ops.empty_rows(y.shape, y.dtype, y.device))
        gw = ctx.plan.waveform_grad(gy, x, w)
        return gw, None, None, None, None, None


class WeightedPowerSum(torch.autograd.Function):
    @staticmethod
    def forward(ctx, pw, wmap):
        ctx.save_for_backward(wmap)
        ctx.shape = pw.shape
        # channel sum first ([nd, R] temporary): one pass over the map in
        # place through its strides (row-padded maps included)
        return (pw.sum(0) * wmap).sum()

    @staticmethod
    def backward(ctx, g):
        (wmap,) = ctx.saved_tensors
        gm = ops.empty_rows(wmap.shape, wmap.dtype, wmap.device)
        torch.mul(wmap, g, out=gm)
        return gm.expand(ctx.shape), None


def radar_power(w, x, plan, nd: int, tpc: str = "valid", win_rows=None):
    """Doppler power map of cube x ([C, P, S] c64) filtered with waveform w
    ([L] c64) by `plan` (ops.PulseCompressionPlan(S, L, n), windowed and
    normalised); differentiable with respect to w."""
    return RadarPower.apply(w, x, plan, nd, tpc, win_rows)


def weighted_power_sum(pw, wmap):
    return WeightedPowerSum.apply(pw, wmap)


def contrast_weights(nd: int, R: int, cells, C: int = 1, device=None):
    """[nd, R] f32 map: 1/(C n_clutter) on clutter cells, -1/(C n_target) on
    the target cells (k, r), so weighted_power_sum(pw, map) = mean clutter
    power - mean target power."""
    wm = torch.empty((nd, R), dtype=torch.float32, device=device)
    k = torch.as_tensor([c[0] for c in cells], dtype=torch.int64, device=device)
    r = torch.as_tensor([c[1] for c in cells], dtype=torch.int64, device=device)
    tgt = torch.zeros((nd, R), dtype=torch.bool, device=device)
    tgt[k, r] = True
    nt = int(tgt.sum())
    wm.fill_(1.0 / (C * (nd * R - nt)))
    wm[tgt] = -1.0 / (C * nt)
    return wm


class Isl(torch.autograd.Function):
    """Per-row integrated sidelobe level (mxr_isl / mxr_isl_adjoint)."""

    @staticmethod
    def forward(ctx, r):
        r = r.contiguous()
        ctx.save_for_backward(r)
        return ops.isl(r)

    @staticmethod
    def backward(ctx, g):
        (r,) = ctx.saved_tensors
        return ops.isl_adjoint(g.contiguous(), r)


class PhaseIsl(torch.autograd.Function):
    """Per-row ISL of the n-point autocorrelation of exp(j phi)
    (mxr_phase_isl / mxr_phase_isl_adjoint): Isl(autocorr(polar(1, phi)))
    in one kernel each way."""

    @staticmethod
    def forward(ctx, phi, n: int):
        phi = phi.contiguous()
        ctx.save_for_backward(phi)
        ctx.n = n
        return ops.phase_isl(phi, n)

    @staticmethod
    def backward(ctx, g):
        (phi,) = ctx.saved_tensors
        return ops.phase_isl_adjoint(g.contiguous(), phi, ctx.n), None


class Psl(torch.autograd.Function):
    """Peak-sidelobe ratio of Doppler columns at targets (mxr_psl /
    mxr_psl_adjoint); the gradient is zero outside the target columns."""

    @staticmethod
    def forward(ctx, pw, targets, ml: int, q: int):
        ctx.save_for_backward(pw, targets)
        ctx.ml, ctx.q = ml, q
        return ops.psl(pw, targets, ml, q)

    @staticmethod
    def backward(ctx, g):
        pw, targets = ctx.saved_tensors
        return ops.psl_adjoint(g.contiguous(), pw, targets, ctx.ml, ctx.q), None, None, None


class DopplerPsl(torch.autograd.Function):
    """psl = Psl(DopplerPower(x, taps, nd)) at the targets.  The forward
    computes the whole power map; when x needs no gradient the backward takes
    the taps gradient from the target columns alone (mxr_psl_taps_adjoint)
    instead of a map-sized power gradient and the pulse-axis adjoint."""

    @staticmethod
    def forward(ctx, x, taps, targets, nd: int, ml: int, q: int):
        x = x.contiguous()
        taps = taps.contiguous()
        pw, _ = ops.doppler_taps(x, taps, nd)
        ctx.save_for_backward(x, taps, pw, targets)
        ctx.nd, ctx.ml, ctx.q = nd, ml, q
        return ops.psl(pw, targets, ml, q)

    @staticmethod
    def backward(ctx, g):
        x, taps, pw, targets = ctx.saved_tensors
        g = g.contiguous()
        if ctx.needs_input_grad[0]:
            gp = ops.psl_adjoint(g, pw, targets, ctx.ml, ctx.q)
            gx, gt = ops.doppler_taps_adjoint(gp, x, taps, ctx.nd, want_gx=True,
                                              want_gtaps=ctx.needs_input_grad[1])
            return gx, gt, None, None, None, None
        gt = ops.psl_taps_adjoint(g, pw, x, taps, targets, ctx.ml, ctx.q) if ctx.needs_input_grad[1] else None
        return None, gt, None, None, None, None


_LOSS_WEIGHTS = {}


def _loss_weights(B: int, T: int, lam: float, device) -> torch.Tensor:
    """[B + T] f32: 1/B for the ISL entries, lam/T for the PSL entries (built
    once per shape and weight, so a captured step allocates nothing)."""
    key = (B, T, float(lam), str(device))
    w = _LOSS_WEIGHTS.get(key)
    if w is None:
        w = torch.cat([torch.full((B,), 1.0 / B), torch.full((T,), float(lam) / T)]).to(
            device=device, dtype=torch.float32)
        _LOSS_WEIGHTS[key] = w
    return w


class DesignLoss(torch.autograd.Function):
    """mean ISL + lam * mean PSL (design_loss) with the per-row heads written
    into one [B + T] vector whose weighted sum is the loss: one dot product
    forward and one scale backward around the two heads and their adjoints
    (PhaseIsl's and DopplerPsl's kernels), instead of the means, their
    weighting and their backward as separate elementwise launches."""

    @staticmethod
    def forward(ctx, phi, taps, cube, targets, n: int, nd: int, lam: float, ml: int, q: int):
        phi, taps, cube = phi.contiguous(), taps.contiguous(), cube.contiguous()
        B, T = phi.shape[0], targets.shape[0]
        v = torch.empty((B + T,), dtype=torch.float32, device=phi.device)
        ops.phase_isl(phi, n, out=v[:B])
        pw, _ = ops.doppler_taps(cube, taps, nd)
        ops.psl(pw, targets, ml, q, out=v[B:])
        w = _loss_weights(B, T, lam, phi.device)
        ctx.save_for_backward(phi, taps, cube, pw, targets, w)
        ctx.n, ctx.nd, ctx.ml, ctx.q, ctx.B = n, nd, ml, q, B
        return torch.dot(v, w)

    @staticmethod
    def backward(ctx, g):
        phi, taps, cube, pw, targets, w = ctx.saved_tensors
        B = ctx.B
        gv = w * g
        gphi = ops.phase_isl_adjoint(gv[:B], phi, ctx.n) if ctx.needs_input_grad[0] else None
        gt = gx = None
        if ctx.needs_input_grad[2]:
            gp = ops.psl_adjoint(gv[B:], pw, targets, ctx.ml, ctx.q)
            gx, gt = ops.doppler_taps_adjoint(gp, cube, taps, ctx.nd, want_gx=True,
                                              want_gtaps=ctx.needs_input_grad[1])
        elif ctx.needs_input_grad[1]:
            gt = ops.psl_taps_adjoint(gv[B:], pw, cube, taps, targets, ctx.ml, ctx.q)
        return gphi, gt, gx, None, None, None, None, None, None


def fft_conv(x, w, n: int, conj: bool = True):
    return FFTConv.apply(x, w, n, conj)


def doppler_power(x, taps, nd: int):
    return DopplerPower.apply(x, taps, nd)


def autocorr(s, n: int):
    return AutoCorr.apply(s, n)


def sidelobe_mask(nd: int, kd: torch.Tensor, ml: int) -> torch.Tensor:
    """[T, nd] bool: Doppler bins farther than ml (circularly) from kd."""
    k = torch.arange(nd, device=kd.device)[None, :]
    d = torch.minimum((k - kd[:, None]) % nd, (kd[:, None] - k) % nd)
    return d > ml


def design_loss(phi, taps, cube, targets, n: int, nd: int, lam: float = 1.0, ml: int = 2,
                q: int = 8):
    """phi: [B, Lw] phases; taps: [P] Doppler window; cube: [C, P, R] c64;
    targets: [T, 3] int (channel, range bin, Doppler bin; int32 avoids a
    conversion per call).  Returns mean ISL + lam * mean PSL (the
    definitions of oracle/grad.py); both heads are single HIP kernels, the
    waveform term (phases -> samples -> autocorrelation -> ISL) is one
    kernel forward and one backward, and the two means are one weighted sum
    (DesignLoss)."""
    tg = targets if targets.dtype == torch.int32 else targets.to(torch.int32)
    return DesignLoss.apply(phi, taps, cube, tg, n, nd, lam, ml, q)


def make_design_problem(seed: int, B: int = 64, Lw: int = 1024, C: int = 16, P: int = 128,
                        R: int = 4096, T: int = 24):
    """Synthetic design problem (numpy): B phase-only waveforms of Lw samples
    (phases uniform in [0, 2 pi)), a Hamming-initialised P-tap window, and a
    [C, P, R] cube of unit complex noise with T Doppler tones (amplitude
    uniform in [5, 50]) at seeded (channel, range bin, Doppler bin)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    phi = rng.uniform(0, 2 * np.pi, (B, Lw)).astype(np.float32)
    h = np.hamming(P).astype(np.float32)
    X = ((rng.standard_normal((C, P, R)) + 1j * rng.standard_normal((C, P, R))) /
         np.sqrt(2)).astype(np.complex64)
    targets = []
    for _ in range(T):
        c, r, kd = int(rng.integers(C)), int(rng.integers(R)), int(rng.integers(P))
        X[c, :, r] += np.float32(rng.uniform(5, 50)) * np.exp(2j * np.pi * kd * np.arange(P) / P)
        targets.append((c, r, kd))
    return phi, h, X.astype(np.complex64), targets
