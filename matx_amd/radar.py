"""Radar pipelines over libmxr.so.

``RadarPipeline`` mirrors MatX's ``RadarPipeline`` class
(examples/simple_radar_pipeline.h:157-520): the same buffers (inputView
{C,P,N}, tpcView {C,Pr,R}, waveformView {N}, ba/normT {C,Pr+4,R+12}, xPow,
dets), the same size rules and the same four stage calls, so a caller of the
MatX example finds the same dataflow.  Each stage is one or two HIP launches.

``FusedRadarPipeline`` is the throughput form: three HBM passes per cube
(pulse compression -> canceller+window+Doppler FFT+|.|^2 -> CFAR), with the
Doppler size and canceller alignment as parameters.
"""
from __future__ import annotations

import torch

from . import _lib as L
from . import ops

# MatX CFAR mask: 5 (Doppler) x 13 (range) window, 3 x 3 hole (CUT + guards),
# examples/simple_radar_pipeline.h:237-241.
CFAR_WIN = (5, 13)
CFAR_GUARD = (3, 3)
PFA = 1e-5


def next_pow2(n: int, strict: bool = False) -> int:
    """simple_radar_pipeline.h:180-188: samples round up (non-strict), pulses
    to the next power of two strictly above."""
    r = 1
    while (r <= n) if strict else (r < n):
        r *= 2
    return r


def cfar_mask(device=None) -> torch.Tensor:
    m = torch.ones(CFAR_WIN, dtype=torch.float32, device=device)
    m[1:4, 5:8] = 0
    return m


class RadarPipeline:
    """Stage-for-stage MatX RadarPipeline on one stream."""

    def __init__(self, num_pulses: int, num_samples: int, wf_len: int, num_channels: int,
                 stream=None, device="cuda"):
        self.numPulses = num_pulses
        self.numSamples = num_samples
        self.waveformLength = wf_len
        self.numChannels = num_channels
        self.stream = stream
        self.numSamplesRnd = next_pow2(num_samples)
        self.numPulsesRnd = next_pow2(num_pulses, strict=True)
        self.numCompressedSamples = num_samples - wf_len + 1
        C, P, N = num_channels, num_pulses, self.numSamplesRnd
        Pr, R = self.numPulsesRnd, self.numCompressedSamples
        my, mx = CFAR_WIN
        kw = dict(device=device)
        # zero-initialised like the constructor's cudaMemset (:211-214)
        self.waveformView = torch.zeros(N, dtype=torch.complex64, **kw)
        self.inputView = torch.zeros((C, P, N), dtype=torch.complex64, **kw)
        self.tpcView = torch.zeros((C, Pr, R), dtype=torch.complex64, **kw)
        self.cancelMask = torch.tensor([1.0, -2.0, 1.0], dtype=torch.float32, **kw)
        self.normT = torch.zeros((C, Pr + my - 1, R + mx - 1), dtype=torch.float32, **kw)
        self.ba = torch.zeros((C, Pr + my - 1, R + mx - 1), dtype=torch.float32, **kw)
        self.dets = torch.zeros((C, Pr, R), dtype=torch.int32, **kw)
        self.xPow = torch.zeros((C, Pr, R), dtype=torch.float32, **kw)
        self.cfarMaskView = cfar_mask(device)
        self.cfarNormReady = False

    # -- stages -----------------------------------------------------------
    def PulseCompression(self):
        """:273-301.  Re-windows and re-normalises waveformView in place (as
        MatX does on every call), then matched-filters inputView in place;
        lags 0..R-1 of every pulse are written (the only lags any later stage
        or the reference test reads)."""
        S, L_, R = self.numSamples, self.waveformLength, self.numCompressedSamples
        x = self.inputView[:, :, :S]
        y = self.inputView[:, :, :R]
        w = self.waveformView[:L_]
        ops.pulse_compress(x, w, self.numSamplesRnd, nout=R, window=True, normalize=True,
                           writeback=True, out=y, stream=self.stream)

    def ThreePulseCanceller(self):
        """:321-328: SAME conv1d with [1,-2,1] along pulses of permuted views."""
        C, P, R = self.numChannels, self.numPulses, self.numCompressedSamples
        x = self.inputView.permute(0, 2, 1)[:, :R, :P]
        xo = self.tpcView.permute(0, 2, 1)[:, :R, :P]
        ops.conv1d(x, self.cancelMask, "same", out=xo, stream=self.stream)

    def DopplerProcessing(self):
        """:344-357: Hamming(P-2) on rows 0..P-3 of tpcView, then a Pr-point
        FFT along pulses, in place."""
        ops.doppler(self.tpcView, self.numPulsesRnd, tpc="none",
                    win_rows=self.numPulses - 2, rd=self.tpcView, want_pow=False,
                    stream=self.stream)

    def CFARDetections(self):
        """:394-436: xPow = |tpcView|^2, ba (trimmed view) = masked sum/norm,
        dets = xPow > alpha*ba, fused in one pass."""
        self._compute_cfar_norm()
        my, mx = CFAR_WIN
        Pr, R = self.numPulsesRnd, self.numCompressedSamples
        baTrim = self.ba[:, my // 2: Pr + my // 2, mx // 2: R + mx // 2]
        ops.cfar(self.tpcView, CFAR_WIN, CFAR_GUARD, PFA, dets=self.dets, ba=baTrim,
                 pw=self.xPow, stream=self.stream)

    def _compute_cfar_norm(self):
        """:485-492: normT = conv2d(ones, mask, FULL), once."""
        if not self.cfarNormReady:
            ones = torch.ones((self.numChannels, self.numPulsesRnd, self.numCompressedSamples),
                              dtype=torch.float32, device=self.tpcView.device)
            ops.conv2d(ones, self.cfarMaskView, "full", out=self.normT, stream=self.stream)
            self.cfarNormReady = True

    def run(self):
        self.PulseCompression()
        self.ThreePulseCanceller()
        self.DopplerProcessing()
        self.CFARDetections()

    # -- getters (:443-482) -------------------------------------------------
    def GetInputView(self):
        return self.inputView

    def GetwaveformView(self):
        return self.waveformView

    def GetTPCView(self):
        return self.tpcView

    def GetDetections(self):
        return self.dets

    def GetBackgroundAverages(self):
        return self.ba

    def GetnormT(self):
        self._compute_cfar_norm()
        return self.normT

    def sync(self):
        (self.stream or torch.cuda.current_stream()).synchronize()


class FusedRadarPipeline:
    """Two-kernel throughput form of the radar chain.  x: [C, P, S] resident
    input cube (a view of MatX's zero-padded inputView {C, P, N} works: only
    the first S samples of each pulse are read).

    Per cube: the pulse-compression plan's block kernel (overlap-save,
    spectrum prepared once per waveform) writes the compressed cube; the
    fused pulse-axis pass (canceller, window, Doppler FFT, |.|^2, CA-CFAR)
    reads it once and writes the detection mask.  The power map and the
    background averages are written only when asked (want_pow / want_ba).

    tpc="valid", nd=128 is the 16 x 128 x 9000 benchmark pipeline (126 canceller rows);
    tpc="same", nd=next_pow2(P, strict) is MatX's stage dataflow.
    """

    def __init__(self, C: int, P: int, S: int, L: int, nd: int | None = None,
                 tpc: str = "valid", n: int | None = None, win=CFAR_WIN, guard=CFAR_GUARD,
                 pfa: float = PFA, want_pow: bool = False, want_ba: bool = False,
                 device="cuda", stream=None):
        self.C, self.P, self.S, self.L = C, P, S, L
        self.R = S - L + 1
        self.n = next_pow2(S) if n is None else n
        self.tpc = tpc
        rows = P - 2 if tpc == "valid" else P
        self.nd = next_pow2(rows) if nd is None else nd
        self.win_rows = P - 2 if tpc != "none" else P
        self.win, self.guard, self.pfa = win, guard, pfa
        self.stream = stream
        kw = dict(device=device)
        # rows padded to 64 elements: every row starts on a 256/512-byte
        # boundary, so tile reads and writes cover whole cache lines
        self.ld = (self.R + 63) // 64 * 64
        R = self.R
        self.plan = ops.PulseCompressionPlan(S, L, self.n, R)
        self.y = torch.empty((C, P, self.ld), dtype=torch.complex64, **kw)[:, :, :R]
        self.dets = torch.empty((C, self.nd, self.ld), dtype=torch.int32, **kw)[:, :, :R]
        self.pw = self.ba = None
        self.set_outputs(want_pow, want_ba)

    def set_outputs(self, want_pow: bool, want_ba: bool):
        """Choose whether each cube also writes the power map / averages."""
        kw = dict(dtype=torch.float32, device=self.dets.device)
        shape = (self.C, self.nd, self.ld)
        self.pw = torch.empty(shape, **kw)[:, :, :self.R] if want_pow else None
        self.ba = torch.empty(shape, **kw)[:, :, :self.R] if want_ba else None

    def set_waveform(self, w):
        self.plan.set_waveform(w, stream=self.stream)

    def step(self, x, w=None):
        """One cube.  Passing w re-prepares the waveform first (MatX's
        per-call form); otherwise the spectrum from set_waveform is used."""
        s = self.stream
        if w is not None:
            self.set_waveform(w)
        self.plan(x, out=self.y, stream=s)
        ops.doppler_cfar(self.y, self.nd, tpc=self.tpc, win_rows=self.win_rows, win=self.win,
                         guard=self.guard, pfa=self.pfa, dets=self.dets, ba=self.ba, pw=self.pw,
                         stream=s)
        return self.dets

    def sync(self):
        """Waits for every cube issued so far on the pipeline's stream."""
        (self.stream or torch.cuda.current_stream()).synchronize()

    def algorithmic_bytes(self) -> int:
        """Stage-boundary byte model per cube: read x, write+read the
        compressed cube, write+read the power map, write int32 detections."""
        C, P, S, R, nd = self.C, self.P, self.S, self.R, self.nd
        return 8 * C * P * S + 16 * C * P * R + 8 * C * nd * R + 4 * C * nd * R
