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
    """Three-pass throughput form.  x: [C, P, >=S] resident input cube (only
    the first S samples of each pulse are read, e.g. MatX's inputView
    {C, P, N}).

    tpc="valid", nd=P (rounded up) is the BASELINE workload (126 canceller
    rows, 128-point Doppler); tpc="same", nd=next_pow2(P, strict) is MatX's.
    """

    def __init__(self, C: int, P: int, S: int, L: int, nd: int | None = None,
                 tpc: str = "valid", n: int | None = None, win=CFAR_WIN, guard=CFAR_GUARD,
                 pfa: float = PFA, device="cuda", stream=None):
        self.C, self.P, self.S, self.L = C, P, S, L
        self.R = S - L + 1
        self.n = next_pow2(S) if n is None else n
        self.tpc = tpc
        rows = P - 2 if tpc == "valid" else P
        self.nd = next_pow2(rows) if nd is None else nd
        self.win, self.guard, self.pfa = win, guard, pfa
        self.stream = stream
        kw = dict(device=device)
        # intermediate rows padded to 64 elements: every row starts on a
        # 256/512-byte boundary, so tile reads and writes cover whole lines
        self.ld = (self.R + 63) // 64 * 64
        R = self.R
        self.y = torch.empty((C, P, self.ld), dtype=torch.complex64, **kw)[:, :, :R]
        self.pw = torch.empty((C, self.nd, self.ld), dtype=torch.float32, **kw)[:, :, :R]
        self.dets = torch.empty((C, self.nd, self.ld), dtype=torch.int32, **kw)[:, :, :R]

    def step(self, x, w):
        s = self.stream
        ops.pulse_compress(x, w, self.n, nout=self.R, out=self.y, stream=s)
        ops.doppler(self.y, self.nd, tpc=self.tpc, win_rows=self.P - 2, pw=self.pw, stream=s)
        ops.cfar(self.pw, self.win, self.guard, self.pfa, dets=self.dets, stream=s)
        return self.dets

    def algorithmic_bytes(self) -> int:
        """Stage-boundary byte model per cube: read x, write+read the
        compressed cube, write+read the power map, write int32 detections."""
        C, P, S, R, nd = self.C, self.P, self.S, self.R, self.nd
        return 8 * C * P * S + 16 * C * P * R + 8 * C * nd * R + 4 * C * nd * R
