"""Ragged batch of pulse-compression + Doppler jobs (mxr_ragged_*).

Each job is (x [C, P, S] c64, w [L] c64): the plan computes
pw = |FFT_nd(hamming_P .* pulse_compress(x, w))|^2 along pulses,
nd = next_pow2(P), into pw [C, nd, S-L+1] f32 -- the result of
ops.pulse_compress(x, w, next_pow2(S)) then ops.doppler(y, nd, tpc="none",
win_rows=P) -- for every job in a few launches.  Jobs need 513 <= S <=
16384 (the batch's overlap-save blocks hold 1024 to 16384 samples) and
2 <= P <= 1024; the plan raises MatxError (NOT_SUPPORTED) otherwise.
`graph()` captures one execution into a torch CUDA (HIP) graph to replay.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib as L


def next_pow2(v: int) -> int:
    p = 1
    while p < v:
        p *= 2
    return p


class RaggedBatch:
    def __init__(self, jobs, autotune: bool = True, persistent: bool = False, stream=None):
        """jobs: list of (x, w) device tensors, x [C, P, S] c64 with unit sample
        stride, w [L] c64 contiguous.  Allocates pw per job."""
        self.inputs = []
        self.pw = []
        descs = (L.mxr_ragged_job * len(jobs))()
        for i, (x, w) in enumerate(jobs):
            C, P, S = x.shape
            Lw = w.shape[0]
            assert x.stride(2) == 1 and w.is_contiguous()
            pw = torch.empty((C, next_pow2(P), S - Lw + 1), dtype=torch.float32, device=x.device)
            d = descs[i]
            d.x, d.C, d.P, d.S = x.data_ptr(), C, P, S
            d.x_cs, d.x_ps = x.stride(0), x.stride(1)
            d.w, d.L = w.data_ptr(), Lw
            d.pw, d.pw_cs, d.pw_ks = pw.data_ptr(), pw.stride(0), pw.stride(1)
            self.inputs.append((x, w))
            self.pw.append(pw)
        flags = (L.RAGGED_AUTOTUNE if autotune else 0) | (L.RAGGED_PERSISTENT if persistent else 0)
        h = ctypes.c_void_p()
        L.check(L.lib().mxr_ragged_plan_create(ctypes.byref(h), descs, len(jobs), flags,
                                               L.stream_handle(stream)))
        self._h = h

    def block_size(self, job: int) -> int:
        m = ctypes.c_int()
        L.check(L.lib().mxr_ragged_plan_block_size(self._h, job, ctypes.byref(m)))
        return m.value

    def execute(self, stream=None):
        L.check(L.lib().mxr_ragged_execute(self._h, L.stream_handle(stream)))
        return self.pw

    def graph(self, stream=None):
        """Captures one execute() into a CUDA graph; returns the graph,
        replayed with .replay().  The plan builds its tables and streams when
        it is created, so the first execute can be the captured one."""
        s = stream or torch.cuda.current_stream()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            self.execute()
        return g

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            L.check(L.lib().mxr_ragged_plan_destroy(self._h))
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
