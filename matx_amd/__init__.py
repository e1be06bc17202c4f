"""matx_amd -- MI355X-native radar signal-processing path with MatX's
operator semantics (fft/ifft, fft_conv, conv1d/conv2d, pulse compression,
three-pulse canceller + Doppler, CA-CFAR) executed by hand-written gfx950 HIP
kernels in libmxr.so (C ABI: include/mxr.h; C++ front end:
include/matx_amd.h)."""
from ._lib import MatxError, lib  # noqa: F401
from . import expr, ops, ragged  # noqa: F401
from .ops import (cfar, conv1d, conv2d, corr, doppler, doppler_cfar, fft, fft_conv, ifft,  # noqa: F401
                  pulse_compress)
from .radar import FusedRadarPipeline, RadarPipeline  # noqa: F401

__all__ = ["MatxError", "lib", "ops", "expr", "fft", "ifft", "fft_conv", "pulse_compress",
           "doppler", "doppler_cfar", "cfar", "conv1d", "conv2d", "corr", "RadarPipeline", "FusedRadarPipeline"]
