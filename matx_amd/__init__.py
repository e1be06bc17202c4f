"""matx_amd -- MI355X-native radar signal-processing path with MatX's
operator semantics (fft/ifft, fft_conv, conv1d/conv2d, pulse compression,
three-pulse canceller + Doppler, CA-CFAR) executed by hand-written gfx950 HIP
kernels in libmxr.so (C ABI: include/mxr.h; C++ front end:
include/matx_amd.h)."""
from ._lib import MatxError, lib  # noqa: F401
from . import expr, ops, ragged, trace  # noqa: F401
from .ops import (cfar, conv1d, conv2d, corr, doppler, doppler_cfar, fft, fft2, fft_conv,  # noqa: F401
                  ifft, ifft2, irfft, irfft2, pulse_compress, rfft, rfft2)
from .radar import FusedRadarPipeline, RadarPipeline  # noqa: F401

__all__ = ["MatxError", "lib", "ops", "expr", "trace", "fft", "ifft", "rfft", "irfft", "fft2", "ifft2",
           "rfft2", "irfft2", "fft_conv", "pulse_compress",
           "doppler", "doppler_cfar", "cfar", "conv1d", "conv2d", "corr", "RadarPipeline", "FusedRadarPipeline"]
