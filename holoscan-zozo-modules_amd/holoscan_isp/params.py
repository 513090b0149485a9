"""Python-side settings of the ISP, mirroring holoscan::isp::IspParams.

Field names, defaults and meanings follow the reference
(modules/isp/include/holoscan/isp/isp_params.hpp:73-121); `to_c()` packs them
into the C struct of include/isp_c.h, keeping the tables alive.
"""
import ctypes as C
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

from . import _abi


@dataclass
class LensShading:
    """LensShading (isp_params.hpp:73-79): gains [height][width][channels]."""
    width: int
    height: int
    channels: int = 1
    gains: Sequence[float] = ()


@dataclass
class AutoExposureParams:
    """AutoExposureParams (isp_params.hpp:82-93)."""
    target: float = 0.18
    damping: float = 0.5
    min_exposure_ns: int = 10_000
    max_exposure_ns: int = 33_000_000
    min_gain: float = 1.0
    max_gain: float = 16.0


@dataclass
class IspParams:
    """IspParams (isp_params.hpp:95-121), plus the histogram_bins extension."""
    black_level: float = 0.0
    channel_black_levels: Optional[Sequence[float]] = None
    white_level: float = 0.0
    lens_shading: Optional[LensShading] = None
    white_balance: int = _abi.WB_MANUAL
    wb_gains: Sequence[float] = (1.0, 1.0, 1.0)
    awb_smoothing: float = 0.25
    white_patch_percentile: float = 0.99
    demosaic: int = _abi.MALVAR
    ccm: Sequence[float] = (1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0)
    tone_curve: int = _abi.TONE_SRGB
    gamma: float = 2.2
    lut: Sequence[float] = ()
    sharpen: float = 0.0
    auto_exposure: AutoExposureParams = field(default_factory=AutoExposureParams)
    histogram_bins: int = 256

    def to_c(self):
        """Returns (ctypes Params, keepalive list)."""
        keep: List[object] = []
        p = _abi.Params()
        p.black_level = self.black_level
        if self.channel_black_levels is not None:
            p.has_channel_black_levels = 1
            for k in range(4):
                p.channel_black_levels[k] = self.channel_black_levels[k]
        p.white_level = self.white_level
        if self.lens_shading is not None:
            s = self.lens_shading
            gains = np.ascontiguousarray(np.asarray(s.gains, dtype=np.float32).ravel())
            keep.append(gains)
            p.has_lens_shading = 1
            p.lens_shading.width = s.width
            p.lens_shading.height = s.height
            p.lens_shading.channels = s.channels
            p.lens_shading.gains = gains.ctypes.data_as(C.POINTER(C.c_float))
            p.lens_shading.gains_count = gains.size
        p.white_balance = self.white_balance
        for k in range(3):
            p.wb_gains[k] = self.wb_gains[k]
        p.awb_smoothing = self.awb_smoothing
        p.white_patch_percentile = self.white_patch_percentile
        p.demosaic = self.demosaic
        for k in range(9):
            p.ccm[k] = self.ccm[k]
        p.tone_curve = self.tone_curve
        p.gamma = self.gamma
        if len(self.lut):
            lut = np.ascontiguousarray(np.asarray(self.lut, dtype=np.float32))
            keep.append(lut)
            p.lut = lut.ctypes.data_as(C.POINTER(C.c_float))
            p.lut_size = lut.size
        p.sharpen = self.sharpen
        ae = self.auto_exposure
        p.auto_exposure.target = ae.target
        p.auto_exposure.damping = ae.damping
        p.auto_exposure.min_exposure_ns = ae.min_exposure_ns
        p.auto_exposure.max_exposure_ns = ae.max_exposure_ns
        p.auto_exposure.min_gain = ae.min_gain
        p.auto_exposure.max_gain = ae.max_gain
        p.histogram_bins = self.histogram_bins
        return p, keep

    @staticmethod
    def from_c(p):
        """The IspParams a C struct holds (tables copied)."""
        out = IspParams()
        out.black_level = p.black_level
        if p.has_channel_black_levels:
            out.channel_black_levels = tuple(p.channel_black_levels)
        out.white_level = p.white_level
        if p.has_lens_shading:
            s = p.lens_shading
            n = s.width * s.height * s.channels
            out.lens_shading = LensShading(s.width, s.height, s.channels,
                                           [s.gains[i] for i in range(n)])
        out.white_balance = p.white_balance
        out.wb_gains = tuple(p.wb_gains)
        out.awb_smoothing = p.awb_smoothing
        out.white_patch_percentile = p.white_patch_percentile
        out.demosaic = p.demosaic
        out.ccm = tuple(p.ccm)
        out.tone_curve = p.tone_curve
        out.gamma = p.gamma
        if p.lut and p.lut_size > 0:
            out.lut = [p.lut[i] for i in range(p.lut_size)]
        out.sharpen = p.sharpen
        a = p.auto_exposure
        out.auto_exposure = AutoExposureParams(a.target, a.damping, a.min_exposure_ns,
                                               a.max_exposure_ns, a.min_gain, a.max_gain)
        out.histogram_bins = p.histogram_bins or 256
        return out


def raw_row_bytes(fmt: int, width: int) -> int:
    """raw_row_bytes (isp_params.cpp:184-197)."""
    w = max(width, 0)
    return {_abi.RAW_U8: w, _abi.RAW_U16: 2 * w, _abi.RAW_RAW10: 5 * ((w + 3) // 4),
            _abi.RAW_RAW12: 3 * ((w + 1) // 2)}[fmt]


def output_channels(fmt: int) -> int:
    return 4 if fmt == _abi.OUT_RGBA8 else 3


def output_channel_bytes(fmt: int) -> int:
    return 2 if fmt == _abi.OUT_RGB16 else 4 if fmt == _abi.OUT_RGB32F else 1
