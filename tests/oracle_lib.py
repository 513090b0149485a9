"""Loaders of the CPU checkers: the oracle (oracle/libisp_oracle.so, this
repository's C restatement) and, where it was built, the reference's own CPU
backend (oracle/_ref/libisp_ref.so). Test infrastructure only."""
import ctypes as C
import os

import numpy as np

from holoscan_isp import _abi
from holoscan_isp.params import output_channel_bytes, output_channels

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "libisp_oracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libisp_ref.so")
CC_REF_SO = os.path.join(ROOT, "oracle", "_ref", "libcc_ref.so")

P = C.POINTER
_v = C.c_void_p

ORACLE_PROTOTYPES = [
    ("oracle_create", C.c_int, [P(_abi.Params), P(_v)]),
    ("oracle_destroy", None, [_v]),
    ("oracle_reset", None, [_v]),
    ("oracle_process", C.c_int, [_v, P(_abi.RawView), P(_abi.RgbView), P(_abi.StatsOut),
                                 P(_abi.FrameExposure), C.c_int]),
    ("oracle_awb_state", None, [_v, P(C.c_float), P(C.c_int)]),
    ("oracle_awb_measure", _v, [_v, P(_abi.RawView), C.c_int]),
    ("oracle_awb_apply", None, [_v, _v]),
    ("oracle_awb_free", None, [_v]),
    ("oracle_tone_lut", C.c_int, [P(_abi.Params), P(C.c_float), C.c_int]),
    ("oracle_mirror", C.c_int, [C.c_int, C.c_int]),
    ("oracle_load_raw", C.c_int, [P(_abi.RawView), C.c_int, C.c_int]),
    ("oracle_finish_stats", None, [P(C.c_uint64), C.c_int64, P(C.c_float), P(_abi.AutoExposure),
                                   P(_abi.FrameExposure), P(C.c_float)]),
    ("oracle_finish_stats_cases", None, [_v, _v, C.c_int]),
    ("oracle_histogram_percentile", C.c_float, [P(C.c_uint32), C.c_float]),
    ("oracle_frame_check", None, [P(C.c_float), C.c_int, C.c_int, P(C.c_float), C.c_int, _v, C.c_size_t,
                                  P(C.c_uint32), P(C.c_float), P(_abi.FrameCheckResult)]),
    ("oracle_frame_psnr", C.c_double, [P(_abi.FrameCheckResult), C.c_int, C.c_int]),
    ("oracle_color_convert", C.c_int, [P(_abi.ImageView), P(_abi.ImageView), C.c_int, C.c_int]),
]

CC_REF_PROTOTYPES = [
    ("ref_cc_last_error", C.c_char_p, []),
    ("ref_cc_convert", C.c_int, [P(_abi.ImageView), P(_abi.ImageView), C.c_int, C.c_int, C.c_int]),
]

REF_PROTOTYPES = [
    ("ref_last_error", C.c_char_p, []),
    ("ref_validate_params", C.c_int, [P(_abi.Params)]),
    ("ref_validate_raw", C.c_int, [P(_abi.RawView), P(_abi.Params)]),
    ("ref_raw_row_bytes", C.c_size_t, [C.c_int, C.c_int]),
    ("ref_tone_lut", C.c_int, [P(_abi.Params), P(C.c_float), C.c_int]),
    ("ref_create", _v, [P(_abi.Params)]),
    ("ref_destroy", None, [_v]),
    ("ref_reset", None, [_v]),
    ("ref_process", C.c_int, [_v, P(_abi.RawView), P(_abi.RgbView), P(_abi.StatsOut),
                              P(_abi.FrameExposure)]),
    ("ref_awb_state", None, [_v, P(C.c_float), P(C.c_int)]),
    ("ref_synthetic_render", C.c_int, [P(_abi.SyntheticConfig), _v, C.c_size_t]),
    ("ref_synthetic_params", C.c_int, [P(_abi.SyntheticConfig), C.c_int, P(_abi.Params),
                                       P(C.c_float)]),
    ("ref_synthetic_truth", C.c_int, [P(_abi.SyntheticConfig), C.c_int, C.c_int, P(C.c_float)]),
]

_oracle = None
_ref = None
_cc_ref = None


def oracle():
    global _oracle
    if _oracle is None:
        _oracle = _abi.bind(C.CDLL(ORACLE_SO), ORACLE_PROTOTYPES)
    return _oracle


def ref_available():
    return os.path.exists(REF_SO)


def ref():
    global _ref
    if _ref is None:
        _ref = _abi.bind(C.CDLL(REF_SO), REF_PROTOTYPES)
    return _ref


class Frame:
    """A raw frame in host memory: bytes (uint8 array of height*pitch) + its view fields."""

    def __init__(self, data, width, height, pitch, fmt, bit_depth, pattern):
        self.data = np.ascontiguousarray(data, dtype=np.uint8)
        self.width, self.height, self.pitch = width, height, pitch
        self.format, self.bit_depth, self.pattern = fmt, bit_depth, pattern

    def view(self, data_ptr=None):
        return _abi.RawView(data_ptr if data_ptr is not None else self.data.ctypes.data,
                            self.width, self.height, self.pitch, self.format, self.bit_depth,
                            self.pattern)


class Result:
    def __init__(self, image, histogram, stats):
        self.image, self.histogram, self.stats = image, histogram, stats


def _outputs(frame, out_format, with_stats, bins):
    pitch = frame.width * output_channels(out_format) * output_channel_bytes(out_format)
    image = np.zeros(pitch * frame.height, dtype=np.uint8)
    hist = np.zeros(4 * bins, dtype=np.uint32) if with_stats else None
    stats = np.zeros(12, dtype=np.float32) if with_stats else None
    so = _abi.StatsOut(hist.ctypes.data if with_stats else None,
                       stats.ctypes.data if with_stats else None)
    return pitch, image, hist, stats, so


class OraclePipeline:
    """The oracle as a pipeline: process() runs one frame in host memory."""

    def __init__(self, params):
        self.params = params
        self._c, self._keep = params.to_c()
        h = C.c_void_p()
        if oracle().oracle_create(C.byref(self._c), C.byref(h)) != 0:
            raise RuntimeError("oracle_create failed")
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None):
            oracle().oracle_destroy(self._h)
            self._h = None

    def reset(self):
        oracle().oracle_reset(self._h)

    def process(self, frame, out_format=_abi.OUT_RGB8, with_stats=True, exposure=(0, 0.0),
                threads=1):
        bins = self.params.histogram_bins or 256
        pitch, image, hist, stats, so = _outputs(frame, out_format, with_stats, bins)
        out = _abi.RgbView(image.ctypes.data, pitch, out_format)
        e = _abi.FrameExposure(int(exposure[0]), float(exposure[1]))
        raw = frame.view()
        rc = oracle().oracle_process(self._h, C.byref(raw), C.byref(out), C.byref(so), C.byref(e),
                                     threads)
        if rc != 0:
            raise RuntimeError("oracle_process failed")
        return Result(image, hist, stats)

    def awb_state(self):
        g = (C.c_float * 3)()
        v = C.c_int()
        oracle().oracle_awb_state(self._h, g, C.byref(v))
        return list(g), v.value

    def replay_white_balance(self, frames, sequence, threads=1):
        """The white balance state after running frames[i] for each i in `sequence`, in order,
        without their main passes: each distinct frame's statistics are measured once."""
        stats = {}
        try:
            for i in set(sequence):
                s = oracle().oracle_awb_measure(self._h, C.byref(frames[i].view()), threads)
                if not s:
                    raise RuntimeError("oracle_awb_measure failed")
                stats[i] = s
            for i in sequence:
                oracle().oracle_awb_apply(self._h, stats[i])
        finally:
            for s in stats.values():
                oracle().oracle_awb_free(s)


class RefPipeline:
    """The reference CPU backend (compiled from its own sources) as a pipeline."""

    def __init__(self, params):
        self.params = params
        self._c, self._keep = params.to_c()
        h = ref().ref_create(C.byref(self._c))
        if not h:
            raise ValueError(ref().ref_last_error().decode())
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None):
            ref().ref_destroy(self._h)
            self._h = None

    def reset(self):
        ref().ref_reset(self._h)

    def process(self, frame, out_format=_abi.OUT_RGB8, with_stats=True, exposure=(0, 0.0)):
        pitch, image, hist, stats, so = _outputs(frame, out_format, with_stats, 256)
        out = _abi.RgbView(image.ctypes.data, pitch, out_format)
        e = _abi.FrameExposure(int(exposure[0]), float(exposure[1]))
        raw = frame.view()
        rc = ref().ref_process(self._h, C.byref(raw), C.byref(out), C.byref(so), C.byref(e))
        if rc != 0:
            raise ValueError(ref().ref_last_error().decode())
        return Result(image, hist, stats)


def frame_check(truth, lut, rgb, histogram, stats=None):
    """The oracle's frame check (check_frame_host): truth float32 [H, W, 3], lut float32, rgb
    uint8 [H, W, 3] (packed rows), histogram uint32 [4 * 256], stats float32 [12] or None."""
    h, w = truth.shape[:2]
    truth = np.ascontiguousarray(truth, dtype=np.float32)
    lut = np.ascontiguousarray(lut, dtype=np.float32)
    rgb = np.ascontiguousarray(rgb, dtype=np.uint8)
    histogram = np.ascontiguousarray(histogram, dtype=np.uint32)
    fp = C.POINTER(C.c_float)
    st = None if stats is None else np.ascontiguousarray(stats, dtype=np.float32)
    r = _abi.FrameCheckResult()
    oracle().oracle_frame_check(truth.ctypes.data_as(fp), w, h, lut.ctypes.data_as(fp), lut.size,
                                rgb.ctypes.data, 3 * w, histogram.ctypes.data_as(C.POINTER(C.c_uint32)),
                                None if st is None else st.ctypes.data_as(fp), C.byref(r))
    return r


def frame_psnr(result, width, height):
    return oracle().oracle_frame_psnr(C.byref(result), width, height)


def cc_ref_available():
    return os.path.exists(CC_REF_SO)


def cc_ref():
    """The colour conversion module's own CPU converter, compiled from its sources."""
    global _cc_ref
    if _cc_ref is None:
        _cc_ref = _abi.bind(C.CDLL(CC_REF_SO), CC_REF_PROTOTYPES)
    return _cc_ref


def color_convert(source, destination, matrix, range_):
    """The oracle's conversion of host Pictures (holoscan_isp.color.Picture)."""
    rc = oracle().oracle_color_convert(C.byref(source.view()), C.byref(destination.view()), matrix, range_)
    if rc != 0:
        raise ValueError("oracle_color_convert: pair or size not converted")


def ref_color_convert(source, destination, matrix, range_, mode=1):
    rc = cc_ref().ref_cc_convert(C.byref(source.view()), C.byref(destination.view()), matrix, range_, mode)
    if rc != 0:
        raise ValueError(cc_ref().ref_cc_last_error().decode())
