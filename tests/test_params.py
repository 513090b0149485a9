"""Settings without a GPU: names, the auto backend, validation of settings and
views (every case of modules/isp/tests/host/params_test.cpp:21-119 and the
rejects of pipeline_test.cpp:706-745, each checked against the reference's
compiled validate_* where built: same verdict, same message), row sizes."""
import ctypes as C

import numpy as np
import pytest

import holoscan_isp as isp
from holoscan_isp import _abi

import oracle_lib

L = isp.lib


def test_names_round_trip():
    for parse, name, values in ((L().isp_parse_backend, L().isp_backend_name, range(3)),
                                (L().isp_parse_demosaic, L().isp_demosaic_name, range(2)),
                                (L().isp_parse_pattern, L().isp_pattern_name, range(4)),
                                (L().isp_parse_white_balance, L().isp_white_balance_name, range(3))):
        for v in values:
            out = C.c_uint8(255)
            assert parse(name(v), C.byref(out)) == 1 and out.value == v
    out = C.c_uint8()
    assert not L().isp_parse_backend(b"vpi", C.byref(out))
    assert not L().isp_parse_demosaic(b"vng", C.byref(out))
    assert not L().isp_parse_pattern(b"rggb", C.byref(out))
    assert L().isp_pattern_name(_abi.GRBG) == b"GRBG"
    assert L().isp_white_balance_name(_abi.WB_GRAY_WORLD) == b"gray-world"


def test_auto_backend():
    assert L().isp_resolve_auto(1) == _abi.BACKEND_GPU
    assert L().isp_resolve_auto(0) == _abi.BACKEND_CPU


def test_defaults_match_python_mirror():
    """isp_params_init() equals IspParams{} (isp_params.hpp:95-121) as the Python mirror states it."""
    c = _abi.Params()
    L().isp_params_init(C.byref(c))
    assert bytes(c) == bytes(isp.IspParams().to_c()[0])


def _validate(params):
    c, keep = params.to_c()
    return L().isp_validate_params(C.byref(c)), isp.last_error()


def _ref_validate(params):
    c, keep = params.to_c()
    return oracle_lib.ref().ref_validate_params(C.byref(c)), oracle_lib.ref().ref_last_error().decode()


BAD_PARAMS = [
    dict(white_level=70000.0), dict(channel_black_levels=(0, 0, -1, 0)),
    dict(lens_shading=isp.LensShading(1, 2, 1, [1.0, 1.0])),
    dict(lens_shading=isp.LensShading(2, 2, 3, [1.0] * 12)),
    dict(lens_shading=isp.LensShading(2, 2, 1, [1.0, 1.0, 1.0, 0.0])),
    dict(lens_shading=isp.LensShading(2, 2, 1, [1.0, 1.0, 1.0])),
    dict(white_patch_percentile=1.0), dict(tone_curve=_abi.TONE_GAMMA, gamma=0.0),
    dict(tone_curve=_abi.TONE_LUT, lut=[0.0, 1.5]), dict(tone_curve=_abi.TONE_LUT),
    dict(auto_exposure=isp.AutoExposureParams(target=1.0)),
    dict(auto_exposure=isp.AutoExposureParams(min_gain=4.0, max_gain=2.0)),
    dict(auto_exposure=isp.AutoExposureParams(max_exposure_ns=1)),
    dict(black_level=-1.0), dict(wb_gains=(1.0, 0.0, 1.0)), dict(ccm=(1, 0, 0, 0, float("nan"), 0, 0, 0, 1)),
    dict(awb_smoothing=0.0), dict(sharpen=5.0), dict(sharpen=-1.0),
]
GOOD_PARAMS = [dict(), dict(lens_shading=isp.LensShading(2, 2, 4, [2.0] * 16)),
               dict(tone_curve=_abi.TONE_LUT, lut=[0.0, 0.25, 1.0]), dict(histogram_bins=1024)]


@pytest.mark.parametrize("kw", BAD_PARAMS)
def test_rejected_params(kw):
    status, message = _validate(isp.IspParams(**kw))
    assert status == _abi.INVALID_ARGUMENT and message
    if oracle_lib.ref_available():
        ref_status, ref_message = _ref_validate(isp.IspParams(**kw))
        assert ref_status == 1 and ref_message == message


@pytest.mark.parametrize("kw", GOOD_PARAMS)
def test_accepted_params(kw):
    assert _validate(isp.IspParams(**kw))[0] == _abi.OK


def test_histogram_bins_extension():
    assert _validate(isp.IspParams(histogram_bins=512))[0] == _abi.INVALID_ARGUMENT


def _raw_status(view_kw, params=None):
    data = np.zeros(64, dtype=np.uint8)
    v = dict(width=4, height=4, pitch=8, format=_abi.RAW_U16, bit_depth=12, pattern=_abi.RGGB)
    v.update(view_kw)
    view = _abi.RawView(v.pop("data", data.ctypes.data), v["width"], v["height"], v["pitch"], v["format"],
                        v["bit_depth"], v["pattern"])
    c, keep = (params or isp.IspParams()).to_c()
    status = L().isp_validate_raw(C.byref(view), C.byref(c))
    message = isp.last_error()
    if oracle_lib.ref_available():
        ref_status = oracle_lib.ref().ref_validate_raw(C.byref(view), C.byref(c))
        assert ref_status == (0 if status == 0 else 1)
        if status:
            assert oracle_lib.ref().ref_last_error().decode() == message
    return status


def test_raw_views():
    """params_test.cpp:309-326 and pipeline_test.cpp:723-744."""
    assert _raw_status({}) == _abi.OK
    for kw in (dict(height=1), dict(pitch=7), dict(format=_abi.RAW_U8), dict(format=_abi.RAW_RAW12, bit_depth=10),
               dict(width=1), dict(pitch=6), dict(bit_depth=17), dict(format=_abi.RAW_RAW10), dict(data=None),
               dict(width=16385)):
        assert _raw_status(kw) == _abi.INVALID_ARGUMENT, kw
    assert _raw_status(dict(format=_abi.RAW_RAW12)) == _abi.OK
    assert _raw_status({}, isp.IspParams(white_level=5000.0)) == _abi.INVALID_ARGUMENT
    assert _raw_status({}, isp.IspParams(black_level=4095.0)) == _abi.INVALID_ARGUMENT


def test_row_bytes():
    """params_test.cpp:95-100."""
    rb = L().isp_raw_row_bytes
    assert rb(_abi.RAW_U8, 7) == 7 and rb(_abi.RAW_U16, 7) == 14
    assert rb(_abi.RAW_RAW10, 4) == 5 and rb(_abi.RAW_RAW10, 5) == 10
    assert rb(_abi.RAW_RAW12, 2) == 3 and rb(_abi.RAW_RAW12, 3) == 6
    for fmt in range(4):
        for w in (0, 1, 2, 3, 4, 5, 3840):
            assert rb(fmt, w) == isp.raw_row_bytes(fmt, w)
            if oracle_lib.ref_available():
                assert rb(fmt, w) == oracle_lib.ref().ref_raw_row_bytes(fmt, w)


def test_tone_tables():
    """params_test.cpp:102-119: endpoints, rising, sRGB codes 18 % gray as about 0.46."""
    for curve in (_abi.TONE_SRGB, _abi.TONE_GAMMA, _abi.TONE_LINEAR):
        lut = isp.tone_lut(isp.IspParams(tone_curve=curve))
        assert lut.size == 4097 and lut[0] == 0.0 and lut[-1] == 1.0
        assert np.all(np.diff(lut) > 0)
    lut = isp.tone_lut(isp.IspParams())
    assert abs(lut[int(0.18 * 4096)] - 0.4613) < 2e-3
    assert list(isp.tone_lut(isp.IspParams(tone_curve=_abi.TONE_LUT, lut=[0.0, 0.25, 1.0]))) == [0.0, 0.25, 1.0]


def test_create_rejects_cpu_backend():
    """This library runs on the GPU only: asking for the cpu backend is refused, loudly."""
    with pytest.raises(ValueError):
        isp.IspPipeline(isp.IspParams(), backend=_abi.BACKEND_CPU)
    with pytest.raises(ValueError):
        isp.IspPipeline(isp.IspParams(sharpen=-1.0))
