"""The colour conversion's oracle (oracle_color_convert) without a GPU.

Pins: the published 100 % colour bars (ITU-R BT.801 for BT.601 limited range,
SMPTE RP 219 for BT.709 limited range, JFIF for BT.601 full range), as the
reference's host tests check them (modules/color-convert/tests/host/
conversion_math_test.cpp:149-255); closed forms (grays carry no chroma, luma is
the matrix's weighted sum, a 2 x 2 block's chroma is its pixels' mean); the
layouts' plane sizes at odd sizes (conversion_math_test.cpp:257-262); and, where
oracle/_ref is built, the module's own CPU converter compiled from its sources,
byte for byte over every pair, matrix, range and edge size."""
import ctypes as C
import itertools

import numpy as np
import pytest

import holoscan_isp as isp
from holoscan_isp import _abi, color
from holoscan_isp.color import Picture

import color_cases as cc
import oracle_lib

needs_cc_ref = pytest.mark.skipif(not oracle_lib.cc_ref_available(), reason="reference converter not built here")

BARS_RGB = [(255, 255, 255), (255, 255, 0), (0, 255, 255), (0, 255, 0), (255, 0, 255), (255, 0, 0), (0, 0, 255),
            (0, 0, 0)]
# Y, Cb, Cr of the 100 % bars as the standards publish them.
PUBLISHED = {
    (_abi.MATRIX_BT601, _abi.RANGE_LIMITED): [(235, 128, 128), (210, 16, 146), (170, 166, 16), (145, 54, 34),
                                              (106, 202, 222), (81, 90, 240), (41, 240, 110), (16, 128, 128)],
    (_abi.MATRIX_BT709, _abi.RANGE_LIMITED): [(235, 128, 128), (219, 16, 138), (188, 154, 16), (173, 42, 26),
                                              (78, 214, 230), (63, 102, 240), (32, 240, 118), (16, 128, 128)],
    (_abi.MATRIX_BT601, _abi.RANGE_FULL): [(255, 128, 128), (226, 0, 149), (179, 171, 0), (150, 44, 21),
                                           (105, 212, 235), (76, 85, 255), (29, 255, 107), (0, 128, 128)],
}
# Samples exactly half way between two codes, which single precision may round
# either way: full-range yellow's Cb and cyan's Cr (JFIF publishes 0).
TIES = {(_abi.MATRIX_BT601, _abi.RANGE_FULL): {(1, 1), (2, 2)}}


def bars(fmt=_abi.PIX_RGB8):
    pic = Picture(fmt, 16, 2)
    data = pic.buf.reshape(2, -1)
    bpp = 4 if fmt in (_abi.PIX_RGBA8, _abi.PIX_BGRA8) else 3
    for x in range(16):
        r, g, b = BARS_RGB[x // 2]
        px = (b, g, r) if fmt in (_abi.PIX_BGR8, _abi.PIX_BGRA8) else (r, g, b)
        for y in range(2):
            data[y, bpp * x:bpp * x + 3] = px
            if bpp == 4:
                data[y, bpp * x + 3] = 77
    return pic


@pytest.mark.parametrize("dst", [_abi.PIX_NV12, _abi.PIX_NV16, _abi.PIX_I420, _abi.PIX_YUYV, _abi.PIX_UYVY])
@pytest.mark.parametrize("key", list(PUBLISHED))
def test_published_colour_bars(dst, key):
    for src in cc.SOURCES:
        out = Picture(dst, 16, 2)
        oracle_lib.color_convert(bars(src), out, *key)
        for bar, want in enumerate(PUBLISHED[key]):
            got = cc.codes(out, 2 * bar, 1)
            for i in range(3):
                slack = 1 if (bar, i) in TIES.get(key, ()) else 0
                assert abs(got[i] - want[i]) <= slack, (src, dst, key, bar, i, got, want)


def test_grays_have_no_chroma_and_luma_is_the_weighted_sum():
    pic = Picture(_abi.PIX_RGB8, 256, 2)
    data = pic.buf.reshape(2, -1)
    for v in range(256):
        data[:, 3 * v:3 * v + 3] = v
    weights = {_abi.MATRIX_BT601: (0.299, 0.114), _abi.MATRIX_BT709: (0.2126, 0.0722),
               _abi.MATRIX_BT2020: (0.2627, 0.0593)}
    for matrix in cc.MATRICES:
        for range_ in cc.RANGES:
            out = Picture(_abi.PIX_NV16, 256, 2)
            oracle_lib.color_convert(pic, out, matrix, range_)
            for v in range(0, 256, 2):
                y, cb, cr = cc.codes(out, v, 0)
                assert cb == 128 and cr == 128
                ys, yo = (219.0, 16.0) if range_ == _abi.RANGE_LIMITED else (255.0, 0.0)
                assert abs(y - (yo + ys * v / 255.0)) <= 0.5 + 1e-6
        mono = Picture(_abi.PIX_MONO8, 4, 1)
        src = Picture(_abi.PIX_RGB8, 4, 1)
        src.buf[:] = [255, 0, 0, 0, 255, 0, 0, 0, 255, 200, 100, 50]
        oracle_lib.color_convert(src, mono, matrix, _abi.RANGE_LIMITED)
        kr, kb = weights[matrix]
        kg = 1 - kr - kb
        want = [kr * 255, kg * 255, kb * 255, kr * 200 + kg * 100 + kb * 50]
        assert [int(v) for v in mono.buf] == [int(np.floor(w + 0.5)) for w in want]


def test_chroma_is_the_mean_of_its_block():
    src = Picture(_abi.PIX_RGB8, 2, 2)
    src.buf[:] = [255, 0, 0, 0, 0, 255, 0, 255, 0, 255, 255, 255]
    out = Picture(_abi.PIX_NV12, 2, 2)
    oracle_lib.color_convert(src, out, _abi.MATRIX_BT709, _abi.RANGE_FULL)
    kr, kb = 0.2126, 0.0722
    kg = 1 - kr - kb
    px = [(255, 0, 0), (0, 0, 255), (0, 255, 0), (255, 255, 255)]
    cbs = [(-kr * r - kg * g + (1 - kb) * b) / (2 * (1 - kb)) + 128 for r, g, b in px]
    crs = [((1 - kr) * r - kg * g - kb * b) / (2 * (1 - kr)) + 128 for r, g, b in px]
    _, cb, cr = cc.codes(out, 0, 0)
    assert cb == int(np.floor(np.mean(cbs) + 0.5)) and cr == int(np.floor(np.mean(crs) + 0.5))


def test_plane_sizes_at_odd_sizes():
    def packed(fmt, w, h):
        return sum(r * n for r, n in (color.plane_size(fmt, w, h, p) for p in range(color.plane_count(fmt))))
    assert packed(_abi.PIX_NV12, 5, 3) == 5 * 3 + 3 * 2 * 2
    assert packed(_abi.PIX_I420, 5, 3) == 5 * 3 + 2 * 3 * 2
    assert packed(_abi.PIX_NV16, 5, 3) == 5 * 3 + 3 * 2 * 3
    assert color.plane_size(_abi.PIX_YUYV, 6, 3, 0) == (12, 3)
    assert color.plane_size(_abi.PIX_RGBA8, 7, 2, 0) == (28, 2)
    with pytest.raises(ValueError):
        color.plane_size(_abi.PIX_RGB8, 4, 4, 1)


@needs_cc_ref
@pytest.mark.parametrize("dst", cc.DESTINATIONS)
def test_oracle_matches_the_reference_converter(dst):
    for src, matrix, range_ in itertools.product(cc.SOURCES, cc.MATRICES, cc.RANGES):
        for i, (w, h) in enumerate(cc.sizes_for(dst)):
            pad = 3 * (i % 2)
            source = cc.rgb_picture(src, w, h, seed=100 * dst + 10 * src + i, pad=pad)
            mine = Picture(dst, w, h, pad=pad, fill=0xA5)
            theirs = Picture(dst, w, h, pad=pad, fill=0xA5)
            oracle_lib.color_convert(source, mine, matrix, range_)
            for mode in (0, 1, 2):  # at one size every resampling mode reads each pixel once
                theirs.buf[:] = 0xA5
                oracle_lib.ref_color_convert(source, theirs, matrix, range_, mode)
                assert np.array_equal(mine.buf, theirs.buf), (src, dst, matrix, range_, w, h, mode)


def test_library_rejects_what_it_does_not_convert():
    lib = isp.lib()
    src = Picture(_abi.PIX_RGB8, 8, 4)
    dst = Picture(_abi.PIX_NV12, 8, 4)
    p = color.params()

    def status(s, d, params=p):
        return lib.isp_color_convert_device(C.byref(s), C.byref(d), C.byref(params), None)

    assert status(dst.view(), src.view()) == _abi.INVALID_ARGUMENT  # YCbCr into RGB
    assert "rgb8" in isp.last_error()
    assert status(src.view(), Picture(_abi.PIX_NV12, 8, 2).view()) == _abi.INVALID_ARGUMENT
    assert "one size" in isp.last_error()
    odd = Picture(_abi.PIX_RGB8, 7, 4)
    yuyv = Picture(_abi.PIX_YUYV, 8, 4).view()
    yuyv.width = 7
    assert status(odd.view(), yuyv) == _abi.INVALID_ARGUMENT
    assert "needs an even width, not 7" in isp.last_error()
    short = dst.view()
    short.planes[1].pitch = 4
    assert status(src.view(), short) == _abi.INVALID_ARGUMENT
    assert "plane 1 pitch 4 is shorter than a row of 8 bytes" in isp.last_error()
    none = dst.view()
    none.planes[1].data = None
    assert status(src.view(), none) == _abi.INVALID_ARGUMENT
    assert "plane 1 has no data" in isp.last_error()
    bad = color.params()
    bad.matrix = 3
    assert status(src.view(), dst.view(), bad) == _abi.INVALID_ARGUMENT
