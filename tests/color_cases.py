"""Shared cases of the colour conversion tests: the formats this library
converts between, sizes with odd edges, seeded RGB content, and reading codes
back from any destination layout."""
import numpy as np

from holoscan_isp import _abi
from holoscan_isp.color import Picture

SOURCES = [_abi.PIX_RGB8, _abi.PIX_BGR8, _abi.PIX_RGBA8, _abi.PIX_BGRA8]
DESTINATIONS = [_abi.PIX_NV12, _abi.PIX_NV16, _abi.PIX_I420, _abi.PIX_YUYV, _abi.PIX_UYVY, _abi.PIX_MONO8]
MATRICES = [_abi.MATRIX_BT601, _abi.MATRIX_BT709, _abi.MATRIX_BT2020]
RANGES = [_abi.RANGE_LIMITED, _abi.RANGE_FULL]
PACKED_422 = (_abi.PIX_YUYV, _abi.PIX_UYVY)


# Sizes whose rows hold whole 16-pixel spans and a ragged last one, with odd and
# even heights, for pitches aligned to the vector loads.
VECTOR_SIZES = [(64, 5), (72, 6), (112, 3), (200, 9), (17, 2), (48, 1), (31, 4)]


def sizes_for(dst):
    """Sizes with odd widths and heights where the layout allows them (packed 4:2:2 needs
    an even width), a span boundary (16 pixels) and its neighbours."""
    sizes = [(34, 18), (5, 3), (7, 5), (1, 1), (9, 2), (16, 7), (23, 9), (32, 2), (33, 3)]
    return [(w, h) for w, h in sizes if not (dst in PACKED_422 and w % 2)]


def rgb_picture(fmt, width, height, seed, device=False, pad=0, align=1):
    """A source Picture of seeded bytes (alpha included)."""
    pic = Picture(fmt, width, height, pad=pad, align=align)
    rng = np.random.default_rng(seed)
    pic.buf[:] = rng.integers(0, 256, pic.nbytes, dtype=np.uint8)
    if device:
        dev = Picture(fmt, width, height, device=True, pad=pad, align=align)
        import torch
        dev.buf.copy_(torch.from_numpy(pic.buf))
        return pic, dev
    return pic


def codes(pic, x, y, data=None):
    """(Y, Cb, Cr) at pixel (x, y) of a YCbCr Picture (Cb = Cr = None for MONO8)."""
    planes = pic.planes_of(data)
    f = pic.format
    if f == _abi.PIX_MONO8:
        return int(planes[0][y, x]), None, None
    if f in PACKED_422:
        row = planes[0][y]
        pair = row[4 * (x >> 1):4 * (x >> 1) + 4]
        if f == _abi.PIX_YUYV:
            return int(pair[2 * (x & 1)]), int(pair[1]), int(pair[3])
        return int(pair[1 + 2 * (x & 1)]), int(pair[0]), int(pair[2])
    cy = y >> 1 if f in (_abi.PIX_NV12, _abi.PIX_I420) else y
    if f == _abi.PIX_I420:
        return int(planes[0][y, x]), int(planes[1][cy, x >> 1]), int(planes[2][cy, x >> 1])
    return int(planes[0][y, x]), int(planes[1][cy, 2 * (x >> 1)]), int(planes[1][cy, 2 * (x >> 1) + 1])
