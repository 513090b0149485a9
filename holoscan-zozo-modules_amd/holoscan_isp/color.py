"""The colour conversion after the ISP (include/isp_color.h): RGB8 / BGR8 / RGBA8 /
BGRA8 into NV12 / NV16 / I420 / YUYV / UYVY / MONO8 at one size, on the GPU, as
the reference's colour conversion module converts them
(modules/color-convert/include/holoscan/color_convert/converter.hpp:60-67)."""
import ctypes as C

import numpy as np

from . import _abi
from . import check, lib

PLANES = {_abi.PIX_NV12: 2, _abi.PIX_NV16: 2, _abi.PIX_I420: 3}


def plane_count(fmt):
    return PLANES.get(fmt, 1)


def plane_size(fmt, width, height, plane):
    """(row bytes, rows) of a packed plane (plane_size, pixel_format.hpp:111-123)."""
    row, rows = C.c_size_t(), C.c_int32()
    check(lib().isp_color_plane_size(fmt, width, height, plane, C.byref(row), C.byref(rows)), "plane_size")
    return row.value, rows.value


def params(matrix=_abi.MATRIX_BT601, range_=_abi.RANGE_LIMITED):
    p = _abi.ColorParams()
    lib().isp_color_params_init(C.byref(p))
    p.matrix, p.range = matrix, range_
    return p


class Picture:
    """An image of `fmt` in one buffer, plane after plane, each row `pad` bytes longer than
    packed and then rounded up to a multiple of `align` bytes (as pitched allocators lay
    rows out): host (numpy uint8) or device (a torch uint8 tensor on cuda)."""

    def __init__(self, fmt, width, height, device=False, pad=0, fill=0, align=1):
        self.format, self.width, self.height = fmt, width, height
        self.layout = []  # (offset, pitch, row bytes, rows) per plane
        offset = 0
        for plane in range(plane_count(fmt)):
            row, rows = plane_size(fmt, width, height, plane)
            pitch = -(-(row + pad) // align) * align
            self.layout.append((offset, pitch, row, rows))
            offset += pitch * rows
        self.nbytes = offset
        if device:
            import torch
            self.buf = torch.full((max(offset, 1),), fill, dtype=torch.uint8, device="cuda")
        else:
            self.buf = np.full(max(offset, 1), fill, dtype=np.uint8)

    def _base(self):
        return self.buf.data_ptr() if hasattr(self.buf, "data_ptr") else self.buf.ctypes.data

    def view(self):
        v = _abi.ImageView()
        v.format, v.width, v.height = self.format, self.width, self.height
        base = self._base()
        for i, (offset, pitch, _, _) in enumerate(self.layout):
            v.planes[i] = _abi.PlaneView(base + offset, pitch)
        return v

    def host(self):
        return self.buf.cpu().numpy() if hasattr(self.buf, "data_ptr") else self.buf

    def planes_of(self, data=None):
        """The packed bytes of each plane (padding dropped), from `data` or this buffer."""
        data = self.host() if data is None else data
        out = []
        for offset, pitch, row, rows in self.layout:
            out.append(np.asarray(data[offset:offset + pitch * rows]).reshape(rows, pitch)[:, :row].copy())
        return out


def convert_device(source, destination, matrix=_abi.MATRIX_BT601, range_=_abi.RANGE_LIMITED, stream=None):
    """isp_color_convert_device on two device Pictures (or ImageViews)."""
    s = source.view() if isinstance(source, Picture) else source
    d = destination.view() if isinstance(destination, Picture) else destination
    check(lib().isp_color_convert_device(C.byref(s), C.byref(d), C.byref(params(matrix, range_)), stream),
          "color_convert")
