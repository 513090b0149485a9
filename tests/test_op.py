"""The IspOp contract without a GPU: IspOpConfig's defaults and validate_config
(modules/isp/src/isp_op.cpp:267-281, the reference's rejects() settings cases,
tests/isp_op_test.cpp:380-393), and the ports setup() declares
(isp_op.cpp:323-344)."""
import pytest

import holoscan_isp as isp
from holoscan_isp import _abi
from holoscan_isp import op


def test_config_defaults():
    c = op.IspOpConfig()
    cc, keep = c.to_c()
    assert (cc.backend, cc.output_format, cc.has_pattern, cc.bit_depth, cc.packing, cc.stats) == \
        (_abi.BACKEND_GPU, _abi.OUT_RGB8, 0, 0, _abi.PACKING_NONE, 0)
    assert (cc.max_width, cc.max_height) == (3840, 2160)
    op.validate_config(c)


@pytest.mark.parametrize("tweak,message", [
    (lambda c: setattr(c, "max_width", 1), "max_width"),
    (lambda c: setattr(c, "max_height", 16385), "max_width"),
    (lambda c: setattr(c, "bit_depth", 17), "bit_depth"),
    (lambda c: (setattr(c, "packing", _abi.PACKING_MIPI_RAW10), setattr(c, "bit_depth", 12)), "10 and 12 bits"),
    (lambda c: (setattr(c, "packing", _abi.PACKING_MIPI_RAW12), setattr(c, "bit_depth", 10)), "10 and 12 bits"),
    (lambda c: (setattr(c.params, "gamma", 0.0), setattr(c.params, "tone_curve", _abi.TONE_GAMMA)), "gamma"),
])
def test_rejected_settings(tweak, message):
    c = op.IspOpConfig()
    tweak(c)
    with pytest.raises(ValueError, match=message):
        op.validate_config(c)
    with pytest.raises(ValueError, match=message):
        op.IspOp(c)


def test_packed_depths_accepted():
    op.validate_config(op.IspOpConfig(packing=_abi.PACKING_MIPI_RAW10, bit_depth=10))
    op.validate_config(op.IspOpConfig(packing=_abi.PACKING_MIPI_RAW12, bit_depth=0))


@pytest.mark.parametrize("fmt,code_bits,pixel", [(_abi.OUT_RGB8, (1, 8), 3), (_abi.OUT_RGBA8, (1, 8), 4),
                                                 (_abi.OUT_RGB16, (1, 16), 6), (_abi.OUT_RGB32F, (2, 32), 12)])
def test_ports(fmt, code_bits, pixel):
    o = op.IspOp(op.IspOpConfig(output_format=fmt, stats=True, max_width=640, max_height=480))
    ports = o.setup()
    assert [p[0] for p in ports] == ["raw", "image", "histogram", "stats"]
    raw, image, hist, stats = ports
    assert raw[1] and raw[2] == 2
    assert image[3] == _abi.MEMORY_DEVICE and image[4] == code_bits and image[5] == 3
    assert image[6] == 640 * 480 * pixel
    assert hist[4] == (1, 32) and hist[5] == 2 and hist[6] == 4 * 4 * 256
    assert stats[4] == (2, 32) and stats[5] == 1 and stats[6] == 48
    # Without `stats` the optional ports do not exist.
    assert [p[0] for p in op.IspOp(op.IspOpConfig(output_format=fmt)).setup()] == ["raw", "image"]


def test_wide_histogram_port():
    o = op.IspOp(op.IspOpConfig(output_format=_abi.OUT_RGB16, stats=True,
                                params=isp.IspParams(histogram_bins=1024)))
    assert o.setup()[2][6] == 4 * 4 * 1024


def test_backend_resolution():
    gpu = bool(isp.lib().isp_probe_platform())
    assert op.IspOp(op.IspOpConfig(backend=_abi.BACKEND_AUTO)).backend() == (_abi.BACKEND_GPU if gpu else _abi.BACKEND_CPU)
    assert op.IspOp(op.IspOpConfig(backend=_abi.BACKEND_CPU)).backend() == _abi.BACKEND_CPU
