"""IspOp through a harness of this repository's own, mirroring the reference's
operator test (modules/isp/tests/isp_op_test.cpp): every output checked bit for
bit against the CPU oracle on the same raw frames (the statistics vector too,
where the reference allows 1e-5); BAYER_* input in device, pinned and pageable
host memory with the CPU waits and copies per frame of each (README.md:146-153);
MONO16 with the configured pattern; packed RAW10; every output format; the
histogram and statistics ports and their emit order; the capture time, frame id,
header, region offsets, exposure, gain and colour space of the outputs; a
staging buffer that grows; a pipeline that cannot start; and unusable inputs
rejected."""
import ctypes as C

import numpy as np
import pytest

import holoscan_isp as isp
from holoscan_isp import _abi
from holoscan_isp import op

from oracle_lib import OraclePipeline

pytestmark = pytest.mark.gpu

W, H = 96, 64
EXPOSURE_NS, GAIN = 8_000_000, 2.0
ROI_X, ROI_Y = 64, 32
HEADER = C.create_string_buffer(b"camera")  # the header passes through as an opaque pointer
BAYER = {_abi.RGGB: _abi.ENCODING_BAYER_RGGB, _abi.BGGR: _abi.ENCODING_BAYER_BGGR,
         _abi.GRBG: _abi.ENCODING_BAYER_GRBG, _abi.GBRG: _abi.ENCODING_BAYER_GBRG}


def capture_time(f):
    return 100_000_000 * (f + 1)


def camera(pattern, fmt, bits, frame, width=W, height=H):
    """camera() (isp_op_test.cpp:88-100)."""
    black = float(1 << (bits - 4))
    return isp.synthetic_config(width=width, height=height, pattern=pattern, format=fmt, bit_depth=bits,
                                black_levels=(black, black, black, black), vignetting=0.3, frame=frame)


class Held:
    """Keeps an input tensor's storage alive."""

    def __init__(self, frame, memory, dtype, shape):
        import torch
        data = frame.data.view(dtype).reshape(shape) if dtype != np.uint8 else frame.data.reshape(shape)
        if memory == _abi.MEMORY_DEVICE:
            self.t = torch.from_numpy(np.ascontiguousarray(data).view(np.int16) if dtype == np.uint16
                                      else np.ascontiguousarray(data)).to("cuda")
            ptr = self.t.data_ptr()
        elif memory == _abi.MEMORY_PINNED:
            src = np.ascontiguousarray(data)
            self.t = torch.empty(src.nbytes, dtype=torch.uint8, pin_memory=True)
            self.t.numpy()[:] = src.view(np.uint8).reshape(-1)
            ptr = self.t.data_ptr()
        else:
            self.t = np.ascontiguousarray(data)
            ptr = self.t.ctypes.data
        self.ptr = ptr


def raw_image(frame, memory, encoding, bits, packed=False, shape3=False):
    """raw_image() (isp_op_test.cpp:70-85) around a tensor in `memory`."""
    u8 = frame.format in (_abi.RAW_U8, _abi.RAW_RAW10, _abi.RAW_RAW12)
    if packed:
        shape = [frame.height, frame.pitch]
    else:
        shape = [frame.height, frame.width] + ([1] if shape3 else [])
    held = Held(frame, memory, np.uint8 if u8 else np.uint16, shape)
    img = _abi.Image()
    img.width, img.height, img.encoding, img.significant_bits = frame.width, frame.height, encoding, bits
    img.color_space = _abi.COLOR_SPACE_SRGB
    img.roi_offset_x, img.roi_offset_y = ROI_X, ROI_Y
    img.exposure_time_ns, img.gain = EXPOSURE_NS, GAIN
    img.header = C.addressof(HEADER)
    img.data = op.tensor(held.ptr, memory, _abi.DTYPE_UINT, 8 if u8 else 16, shape)
    return img, held


def reference(params, frames, out_format):
    """The oracle over the frames, in order (isp_op_test.cpp:103-114)."""
    p = OraclePipeline(params)
    return [p.process(f, out_format, True, (EXPOSURE_NS, GAIN)) for f in frames]


def run(config, frames, images):
    import torch
    o = op.IspOp(config)
    o.start()
    rt = op.TorchRuntime(torch.cuda.current_stream())
    for i, (img, _) in enumerate(images):
        o.compute(img, (capture_time(i), 7 + i), rt, torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    o.stop()
    return rt.emitted


def check_outputs(emitted, expected, config):
    """check_outputs() (isp_op_test.cpp:123-171), the statistics bit for bit."""
    ports = ["image", "histogram", "stats"] if config.stats else ["image"]
    assert [e.port for e in emitted] == ports * len(expected), "emit order"
    per = len(ports)
    channels = isp.output_channels(config.output_format)
    for i, want in enumerate(expected):
        img = emitted[per * i]
        d = img.descriptor
        assert (d["width"], d["height"]) == (W, H)
        assert img.shape == (H, W, channels)
        assert img.data == want.image.tobytes(), f"frame {i} differs from the oracle"
        assert d["header"] == C.addressof(HEADER)
        assert (d["exposure_time_ns"], d["gain"]) == (EXPOSURE_NS, GAIN)
        assert (d["roi_offset_x"], d["roi_offset_y"]) == (ROI_X, ROI_Y)
        assert d["color_space"] == _abi.COLOR_SPACE_UNSPECIFIED
        assert d["encoding"] == {_abi.OUT_RGB8: _abi.ENCODING_RGB8, _abi.OUT_RGBA8: _abi.ENCODING_RGBA8,
                                 _abi.OUT_RGB16: _abi.ENCODING_RGB16,
                                 _abi.OUT_RGB32F: _abi.ENCODING_RGB32F}[config.output_format]
        assert img.metadata == (capture_time(i), 7 + i)
        if config.stats:
            hist, stats = emitted[per * i + 1], emitted[per * i + 2]
            assert hist.shape == (4, 256)
            assert np.frombuffer(hist.data, np.uint32).tolist() == want.histogram.tolist(), "histograms differ"
            assert stats.shape == (12,)
            assert np.array_equal(np.frombuffer(stats.data, np.uint32), want.stats.view(np.uint32)), "stats differ"
            assert hist.metadata == stats.metadata == (capture_time(i), 7 + i)


# Waits and copies per frame (README.md:146-153).
COST = {_abi.MEMORY_DEVICE: (0, 0), _abi.MEMORY_PINNED: (1, 1), _abi.MEMORY_PAGEABLE: (2, 1)}


@pytest.mark.parametrize("memory,fmt,bits,out_format,stats", [
    (_abi.MEMORY_DEVICE, _abi.RAW_U16, 12, _abi.OUT_RGB8, True),
    (_abi.MEMORY_PINNED, _abi.RAW_U8, 8, _abi.OUT_RGBA8, False),
    (_abi.MEMORY_PAGEABLE, _abi.RAW_U16, 12, _abi.OUT_RGB8, True),
], ids=["device u16 RGB8 stats", "pinned u8 RGBA8", "pageable u16 RGB8 stats"])
def test_bayer_input(torch_gpu, memory, fmt, bits, out_format, stats):
    """bayer_input() (isp_op_test.cpp:243-284)."""
    pattern = _abi.BGGR if fmt == _abi.RAW_U8 else _abi.GRBG
    frames = [isp.render_synthetic(camera(pattern, fmt, bits, f)) for f in range(3)]
    images = [raw_image(f, memory, BAYER[pattern], bits) for f in frames]
    params = isp.synthetic_params(camera(pattern, fmt, bits, 0), 9)
    params.white_balance, params.sharpen = _abi.WB_GRAY_WORLD, 0.5
    config = op.IspOpConfig(params=params, output_format=out_format, stats=stats, max_width=128, max_height=128)
    before = op.statistics()
    emitted = run(config, frames, images)
    after = op.statistics()
    check_outputs(emitted, reference(params, frames, out_format), config)
    waits, copies = COST[memory]
    assert after[0] - before[0] == 3 * waits, "CPU waits"
    assert after[1] - before[1] == 3 * copies, "copies"


def test_mono16(torch_gpu):
    """MONO16 [H, W, 1] with the configured pattern and 10 significant bits (isp_op_test.cpp:287-302)."""
    frame = isp.render_synthetic(camera(_abi.GBRG, _abi.RAW_U16, 10, 0))
    params = isp.synthetic_params(camera(_abi.GBRG, _abi.RAW_U16, 10, 0), 9)
    config = op.IspOpConfig(params=params, pattern=_abi.GBRG, output_format=_abi.OUT_RGB16, max_width=W,
                            max_height=H)
    images = [raw_image(frame, _abi.MEMORY_DEVICE, _abi.ENCODING_MONO16, 10, shape3=True)]
    check_outputs(run(config, [frame], images), reference(params, [frame], _abi.OUT_RGB16), config)


def test_packed_raw10(torch_gpu):
    """MIPI RAW10 in a uint8 [H, row bytes] tensor, to float RGB (isp_op_test.cpp:305-327)."""
    frames = [isp.render_synthetic(camera(_abi.RGGB, _abi.RAW_RAW10, 10, f)) for f in range(2)]
    params = isp.synthetic_params(camera(_abi.RGGB, _abi.RAW_RAW10, 10, 0), 9)
    params.demosaic = _abi.BILINEAR
    config = op.IspOpConfig(params=params, packing=_abi.PACKING_MIPI_RAW10, output_format=_abi.OUT_RGB32F,
                            stats=True, max_width=W, max_height=H)
    images = [raw_image(f, _abi.MEMORY_DEVICE, _abi.ENCODING_BAYER_RGGB, 8, packed=True) for f in frames]
    check_outputs(run(config, frames, images), reference(params, frames, _abi.OUT_RGB32F), config)


def test_every_output_format(torch_gpu):
    frames = [isp.render_synthetic(camera(_abi.RGGB, _abi.RAW_RAW12, 12, f)) for f in range(2)]
    params = isp.synthetic_params(camera(_abi.RGGB, _abi.RAW_RAW12, 12, 0), 9)
    for out_format in (_abi.OUT_RGB8, _abi.OUT_RGBA8, _abi.OUT_RGB16, _abi.OUT_RGB32F):
        config = op.IspOpConfig(params=params, packing=_abi.PACKING_MIPI_RAW12, output_format=out_format,
                                stats=True, max_width=W, max_height=H)
        images = [raw_image(f, _abi.MEMORY_DEVICE, _abi.ENCODING_BAYER_RGGB, 12, packed=True) for f in frames]
        check_outputs(run(config, frames, images), reference(params, frames, out_format), config)


def test_staging_grows_once(torch_gpu):
    """The staging buffer for host input grows with the largest frame so far; growing it waits
    for the stream once (README.md:155-157)."""
    import torch
    params = isp.synthetic_params(camera(_abi.RGGB, _abi.RAW_U16, 12, 0), 9)
    o = op.IspOp(op.IspOpConfig(params=params, max_width=256, max_height=256))
    o.start()
    rt = op.TorchRuntime(torch.cuda.current_stream())
    sizes = [(64, 48), (128, 96), (96, 64)]
    before = op.statistics()
    for i, (w, h) in enumerate(sizes):
        f = isp.render_synthetic(camera(_abi.RGGB, _abi.RAW_U16, 12, i, w, h))
        img, held = raw_image(f, _abi.MEMORY_PINNED, _abi.ENCODING_BAYER_RGGB, 12)
        o.compute(img, (capture_time(i), i), rt, torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    after = op.statistics()
    # 3 frames x (1 copy, 1 wait) + 1 wait for the one growth after the first allocation.
    assert after[1] - before[1] == 3
    assert after[0] - before[0] == 3 + 1
    o.stop()


def test_pipeline_that_cannot_start(torch_gpu):
    """start() keeps the error; compute() fails the frame with it (isp_op.cpp:352-376)."""
    import torch
    o = op.IspOp(op.IspOpConfig(backend=_abi.BACKEND_CPU))
    o.start()
    frame = isp.render_synthetic(camera(_abi.RGGB, _abi.RAW_U16, 12, 0))
    img, held = raw_image(frame, _abi.MEMORY_DEVICE, _abi.ENCODING_BAYER_RGGB, 12)
    with pytest.raises(RuntimeError, match="IspOp: the pipeline could not start"):
        o.compute(img, (0, 0), op.TorchRuntime(torch.cuda.current_stream()))


def test_auto_backend(torch_gpu):
    assert op.IspOp(op.IspOpConfig(backend=_abi.BACKEND_AUTO)).backend() == _abi.BACKEND_GPU


def rejected(config, image, match):
    import torch
    o = op.IspOp(config)
    o.start()
    rt = op.TorchRuntime(torch.cuda.current_stream())
    with pytest.raises(ValueError, match=match):
        o.compute(image, (0, 0), rt, torch.cuda.current_stream().cuda_stream)
    assert rt.emitted == []


def test_rejects(torch_gpu):
    """rejects() (isp_op_test.cpp:343-393), inputs part."""
    frame = isp.render_synthetic(camera(_abi.RGGB, _abi.RAW_U16, 12, 0))
    config = op.IspOpConfig(max_width=W, max_height=H)
    img, held = raw_image(frame, _abi.MEMORY_PAGEABLE, _abi.ENCODING_MONO16, 12)
    rejected(config, img, "needs the Bayer pattern")
    img, held = raw_image(frame, _abi.MEMORY_PAGEABLE, _abi.ENCODING_RGB8, 12)
    rejected(config, img, "unsupported raw encoding RGB8")
    img, held = raw_image(frame, _abi.MEMORY_PAGEABLE, _abi.ENCODING_BAYER_RGGB, 12)
    rejected(op.IspOpConfig(max_width=W, max_height=H, pattern=_abi.BGGR), img, "disagrees with the configured")
    rejected(op.IspOpConfig(max_width=W // 2, max_height=H), img, "is outside 2x2 to the configured 48x64")
    dark = op.IspOpConfig(max_width=W, max_height=H, params=isp.IspParams(black_level=5000.0))
    rejected(dark, img, "black levels must be below the white level")
    img.data.strides[1] = 2  # samples of a row not packed
    rejected(config, img, "within a row must be packed")
    img, held = raw_image(frame, _abi.MEMORY_PAGEABLE, _abi.ENCODING_BAYER_RGGB, 12)
    img.width = W + 2  # the descriptor disagrees with its tensor
    rejected(config, img, "descriptor's size")
    packed = isp.render_synthetic(camera(_abi.RGGB, _abi.RAW_RAW12, 12, 0))
    img, held = raw_image(packed, _abi.MEMORY_DEVICE, _abi.ENCODING_BAYER_RGGB, 12, packed=True)
    img.width = W * 2  # rows too short for the width
    rejected(op.IspOpConfig(packing=_abi.PACKING_MIPI_RAW12, max_width=4 * W, max_height=H), img,
             "cannot hold 192 samples")
