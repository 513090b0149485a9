"""isp_check_frame_device_static, the frame check that allocates nothing (its
scratch is a static device array instead of a workspace argument), against
isp_check_frame_device with a zeroed workspace: the same 64-byte result, bit
for bit, on two calls in a row, the second with a smaller frame, with pixel
errors and histogram bins planted so that every field of the result is
exercised. Sizes beyond the array's maximum are refused before any launch."""
import ctypes as C

import numpy as np
import pytest

import holoscan_isp as isp
from holoscan_isp import _abi

import cases

pytestmark = pytest.mark.gpu


def checked_frame(torch, width, height, frame, stream):
    """A demo camera frame rendered and processed on the device, with planted errors: a few
    pixels changed and two histogram bins off by one. Returns the checker's inputs."""
    cam = cases.demo_camera(width=width, height=height, pattern=_abi.GRBG, bits=12, packing="raw12", frame=frame)
    params = cases.demo_params(cam, _abi.MALVAR, _abi.WB_MANUAL, 0.5)
    pipe = isp.IspPipeline(params)
    pitch = (isp.lib().isp_raw_row_bytes(cam.format, cam.width) + 1) & ~1
    raw = torch.empty(pitch * height, dtype=torch.uint8, device="cuda")
    out = torch.empty(3 * width * height, dtype=torch.uint8, device="cuda")
    hist = torch.empty(1024, dtype=torch.int32, device="cuda")
    stats = torch.empty(12, dtype=torch.float32, device="cuda")
    isp.render_synthetic_device(cam, raw, pitch, stream)
    view = _abi.RawView(raw.data_ptr(), width, height, pitch, cam.format, cam.bit_depth, cam.pattern)
    pipe.process(view, out, 3 * width, _abi.OUT_RGB8, hist, stats, cases.DEMO_EXPOSURE, stream)
    torch.cuda.synchronize()
    out[7 * 3 * width + 30:7 * 3 * width + 36].add_(9)
    out[-5:].sub_(3)
    hist[17] += 1
    hist[600] -= 1
    lut = torch.from_numpy(np.asarray(isp.tone_lut(params), dtype=np.float32)).cuda()
    pipe.close()
    return cam, lut, out, hist, stats


def with_workspace(torch, cam, lut, out, hist, stats, stream):
    work = torch.zeros(_abi.FRAME_CHECK_SCRATCH_BYTES, dtype=torch.uint8, device="cuda")
    result = torch.full((C.sizeof(_abi.FrameCheckResult),), 0xEE, dtype=torch.uint8, device="cuda")
    isp.check(isp.lib().isp_check_frame_device(C.byref(cam), lut.data_ptr(), lut.numel(), out.data_ptr(), 0,
                                               hist.data_ptr(), stats.data_ptr(), work.data_ptr(),
                                               result.data_ptr(), stream), "check_frame_device")
    torch.cuda.synchronize()
    assert not work.any(), "the check leaves its workspace zeroed"
    return bytes(result.cpu().numpy())


def without_allocation(torch, cam, lut, out, hist, stats, stream):
    result = torch.full((C.sizeof(_abi.FrameCheckResult),), 0xEE, dtype=torch.uint8, device="cuda")
    isp.check(isp.lib().isp_check_frame_device_static(C.byref(cam), lut.data_ptr(), lut.numel(), out.data_ptr(),
                                                      0, hist.data_ptr(), stats.data_ptr(), result.data_ptr(),
                                                      stream), "check_frame_device_static")
    torch.cuda.synchronize()
    return bytes(result.cpu().numpy())


def test_static_check_equals_the_workspace_check_twice(torch_gpu):
    torch = torch_gpu
    stream = torch.cuda.current_stream().cuda_stream
    for n, (w, h) in enumerate([(640, 480), (333, 201)]):  # the second call is smaller
        print(f"static check call {n}: {w}x{h}", flush=True)
        args = checked_frame(torch, w, h, 3 + n, stream)
        want = with_workspace(torch, *args, stream)
        got = without_allocation(torch, *args, stream)
        assert got == want, f"call {n} ({w}x{h}): the static check differs"
        r = _abi.FrameCheckResult.from_buffer_copy(got)
        assert r.squared_error > 0 and r.histogram_mismatches >= 2
        assert list(r.stats) == list(args[4].cpu().numpy())


def test_static_check_refuses_before_launching(torch_gpu):
    torch = torch_gpu
    lut = torch.zeros(16, dtype=torch.float32, device="cuda")
    buf = torch.zeros(1 << 16, dtype=torch.uint8, device="cuda")
    big = cases.demo_camera(width=64, height=64)
    big.width, big.height = 70000, 70000
    status = isp.lib().isp_check_frame_device_static(C.byref(big), lut.data_ptr(), 16, buf.data_ptr(), 0,
                                                     buf.data_ptr(), None, buf.data_ptr(), None)
    assert status == _abi.INVALID_ARGUMENT
    assert "maximum" in isp.last_error()
    cam = cases.demo_camera(width=64, height=64)
    status = isp.lib().isp_check_frame_device_static(C.byref(cam), lut.data_ptr(), 16, buf.data_ptr(), 0,
                                                     None, None, buf.data_ptr(), None)
    assert status == _abi.INVALID_ARGUMENT
