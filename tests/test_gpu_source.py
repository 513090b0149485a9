"""The device-side raw source and the GPU frame checker (include/isp_source.h).

- The device render of the synthetic camera is byte for byte the host render
  (and the reference's own, where oracle/_ref is built) in every format and
  pattern, with padded and unaligned rows, as the reference's
  synthetic_render_matches (pipeline_test.cpp:365-381) requires; bytes past a
  row's samples stay untouched.
- The device check equals the oracle's check (check_frame_host) bit for bit:
  error sum, mismatched bins and statistics, also for planted errors.
- The whole demo chain on the GPU (device source -> ISP -> device checker)
  reproduces the reference demo's published ci run and passes its app test
  matrix (applications/isp/tests/CMakeLists.txt:18-62); the served form fed
  by device-rendered frames passes the same checks."""
import ctypes as C
import time

import numpy as np
import pytest

import holoscan_isp as isp
from holoscan_isp import _abi

import cases
import gpu_support as gs
import oracle_lib

pytestmark = pytest.mark.gpu

SENTINEL = 0xA5


def device_render(torch, cam, pitch, offset=0):
    """Renders into a sentinel-filled device buffer at byte `offset`; returns the host copy of
    the frame's rows (height x pitch bytes)."""
    buf = torch.full((offset + pitch * cam.height + 64,), SENTINEL, dtype=torch.uint8, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    isp.render_synthetic_device(cam, buf.data_ptr() + offset, pitch, stream)
    torch.cuda.synchronize()
    host = buf.cpu().numpy()
    assert np.all(host[:offset] == SENTINEL) and np.all(host[offset + pitch * cam.height:] == SENTINEL)
    return host[offset:offset + pitch * cam.height]


def compare_rows(cam, got, pitch):
    row = isp.lib().isp_raw_row_bytes(cam.format, cam.width)
    want = isp.render_synthetic(cam, pitch).data.reshape(cam.height, pitch)
    got = got.reshape(cam.height, pitch)
    assert np.array_equal(got[:, :row], want[:, :row])
    assert np.all(got[:, row:] == SENTINEL)


@pytest.mark.parametrize("fi", range(len(cases.FORMATS)))
@pytest.mark.parametrize("pattern", cases.PATTERNS)
def test_device_render_matches_host(torch_gpu, fi, pattern):
    print(f"render format {fi} pattern {pattern}", flush=True)
    for kind, frame_no, exposure in ((_abi.SCENE_CHART, 3, 1.0), (_abi.SCENE_GRAY, 17, 1.0),
                                     (_abi.SCENE_CHART, 40, 4.0)):
        cam = cases.camera(pattern, cases.FORMATS[fi], 101, 37, kind=kind, frame=frame_no, exposure=exposure)
        row = isp.lib().isp_raw_row_bytes(cam.format, cam.width)
        for pitch, offset in ((row, 0), ((row + 1) & ~1, 0), (row + 3, 1), (row + 64, 2)):
            compare_rows(cam, device_render(torch_gpu, cam, pitch, offset), pitch)


@pytest.mark.skipif(not oracle_lib.ref_available(), reason="reference backend not built here")
def test_device_render_matches_the_reference(torch_gpu):
    r = oracle_lib.ref()
    for fmt in (cases.FORMATS[0], cases.FORMATS[3], cases.FORMATS[4], cases.FORMATS[5]):
        cam = cases.camera(_abi.GBRG, fmt, 203, 61, frame=9, vignetting=0.3)
        pitch = isp.lib().isp_raw_row_bytes(cam.format, cam.width) + 5
        theirs = np.full(pitch * cam.height, SENTINEL, dtype=np.uint8)
        assert r.ref_synthetic_render(C.byref(cam), theirs.ctypes.data, pitch) == 0
        assert np.array_equal(device_render(torch_gpu, cam, pitch), theirs)


@pytest.mark.parametrize("w,h,fmt", [(3840, 2160, cases.FORMATS[5]), (4096, 3072, cases.FORMATS[2]),
                                     (1920, 1080, cases.FORMATS[4]), (1283, 721, cases.FORMATS[0])])
def test_device_render_large_frames(torch_gpu, w, h, fmt):
    print(f"render {w}x{h} format {fmt[0]}", flush=True)
    cam = cases.camera(_abi.GRBG, fmt, w, h, frame=11)
    pitch = (isp.lib().isp_raw_row_bytes(cam.format, cam.width) + 1) & ~1
    compare_rows(cam, device_render(torch_gpu, cam, pitch), pitch)


def test_device_render_rejects_what_the_host_rejects(torch_gpu):
    buf = torch_gpu.zeros(1 << 16, dtype=torch_gpu.uint8, device="cuda")
    with pytest.raises(ValueError):
        isp.render_synthetic_device(cases.demo_camera(width=8, height=8), buf)
    cam = cases.demo_camera(width=64, height=16)
    with pytest.raises(ValueError):
        isp.render_synthetic_device(cam, buf, pitch=10)
    with pytest.raises(ValueError):
        isp.render_synthetic_device(cam, 0)


class Demo:
    """The demo's source -> IspOp -> checker chain on the GPU, one stream: the device renders
    each frame, the pipeline processes it, the device checker checks it."""

    def __init__(self, torch, camera_kw, demosaic=_abi.MALVAR, white_balance=_abi.WB_MANUAL, sharpen=0.0):
        self.torch = torch
        self.camera_kw = camera_kw
        cam = cases.demo_camera(frame=0, **camera_kw)
        self.params = cases.demo_params(cam, demosaic, white_balance, sharpen)
        self.pipe = isp.IspPipeline(self.params)
        self.checker = isp.FrameChecker(self.params)
        self.w, self.h = cam.width, cam.height
        self.pitch = (isp.lib().isp_raw_row_bytes(cam.format, cam.width) + 1) & ~1
        self.raw = torch.empty(self.pitch * self.h, dtype=torch.uint8, device="cuda")
        self.out = torch.empty(3 * self.w * self.h, dtype=torch.uint8, device="cuda")
        self.hist = torch.empty(1024, dtype=torch.int32, device="cuda")
        self.stats = torch.empty(12, dtype=torch.float32, device="cuda")
        self.stream = torch.cuda.current_stream().cuda_stream

    def frame(self, n, keep=False):
        cam = cases.demo_camera(frame=n, **self.camera_kw)
        isp.render_synthetic_device(cam, self.raw, self.pitch, self.stream)
        view = _abi.RawView(self.raw.data_ptr(), cam.width, cam.height, self.pitch, cam.format, cam.bit_depth,
                            cam.pattern)
        self.pipe.process(view, self.out, 3 * self.w, _abi.OUT_RGB8, self.hist, self.stats, cases.DEMO_EXPOSURE,
                          self.stream)
        self.checker.check(cam, self.out, self.hist, self.stats, 0, self.stream)
        return cam, self.checker.result()

    def run(self, frames):
        """(psnr per frame, mismatched-bin counts, last statistics)."""
        psnr, miss = [], []
        for n in range(frames):
            _, r = self.frame(n)
            psnr.append(isp.frame_psnr(r, self.w, self.h))
            miss.append(r.histogram_mismatches)
        return psnr, miss, list(r.stats)


def test_device_check_equals_the_oracle(torch_gpu):
    demo = Demo(torch_gpu, dict(width=333, height=201, pattern=_abi.GRBG, bits=12, packing="raw12"),
                sharpen=0.7)
    for n in (0, 7):
        print(f"check frame {n}", flush=True)
        cam, got = demo.frame(n)
        image = demo.out.cpu().numpy().reshape(demo.h, demo.w, 3)
        hist = demo.hist.cpu().numpy().view(np.uint32)
        stats = demo.stats.cpu().numpy()
        lut = isp.tone_lut(demo.params)
        truth = isp.synthetic_truth(cam)
        want = oracle_lib.frame_check(truth, lut, image, hist, stats)
        assert got.squared_error == want.squared_error > 0
        assert got.histogram_mismatches == want.histogram_mismatches == 0
        assert bytes(got.stats) == bytes(want.stats) == stats.tobytes()
    # Planted errors: two pixels changed, two bins changed; no statistics (zeros).
    flat = demo.out.view(-1)
    flat[5] = (int(flat[5]) + 77) % 256
    flat[3 * demo.w * 100 + 2] = 255 - int(flat[3 * demo.w * 100 + 2])
    demo.hist[3] += 1
    demo.hist[700] += 5
    demo.checker.check(cam, demo.out, demo.hist, None, 0, demo.stream)
    got = demo.checker.result()
    image = demo.out.cpu().numpy().reshape(demo.h, demo.w, 3)
    want = oracle_lib.frame_check(truth, lut, image, demo.hist.cpu().numpy().view(np.uint32))
    assert got.squared_error == want.squared_error
    assert got.histogram_mismatches == want.histogram_mismatches >= 2
    assert list(got.stats) == [0.0] * 12


def test_device_check_with_padded_rows(torch_gpu):
    demo = Demo(torch_gpu, dict(width=130, height=67, pattern=_abi.BGGR, bits=10, packing="raw10"))
    cam, r0 = demo.frame(2)
    padded = torch_gpu.zeros((demo.h, 3 * demo.w + 13), dtype=torch_gpu.uint8, device="cuda")
    padded[:, :3 * demo.w] = demo.out.view(demo.h, 3 * demo.w)
    demo.checker.check(cam, padded, demo.hist, demo.stats, padded.stride(0), demo.stream)
    r1 = demo.checker.result()
    assert bytes(r0) == bytes(r1)


def test_demo_ci_run_on_the_gpu_reproduces_the_published_numbers(torch_gpu):
    psnr, miss, stats = Demo(torch_gpu, {}).run(30)
    assert miss == [0] * 30
    assert f"{np.mean(psnr):.2f}" == "30.16"
    assert f"{min(psnr):.2f}" == "30.12"
    assert f"{stats[_abi.STAT_WB_GAIN_R]:.3f},{stats[_abi.STAT_WB_GAIN_G]:.3f},{stats[_abi.STAT_WB_GAIN_B]:.3f}" \
        == "1.900,1.000,1.600"
    assert f"{stats[_abi.STAT_MEAN_LUMA]:.3f}" == "0.275"
    assert f"{stats[_abi.STAT_AE_ERROR_EV]:.3f}" == "-0.613"


# The demo's app tests (tests/CMakeLists.txt:31-51), 10 frames of 640x480 each:
# (name, camera options, demosaic, white balance, sharpen, min PSNR or None for the default).
APP_TESTS = [
    ("malvar", {}, _abi.MALVAR, _abi.WB_MANUAL, 0.0, None),
    ("bilinear", {}, _abi.BILINEAR, _abi.WB_MANUAL, 0.0, None),
    ("8bit", dict(bits=8, pattern=_abi.BGGR), _abi.MALVAR, _abi.WB_MANUAL, 0.0, None),
    ("16bit_sharpened", dict(bits=16, pattern=_abi.GRBG), _abi.MALVAR, _abi.WB_MANUAL, 0.3, 26.0),
    ("raw10", dict(packing="raw10", pattern=_abi.GBRG), _abi.MALVAR, _abi.WB_MANUAL, 0.0, None),
    ("raw12", dict(packing="raw12"), _abi.MALVAR, _abi.WB_MANUAL, 0.0, None),
    ("gray_world", dict(scene=_abi.SCENE_GRAY), _abi.MALVAR, _abi.WB_GRAY_WORLD, 0.0, None),
    ("white_patch", {}, _abi.MALVAR, _abi.WB_WHITE_PATCH, 0.0, None),
]


@pytest.mark.parametrize("name,camera_kw,demosaic,wb,sharpen,bound", APP_TESTS, ids=[t[0] for t in APP_TESTS])
def test_demo_app_matrix(torch_gpu, name, camera_kw, demosaic, wb, sharpen, bound):
    print(f"demo {name}", flush=True)
    demo = Demo(torch_gpu, camera_kw, demosaic, wb, sharpen)
    psnr, miss, stats = demo.run(10)
    cam = cases.demo_camera(**camera_kw)
    if bound is None:
        bound = cases.demo_min_psnr(cam.bit_depth, demosaic)
    assert miss == [0] * 10
    assert min(psnr) > bound, (name, psnr)
    if wb != _abi.WB_MANUAL:
        # The estimate finds the camera's gains where the scene allows it (main.cpp:404-418).
        gains = [stats[_abi.STAT_WB_GAIN_R], stats[_abi.STAT_WB_GAIN_G], stats[_abi.STAT_WB_GAIN_B]]
        for g, want in zip(gains, cam.wb_gains):
            assert abs(g / want - 1.0) < 0.02, (name, gains)
    if name == "gray_world":
        s = f"{stats[_abi.STAT_WB_GAIN_R]:.3f},{stats[_abi.STAT_WB_GAIN_G]:.3f},{stats[_abi.STAT_WB_GAIN_B]:.3f}"
        assert s.startswith("1.9") and ",1.000,1." in s and s.split(",")[2][2] in "56", s


def test_demo_fails_a_bound_no_frame_can_meet(torch_gpu):
    psnr, _, _ = Demo(torch_gpu, {}).run(3)
    assert not min(psnr) > 99.0  # isp_app_fails_below_bound (tests/CMakeLists.txt:58)


def test_checker_keeps_up_with_a_10ms_camera_at_1080p(torch_gpu):
    """isp_app_1080p (tests/CMakeLists.txt:41-44): 60 frames at 1920x1080, each rendered,
    processed and checked well within the 10 ms sensor period."""
    demo = Demo(torch_gpu, dict(width=1920, height=1080))
    demo.run(2)
    t0 = time.perf_counter()
    psnr, miss, _ = demo.run(60)
    per_frame = (time.perf_counter() - t0) / 60
    assert miss == [0] * 60 and min(psnr) > 28.0
    assert per_frame < 0.010, per_frame


def test_served_cameras_fed_from_device_renders(torch_gpu):
    """Four demo cameras of different sizes and formats rendered straight into HBM and
    served (isp_server_*, device frames, no PCIe copy), every result checked on the GPU."""
    torch = torch_gpu
    specs = [dict(width=640, height=480), dict(width=320, height=240, packing="raw10", pattern=_abi.BGGR),
             dict(width=1280, height=720, packing="raw12", pattern=_abi.GRBG), dict(width=256, height=144, bits=8)]
    server = isp.IspServer(hbm_budget=64 << 20, max_batch=8)
    stream = torch.cuda.current_stream().cuda_stream
    cams = []
    for spec in specs:
        cam = cases.demo_camera(**spec)
        params = cases.demo_params(cam)
        cid = server.add_camera(params, cam.width, cam.height, cam.format, cam.bit_depth, cam.pattern,
                                _abi.OUT_RGB8, True)
        pitch = (isp.lib().isp_raw_row_bytes(cam.format, cam.width) + 1) & ~1
        cams.append((spec, cid, pitch, isp.FrameChecker(params)))
    for n in range(4):
        bufs = []
        for spec, cid, pitch, _ in cams:
            cam = cases.demo_camera(frame=n, **spec)
            buf = torch.empty(pitch * cam.height, dtype=torch.uint8, device="cuda")
            isp.render_synthetic_device(cam, buf, pitch, stream)
            bufs.append(buf)
        torch.cuda.synchronize()  # the renders have landed before the server reads them
        for (spec, cid, pitch, _), buf in zip(cams, bufs):
            server.submit(cid, buf, pitch, _abi.LOCATION_DEVICE, tag=n)
        done = 0
        while done < len(cams):
            r = server.poll(2_000_000)
            assert r is not None and r.status == 0
            spec, cid, pitch, checker = next(c for c in cams if c[1] == r.camera)
            cam = cases.demo_camera(frame=int(r.tag), **spec)
            checker.check(cam, r.out, r.histogram, r.stats, r.out_pitch, stream)
            res = checker.result()
            assert res.histogram_mismatches == 0
            assert isp.frame_psnr(res, cam.width, cam.height) > cases.demo_min_psnr(cam.bit_depth, _abi.MALVAR)
            server.release(r.request)
            done += 1
    server.close()
