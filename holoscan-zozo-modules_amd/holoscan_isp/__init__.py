"""holoscan_isp: Python access to the MI355X ISP through its C ABI (include/isp_c.h).

`IspPipeline` mirrors holoscan::isp::IspPipeline (reference
modules/isp/include/holoscan/isp/isp_pipeline.hpp:39-74): construct with
IspParams, then process() frames whose buffers live in device memory (torch
tensors on the ROCm device). Errors raise ValueError (std::invalid_argument)
or RuntimeError (std::runtime_error) with the library's message.

The library (libholoscan_isp_amd.so, built in-tree by `make -C
holoscan-zozo-modules_amd`) must be present: there is no fallback.
"""
import ctypes as C
import os

import numpy as np

from . import _abi
from ._abi import *  # noqa: F401,F403  (enums and constants)
from .params import (AutoExposureParams, IspParams, LensShading, output_channel_bytes,  # noqa: F401
                     output_channels, raw_row_bytes)

PACKAGE_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBRARY = os.path.join(PACKAGE_DIR, "libholoscan_isp_amd.so")

_lib = None


def lib():
    """The bound C library; raises if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBRARY):
            raise RuntimeError(f"{LIBRARY} is missing: build it with `make -C {PACKAGE_DIR}` "
                               "(or __graft_entry__.build())")
        # torch (device memory and streams) must load its HIP runtime first: the
        # library then binds to the same libamdhip64.so.7, so one runtime serves both.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        _lib = _abi.bind(C.CDLL(LIBRARY))
    return _lib


def last_error():
    return lib().isp_last_error().decode()


def check(status, what="isp"):
    if status == _abi.OK:
        return
    message = f"{what}: {last_error()}"
    if status == _abi.INVALID_ARGUMENT:
        raise ValueError(message)
    if status == _abi.OUT_OF_MEMORY:
        raise MemoryError(message)
    raise RuntimeError(message)


def default_params():
    """IspParams as isp_params_init() fills them."""
    p = _abi.Params()
    lib().isp_params_init(C.byref(p))
    return IspParams.from_c(p)


def tone_lut(params):
    c, keep = params.to_c()
    size = C.c_int32()
    check(lib().isp_tone_lut(C.byref(c), None, 0, C.byref(size)), "tone_lut")
    out = np.zeros(size.value, dtype=np.float32)
    check(lib().isp_tone_lut(C.byref(c), out.ctypes.data_as(C.POINTER(C.c_float)), size.value,
                             C.byref(size)), "tone_lut")
    return out


def validate_params(params):
    c, keep = params.to_c()
    check(lib().isp_validate_params(C.byref(c)), "validate_params")


def _ptr(x):
    """Device or host address of a torch tensor / numpy array / int."""
    if x is None:
        return None
    if isinstance(x, int):
        return x
    if hasattr(x, "data_ptr"):
        return x.data_ptr()
    return x.ctypes.data


class RawFrame:
    """A raw mosaic in host memory: `data` uint8 [height * pitch] and its view fields."""

    def __init__(self, data, width, height, pitch, fmt, bit_depth, pattern):
        self.data = np.ascontiguousarray(data, dtype=np.uint8)
        self.width, self.height, self.pitch = int(width), int(height), int(pitch)
        self.format, self.bit_depth, self.pattern = int(fmt), int(bit_depth), int(pattern)

    def view(self, data=None):
        return _abi.RawView(_ptr(data) if data is not None else self.data.ctypes.data,
                            self.width, self.height, self.pitch, self.format, self.bit_depth,
                            self.pattern)

    @property
    def pixels(self):
        return self.width * self.height


def synthetic_config(**kw):
    """SyntheticRawConfig (synthetic_raw.hpp:40-59) with its defaults, fields overridden by kw."""
    c = _abi.SyntheticConfig()
    lib().isp_synthetic_config_init(C.byref(c))
    for key, value in kw.items():
        if isinstance(value, (list, tuple)):
            arr = getattr(c, key)
            for i, v in enumerate(value):
                arr[i] = v
        else:
            setattr(c, key, value)
    return c


def render_synthetic(config, pitch=None):
    """render_synthetic_raw into host memory; rows padded to even bytes as the reference's tests do."""
    row = lib().isp_raw_row_bytes(config.format, config.width)
    if pitch is None:
        pitch = (row + 1) & ~1
    data = np.zeros(pitch * config.height, dtype=np.uint8)
    check(lib().isp_synthetic_render(C.byref(config), data.ctypes.data, pitch), "synthetic_render")
    return RawFrame(data, config.width, config.height, pitch, config.format, config.bit_depth,
                    config.pattern)


def synthetic_params(config, grid=17):
    """synthetic_isp_params(config, grid) as IspParams."""
    p = _abi.Params()
    gains = (C.c_float * (grid * grid))()
    check(lib().isp_synthetic_params(C.byref(config), grid, C.byref(p), gains), "synthetic_params")
    return IspParams.from_c(p)


def synthetic_truth(config):
    """The truth image, float32 [H, W, 3]."""
    out = np.zeros((config.height, config.width, 3), dtype=np.float32)
    check(lib().isp_synthetic_truth_image(C.byref(config), out.ctypes.data_as(C.POINTER(C.c_float))),
          "synthetic_truth")
    return out


def scene_config(**kw):
    c = _abi.SceneConfig()
    lib().isp_scene_config_init(C.byref(c))
    for key, value in kw.items():
        if isinstance(value, (list, tuple)):
            arr = getattr(c, key)
            for i, v in enumerate(value):
                arr[i] = v
        else:
            setattr(c, key, value)
    return c


def render_scene(config, pitch=None):
    row = lib().isp_raw_row_bytes(config.format, config.width)
    if pitch is None:
        pitch = (row + 15) & ~15 if config.format != _abi.RAW_U16 else row
    data = np.zeros(pitch * config.height, dtype=np.uint8)
    check(lib().isp_scene_render(C.byref(config), data.ctypes.data, pitch), "scene_render")
    return RawFrame(data, config.width, config.height, pitch, config.format, config.bit_depth,
                    config.pattern)


def scene_lens_shading(config, gw=17, gh=13):
    gains = np.zeros(gw * gh, dtype=np.float32)
    check(lib().isp_scene_lens_shading(C.byref(config), gw, gh,
                                       gains.ctypes.data_as(C.POINTER(C.c_float))), "scene_lens_shading")
    return LensShading(gw, gh, 1, gains)


class IspPipeline:
    """IspPipeline (isp_pipeline.hpp:39-74) on the GPU; buffers are device addresses."""

    def __init__(self, params, backend=_abi.BACKEND_GPU):
        self.params = params
        self._c, self._keep = params.to_c()
        h = C.c_void_p()
        check(lib().isp_create(C.byref(self._c), backend, C.byref(h)), "IspPipeline")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().isp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # interpreter shutdown
            pass

    def backend(self):
        return lib().isp_backend(self._h)

    def location(self):
        return lib().isp_location(self._h)

    def reset(self):
        check(lib().isp_reset(self._h), "reset")

    def timing_begin(self):
        check(lib().isp_timing_begin(self._h), "timing_begin")

    def timing_end(self):
        """(awb_ms, main_ms, frames): summed kernel times since timing_begin()."""
        a, m, n = C.c_double(), C.c_double(), C.c_int32()
        check(lib().isp_timing_end(self._h, C.byref(a), C.byref(m), C.byref(n)), "timing_end")
        return a.value, m.value, n.value

    def process(self, raw_view, out, out_pitch, out_format, histogram=None, stats=None,
                exposure=(0, 0.0), stream=None):
        """Queues one frame: raw_view an _abi.RawView with a device address; out, histogram and
        stats device tensors/addresses; stream a hipStream_t as int (None: the null stream)."""
        rgb = _abi.RgbView(_ptr(out), out_pitch, out_format)
        so = _abi.StatsOut(_ptr(histogram), _ptr(stats))
        e = _abi.FrameExposure(int(exposure[0]), float(exposure[1]))
        check(lib().isp_process(self._h, C.byref(raw_view), C.byref(rgb), C.byref(so), C.byref(e),
                                stream), "process")

    def set_exposure_feedback(self, initial):
        """Device exposure feedback: each frame runs at its predecessor's suggestion,
        the first at `initial` = (exposure_ns, gain); None switches it off."""
        if initial is None:
            check(lib().isp_set_exposure_feedback(self._h, 0, None), "set_exposure_feedback")
        else:
            e = _abi.FrameExposure(int(initial[0]), float(initial[1]))
            check(lib().isp_set_exposure_feedback(self._h, 1, C.byref(e)), "set_exposure_feedback")

    def exposure_state(self):
        e = _abi.FrameExposure()
        check(lib().isp_get_exposure_state(self._h, C.byref(e)), "exposure_state")
        return int(e.exposure_ns), float(e.gain)


def default_schedule():
    s = _abi.BatchSchedule()
    lib().isp_batch_schedule_init(C.byref(s))
    return s


class IspBatch:
    """Many streams' frames in one persistent launch (isp_batch_*). Each frame is
    (pipeline, raw_view, out, out_pitch, out_format, histogram, stats, exposure)."""

    def __init__(self, schedule=None):
        h = C.c_void_p()
        check(lib().isp_batch_create(C.byref(h)), "IspBatch")
        self._h = h
        self._frames = None
        if schedule is not None:
            self.set_schedule(schedule)

    def close(self):
        if getattr(self, "_h", None):
            lib().isp_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # interpreter shutdown
            pass

    def set_schedule(self, schedule):
        check(lib().isp_batch_set_schedule(self._h, C.byref(schedule)), "batch set_schedule")

    def schedule(self):
        s = _abi.BatchSchedule()
        check(lib().isp_batch_get_schedule(self._h, C.byref(s)), "batch schedule")
        return s

    def set_frames(self, frames):
        arr = (_abi.BatchFrame * len(frames))()
        for i, (pipe, raw_view, out, out_pitch, out_format, histogram, stats, exposure) in enumerate(frames):
            arr[i].pipeline = pipe._h
            arr[i].raw = raw_view
            arr[i].out = _abi.RgbView(_ptr(out), out_pitch, out_format)
            arr[i].stats = _abi.StatsOut(_ptr(histogram), _ptr(stats))
            e = exposure if exposure is not None else (0, 0.0)
            arr[i].exposure = _abi.FrameExposure(int(e[0]), float(e[1]))
        check(lib().isp_batch_set_frames(self._h, arr, len(frames)), "batch set_frames")
        self._frames = frames  # keeps the pipelines alive

    def launch(self, stream=None):
        check(lib().isp_batch_launch(self._h, stream), "batch launch")

    def synchronize(self):
        """Waits for the batch's launches."""
        check(lib().isp_batch_synchronize(self._h), "batch synchronize")

    def tune(self, stream=None, cache_path=None, launches=5):
        s = _abi.BatchSchedule()
        ms = C.c_double()
        path = cache_path.encode() if cache_path else None
        check(lib().isp_batch_tune(self._h, stream, path, launches, C.byref(s), C.byref(ms)), "batch tune")
        return s, ms.value


def server_config(**kw):
    c = _abi.ServerConfig()
    lib().isp_server_config_init(C.byref(c))
    for key, value in kw.items():
        setattr(c, key, value)
    return c


class IspServer:
    """The served form (isp_server_*): cameras submit frames as they arrive; a native
    scheduling thread admits them against an HBM budget and runs them in continuous batches."""

    def __init__(self, config=None, **kw):
        config = config if config is not None else server_config(**kw)
        h = C.c_void_p()
        check(lib().isp_server_create(C.byref(config), C.byref(h)), "IspServer")
        self._h = h
        self.config = config
        self._cameras = []

    def close(self):
        if getattr(self, "_h", None):
            lib().isp_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # interpreter shutdown
            pass

    def add_camera(self, params, width, height, fmt, bit_depth, pattern, out_format, stats=True,
                   initial=(0, 0.0)):
        c = _abi.CameraConfig()
        cp, keep = params.to_c()
        c.params = cp
        c.width, c.height, c.format, c.bit_depth, c.pattern = width, height, fmt, bit_depth, pattern
        c.out_format, c.stats = out_format, 1 if stats else 0
        c.initial = _abi.FrameExposure(int(initial[0]), float(initial[1]))
        cam = C.c_int32()
        check(lib().isp_server_add_camera(self._h, C.byref(c), C.byref(cam)), "add_camera")
        self._cameras.append((params, keep))
        return cam.value

    def submit(self, camera, raw, pitch, location=_abi.LOCATION_HOST, tag=0):
        """`raw`: a host numpy array (pinned for an asynchronous copy) or a device tensor/address;
        it must stay valid until the result is polled. Returns the request id."""
        rid = C.c_int64()
        check(lib().isp_server_submit(self._h, camera, _ptr(raw), pitch, location, tag, C.byref(rid)), "submit")
        return rid.value

    def poll(self, timeout_us=0):
        """The next finished frame's ServerResult, or None."""
        r = _abi.ServerResult()
        status = lib().isp_server_poll(self._h, int(timeout_us), C.byref(r))
        if status == _abi.NOT_READY:
            return None
        check(status, "poll")
        return r

    def release(self, request):
        check(lib().isp_server_release(self._h, request), "release")

    def counters(self):
        c = _abi.ServerCounters()
        check(lib().isp_server_counters_get(self._h, C.byref(c)), "counters")
        return c


def now_ns():
    return lib().isp_server_now_ns()


def render_synthetic_device(config, data, pitch=None, stream=None):
    """render_synthetic_raw_cuda (synthetic_raw_cuda.hpp:17-19): the frame of `config` into
    device memory `data` (a tensor or address), rows `pitch` bytes apart (None: the packed row
    rounded up to even bytes, as render_synthetic), queued on `stream`. Returns the pitch."""
    row = lib().isp_raw_row_bytes(config.format, config.width)
    if pitch is None:
        pitch = (row + 1) & ~1
    check(lib().isp_synthetic_render_device(C.byref(config), _ptr(data), pitch, stream), "render_synthetic_device")
    return pitch


class FrameChecker:
    """The demo checker's device state (check_op.cpp:104-140) around check_frame_cuda
    (frame_check.hpp:70-76): the tone table of `params`, the zeroed scratch and the 64-byte
    result in device memory. check() queues one frame's check; result() reads it back."""

    def __init__(self, params):
        import torch
        self.lut = torch.from_numpy(np.asarray(tone_lut(params), dtype=np.float32)).cuda()
        self.scratch = torch.zeros(FRAME_CHECK_SCRATCH_BYTES, dtype=torch.uint8, device="cuda")
        self.device_result = torch.zeros(C.sizeof(_abi.FrameCheckResult), dtype=torch.uint8, device="cuda")

    def check(self, camera, rgb, histogram, stats=None, pitch=0, stream=None):
        check(lib().isp_check_frame_device(C.byref(camera), _ptr(self.lut), int(self.lut.numel()), _ptr(rgb), pitch,
                                           _ptr(histogram), _ptr(stats), _ptr(self.scratch),
                                           _ptr(self.device_result), stream), "check_frame_device")

    def result(self):
        """The last check's isp_frame_check_result (waits for the device)."""
        host = self.device_result.cpu().numpy()
        return _abi.FrameCheckResult.from_buffer_copy(host.tobytes())


def frame_psnr(result, width, height):
    """PSNR in dB of a checked frame (check_op.cpp:226-228)."""
    return lib().isp_frame_check_psnr(C.byref(result), width, height)
