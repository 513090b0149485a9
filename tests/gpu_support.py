"""Running the HIP pipeline on host frames: stage to device memory, process,
copy back (as the reference's test::run, modules/isp/tests/scene_support.hpp:247-305)."""
import numpy as np

import holoscan_isp as isp
from holoscan_isp import _abi
from holoscan_isp.params import output_channel_bytes, output_channels


class Result:
    def __init__(self, image, histogram, stats):
        self.image, self.histogram, self.stats = image, histogram, stats


def to_device(frame):
    import torch
    return torch.from_numpy(frame.data).to("cuda")


def run(pipeline, frame, out_format=_abi.OUT_RGB8, with_stats=True, exposure=(0, 0.0), d_raw=None):
    import torch
    if d_raw is None:
        d_raw = to_device(frame)
    pitch = frame.width * output_channels(out_format) * output_channel_bytes(out_format)
    bins = pipeline.params.histogram_bins or 256
    out = torch.zeros(pitch * frame.height, dtype=torch.uint8, device="cuda")
    hist = torch.zeros(4 * bins, dtype=torch.int32, device="cuda") if with_stats else None
    stats = torch.zeros(12, dtype=torch.float32, device="cuda") if with_stats else None
    stream = torch.cuda.current_stream().cuda_stream
    pipeline.process(frame.view(d_raw), out, pitch, out_format, hist, stats, exposure, stream)
    torch.cuda.synchronize()
    return Result(out.cpu().numpy(),
                  hist.cpu().numpy().view(np.uint32) if with_stats else None,
                  stats.cpu().numpy() if with_stats else None)


def mismatches(a, b):
    return int(np.count_nonzero(a.image != b.image)) + (0 if a.image.size == b.image.size else 1)


def same_stats(a, b):
    return np.array_equal(a.stats.view(np.uint32), b.stats.view(np.uint32))


_hip = None


def device_bytes(address, nbytes):
    """Copies `nbytes` at a device address to a new host uint8 array (hipMemcpy of the HIP
    runtime torch loaded)."""
    import ctypes as C
    import torch  # noqa: F401  (loads the HIP runtime first)
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so.7")
        _hip.hipMemcpy.restype = C.c_int
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    out = np.empty(nbytes, dtype=np.uint8)
    if nbytes:
        status = _hip.hipMemcpy(out.ctypes.data, address, nbytes, 2)  # hipMemcpyDeviceToHost
        assert status == 0, f"hipMemcpy failed: {status}"
    return out
