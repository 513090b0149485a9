"""IspOp (modules/isp/include/holoscan/isp/isp_op.hpp) over the C ABI (include/isp_op.h).

The operator's config fields, defaults, ports, shapes, dtypes, descriptor passthrough
and emit order are the reference's. With no Holoscan SDK here, the runtime side is a
small object: it lends output tensors (device memory from torch) and records what the
operator emits, in order."""
import ctypes as C
from dataclasses import dataclass, field
from typing import Optional

from . import _abi, check, lib
from .params import IspParams


@dataclass
class IspOpConfig:
    """IspOpConfig (isp_op.hpp:38-55) with its defaults."""
    backend: int = _abi.BACKEND_GPU
    params: IspParams = field(default_factory=IspParams)
    output_format: int = _abi.OUT_RGB8
    pattern: Optional[int] = None
    bit_depth: int = 0
    packing: int = _abi.PACKING_NONE
    stats: bool = False
    max_width: int = 3840
    max_height: int = 2160

    def to_c(self):
        c = _abi.OpConfig()
        lib().isp_op_config_init(C.byref(c))
        p, keep = self.params.to_c()
        c.backend, c.params, c.output_format = self.backend, p, self.output_format
        c.has_pattern = 0 if self.pattern is None else 1
        c.pattern = 0 if self.pattern is None else self.pattern
        c.bit_depth, c.packing, c.stats = self.bit_depth, self.packing, 1 if self.stats else 0
        c.max_width, c.max_height = self.max_width, self.max_height
        return c, keep


def validate_config(config):
    c, keep = config.to_c()
    check(lib().isp_op_validate_config(C.byref(c)), "validate_config")


def statistics():
    """isp_op_statistics(): (host_waits, copies) of every IspOp in the process."""
    s = lib().isp_op_statistics()
    return int(s.host_waits), int(s.copies)


def tensor(data_ptr, memory, dtype_code, dtype_bits, shape, strides=None, ready=None):
    """An _abi.Tensor; `strides` in elements (row-major contiguous when omitted)."""
    t = _abi.Tensor()
    t.data, t.memory, t.dtype_code, t.dtype_bits, t.ndim = data_ptr, memory, dtype_code, dtype_bits, len(shape)
    if strides is None:
        strides, step = [], 1
        for n in reversed(shape):
            strides.insert(0, step)
            step *= n
    for i, (n, st) in enumerate(zip(shape, strides)):
        t.shape[i], t.strides[i] = n, st
    t.ready = ready
    return t


class Emitted:
    """One emitted output: its port, the descriptor (image) or None, the tensor's bytes
    copied to the host at emit time, its shape and dtype, and the sample metadata."""

    def __init__(self, port, descriptor, data, shape, dtype, metadata):
        self.port, self.descriptor, self.data, self.shape, self.dtype = port, descriptor, data, shape, dtype
        self.metadata = metadata


class TorchRuntime:
    """The runtime side of compute(): lends device tensors from torch, records the emits."""

    def __init__(self, stream):
        self.stream = stream
        self.emitted = []
        self._lent = {}
        self._alloc = _abi.ALLOCATE_FN(self._allocate)
        self._emit = _abi.EMIT_FN(self._on_emit)
        self.c = _abi.OpRuntime(None, self._alloc, self._emit)

    def _allocate(self, user, port, request, out):
        import torch
        r = request.contents
        dtype = {(_abi.DTYPE_UINT, 8): torch.uint8, (_abi.DTYPE_UINT, 16): torch.int16,
                 (_abi.DTYPE_UINT, 32): torch.int32, (_abi.DTYPE_FLOAT, 32): torch.float32}[(r.dtype_code, r.dtype_bits)]
        shape = [r.shape[i] for i in range(r.ndim)]
        t = torch.empty(shape, dtype=dtype, device="cuda")
        self._lent[port.decode()] = t
        o = out.contents
        C.memmove(C.addressof(o), C.addressof(r), C.sizeof(_abi.Tensor))
        o.data = t.data_ptr()
        return 0

    def _on_emit(self, user, port, image, tens, metadata):
        import torch
        name = port.decode()
        t = self._lent.pop(name)
        torch.cuda.current_stream().synchronize()
        desc = None
        if image:
            d = image.contents
            desc = {f: getattr(d, f) for f, _ in _abi.Image._fields_ if f != "data"}
        md = metadata.contents
        self.emitted.append(Emitted(name, desc, t.cpu().numpy().tobytes(), tuple(t.shape), t.dtype,
                                    (md.capture_timestamp_ns, md.frame_id)))
        return 0


class IspOp:
    """IspOp (isp_op.hpp:74-121): construct, setup() ports, start(), compute() per raw
    frame, stop()."""

    def __init__(self, config):
        self.config = config
        self._c, self._keep = config.to_c()
        h = C.c_void_p()
        check(lib().isp_op_create(C.byref(self._c), C.byref(h)), "IspOp")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().isp_op_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # interpreter shutdown
            pass

    def backend(self):
        return lib().isp_op_backend(self._h)

    def setup(self):
        """The declared ports: [(name, is_input, queue_depth, memory, (code, bits), rank, max_bytes)]."""
        n = C.c_int32()
        check(lib().isp_op_ports(self._h, None, 0, C.byref(n)), "setup")
        ports = (_abi.OpPort * n.value)()
        check(lib().isp_op_ports(self._h, ports, n.value, C.byref(n)), "setup")
        return [(p.name.decode(), bool(p.is_input), p.queue_depth, p.memory, (p.dtype_code, p.dtype_bits), p.rank,
                 p.max_bytes) for p in ports]

    def start(self):
        check(lib().isp_op_start(self._h), "start")

    def stop(self):
        check(lib().isp_op_stop(self._h), "stop")

    def compute(self, image, metadata, runtime, stream=None):
        """One raw sample: `image` an _abi.Image, `metadata` (capture_timestamp_ns, frame_id)."""
        md = _abi.SampleMetadata(int(metadata[0]), int(metadata[1]))
        check(lib().isp_op_compute(self._h, C.byref(image), C.byref(md), stream, C.byref(runtime.c)), "compute")
