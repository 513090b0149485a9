"""The kernels' statistics epilogue arithmetic, evaluated on the host by the
same code the device runs (include/isp_c.h isp_epilogue_*), against the host
libm that the reference and the oracle call: the float statistics agree bit
for bit on randomised sums and settings; log2/exp2 agree to within one ulp."""
import ctypes as C
import math

import numpy as np

import holoscan_isp as isp
from holoscan_isp import _abi

import oracle_lib

_libm = C.CDLL("libm.so.6")
_libm.log2.restype = _libm.exp2.restype = C.c_double
_libm.log2.argtypes = _libm.exp2.argtypes = [C.c_double]


def test_log2_exp2_within_one_ulp():
    rng = np.random.default_rng(5)
    L = isp.lib()
    for x in np.exp2(rng.uniform(-40, 40, 20000)):
        a, b = L.isp_epilogue_log2(float(x)), _libm.log2(float(x))
        assert abs(a - b) <= math.ulp(b)
    for y in rng.uniform(-8, 8, 20000):
        a, b = L.isp_epilogue_exp2(float(y)), _libm.exp2(float(y))
        assert abs(a - b) <= math.ulp(b)
    for x in (1.0, 2.0, 0.5, 1024.0, 3.0):
        assert L.isp_epilogue_log2(x) == _libm.log2(x)
    for y in (0.0, 1.0, -1.0, 8.0, -8.0, 0.5):
        assert L.isp_epilogue_exp2(y) == _libm.exp2(y)


def test_finish_stats_bit_exact_with_libm():
    rng = np.random.default_rng(9)
    L = isp.lib()
    o = oracle_lib.oracle()
    bad = 0
    for i in range(100000):
        pixels = int(rng.integers(1, 3840 * 2160))
        sums = (C.c_uint64 * 4)(*[int(rng.integers(0, pixels * 1048576 + 1)) for _ in range(3)],
                                int(rng.integers(0, pixels + 1)))
        if i % 500 == 0:
            sums[0] = sums[1] = sums[2] = 0
        gains = (C.c_float * 3)(1.5, 1.0, 2.0)
        ae = _abi.AutoExposure(float(rng.integers(1, 1000)) / 1000, float(rng.integers(0, 1001)) / 1000,
                               10_000, int(rng.integers(10_000, 50_000_000)), 1.0, 16.0)
        ex = _abi.FrameExposure(int(rng.integers(0, 40_000_000)), float(rng.integers(0, 3000)) / 100)
        a = (C.c_float * 12)()
        b = (C.c_float * 12)()
        L.isp_epilogue_finish_stats(sums, pixels, gains, C.byref(ae), C.byref(ex), a)
        o.oracle_finish_stats(sums, pixels, gains, C.byref(ae), C.byref(ex), b)
        bad += bytes(a) != bytes(b)
    assert bad == 0
