"""The statistics epilogue as the device evaluates it: isp_epilogue_finish_stats_device
runs finish_stats (isp_math.hpp:387-425), the code the main pass's last block runs, with
its double-double log2 / exp2, one case per thread. Every float of the statistics vector,
the auto-exposure fields included, must equal the oracle's, which calls the host libm as
the reference does, on a million randomised sums and settings. (The reference itself
allows 1e-5 on the auto-exposure fields, modules/isp/tests/pipeline_test.cpp:224-228.)"""
import ctypes as C

import numpy as np
import pytest

import holoscan_isp as isp
from holoscan_isp import _abi

import oracle_lib

pytestmark = pytest.mark.gpu

CASE = np.dtype(_abi.EpilogueCase)


def random_cases(n, seed):
    """Sums of frames from 1 pixel to 16384 x 16384 (uniform in the log of the size), means
    anywhere in [0, 1] (a tenth of them near black, where log2 and the clamp at 8 EV meet),
    saturated counts up to the frame, white balance gains in [1/16, 16], every auto-exposure
    setting validation allows, unknown (0) and known capture exposures."""
    rng = np.random.default_rng(seed)
    c = np.zeros(n, dtype=CASE)
    pixels = np.exp(rng.uniform(0.0, np.log(16384.0 * 16384.0), n)).astype(np.int64).clip(1)
    c["pixels"] = pixels
    top = pixels.astype(np.float64) * 1048576.0
    mean = rng.uniform(0.0, 1.0, (n, 3))
    dark = rng.random(n) < 0.1
    mean[dark] *= np.exp2(rng.uniform(-30.0, -6.0, (int(dark.sum()), 1)))
    sums = np.floor(mean * top[:, None]).astype(np.uint64)
    zero = rng.random(n) < 0.01
    sums[zero] = 0
    c["sums"][:, :3] = sums
    c["sums"][:, 3] = np.floor(rng.random(n) * (pixels + 1)).astype(np.uint64)
    c["gains"] = np.exp2(rng.uniform(-4.0, 4.0, (n, 3))).astype(np.float32)
    ae = c["ae"]
    ae["target"] = rng.uniform(0.001, 1.0, n).astype(np.float32)
    ae["damping"] = rng.uniform(0.0, 1.0, n).astype(np.float32)
    ae["min_exposure_ns"] = rng.integers(1, 100_000, n)
    ae["max_exposure_ns"] = ae["min_exposure_ns"] + rng.integers(0, 100_000_000, n)
    ae["min_gain"] = rng.uniform(0.5, 2.0, n).astype(np.float32)
    ae["max_gain"] = ae["min_gain"] * rng.uniform(1.0, 32.0, n).astype(np.float32)
    ex = c["exposure"]
    ex["exposure_ns"] = np.where(rng.random(n) < 0.05, 0, rng.integers(1, 200_000_000, n))
    ex["gain"] = np.where(rng.random(n) < 0.05, 0.0, rng.uniform(0.0, 64.0, n)).astype(np.float32)
    return c


def device_epilogue(cases):
    import torch
    n = cases.size
    d_cases = torch.from_numpy(cases.view(np.uint8).copy()).to("cuda")
    d_out = torch.full((n * _abi.STATS_SIZE,), float("nan"), dtype=torch.float32, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    isp.check(isp.lib().isp_epilogue_finish_stats_device(d_cases.data_ptr(), d_out.data_ptr(), n, stream),
              "epilogue")
    torch.cuda.synchronize()
    return d_out.cpu().numpy().reshape(n, _abi.STATS_SIZE)


def oracle_epilogue(cases):
    out = np.zeros((cases.size, _abi.STATS_SIZE), dtype=np.float32)
    oracle_lib.oracle().oracle_finish_stats_cases(cases.ctypes.data, out.ctypes.data, cases.size)
    return out


def test_case_layout_matches_the_struct():
    assert CASE.itemsize == C.sizeof(_abi.EpilogueCase)


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_device_epilogue_bit_exact_with_libm(torch_gpu, seed):
    cases = random_cases(250_000, seed)
    got = device_epilogue(cases)
    want = oracle_epilogue(cases)
    bad = np.nonzero(np.any(got.view(np.uint32) != want.view(np.uint32), axis=1))[0]
    assert bad.size == 0, (f"{bad.size} of {cases.size} cases differ; first {bad[:5]}: "
                           f"{got[bad[0]]} vs {want[bad[0]]}")
    # The sweep reaches every branch of the suggestion: clamped errors, both bounds.
    err = want[:, _abi.STAT_AE_ERROR_EV]
    assert (err == 8.0).any() and (err == -8.0).any() and ((err > -8.0) & (err < 8.0)).any()
    assert (want[:, _abi.STAT_SUGGESTED_EXPOSURE_NS] == 0.0).any()


def test_device_epilogue_empty_and_errors(torch_gpu):
    assert isp.lib().isp_epilogue_finish_stats_device(None, None, 0, None) == _abi.OK
    assert isp.lib().isp_epilogue_finish_stats_device(None, None, -1, None) == _abi.INVALID_ARGUMENT
    assert isp.lib().isp_epilogue_finish_stats_device(None, None, 4, None) == _abi.INVALID_ARGUMENT
