"""Frames queued on two streams in turn, never waited for in between, come out
as the CPU oracle's in call order (the reference's two_streams,
modules/isp/tests/pipeline_test.cpp:270-362): each frame's GPU work waits for
the previous frame's, which shares the scratch and the white balance state, so
the smoothed gray-world estimate of every frame depends on the order. Unlike
the reference, which leaves the auto-exposure fields out, the whole statistics
vector is compared bit for bit."""
import numpy as np
import pytest

import holoscan_isp as isp
from holoscan_isp import _abi

import cases
from oracle_lib import OraclePipeline

pytestmark = pytest.mark.gpu

FRAMES = 8


def two_stream_frames():
    config = cases.camera(_abi.RGGB, cases.FORMATS[2], 1920, 1080)
    params = isp.synthetic_params(config, 9)
    params.white_balance = _abi.WB_GRAY_WORLD
    params.sharpen = 0.5
    frames = []
    for f in range(FRAMES):
        config.frame = f
        # Gains that change from frame to frame, so the smoothed estimate depends on the order.
        config.wb_gains[:] = [1.9 if f % 2 == 0 else 1.4, 1.0, 1.6 if f % 2 == 0 else 2.1]
        frames.append(isp.render_synthetic(config))
    return params, frames


@pytest.mark.parametrize("exposure", [(0, 0.0), (10_000_000, 1.5)], ids=["no exposure", "exposure"])
def test_two_streams_in_call_order(torch_gpu, exposure):
    import torch
    params, frames = two_stream_frames()
    cpu = OraclePipeline(params)
    expected = [cpu.process(fr, _abi.OUT_RGB8, True, exposure, threads=16) for fr in frames]
    width, height = frames[0].width, frames[0].height
    pitch = 3 * width
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]  # default priority
    gpu = isp.IspPipeline(params)
    raws = [torch.from_numpy(fr.data).to("cuda") for fr in frames]
    images = [torch.zeros(pitch * height, dtype=torch.uint8, device="cuda") for _ in frames]
    hists = [torch.zeros(4 * 256, dtype=torch.int32, device="cuda") for _ in frames]
    stats = [torch.zeros(12, dtype=torch.float32, device="cuda") for _ in frames]
    torch.cuda.synchronize()
    for f, fr in enumerate(frames):
        gpu.process(fr.view(raws[f]), images[f], pitch, _abi.OUT_RGB8, hists[f], stats[f], exposure,
                    streams[f % 2].cuda_stream)
    torch.cuda.synchronize()
    for f in range(FRAMES):
        h = expected[f]
        image = images[f].cpu().numpy()
        assert np.count_nonzero(image != h.image) == 0, f"frame {f}: the image differs"
        assert np.array_equal(hists[f].cpu().numpy().view(np.uint32), h.histogram), f"frame {f}: histograms differ"
        got = stats[f].cpu().numpy()
        assert np.array_equal(got.view(np.uint32), h.stats.view(np.uint32)), f"frame {f}: stats {got} vs {h.stats}"
    # The gray-world state moved from frame to frame (the test would not see a reordering otherwise).
    gains = {tuple(e.stats[8:11]) for e in expected}
    assert len(gains) == FRAMES
