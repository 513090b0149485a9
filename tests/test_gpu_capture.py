"""The per-frame call captured into a HIP graph (torch.cuda.graph on a side stream) and
replayed: each replay is the pipeline's next frame, bit for bit the oracle's, with eager GPU
work and host synchronisations between the replays. The reference's CUDA backend leaves its
process() capturable (kernel launches, an event record and, across streams, an event wait,
cuda_pipeline.cu:294-364); so does this one."""
import numpy as np
import pytest

import holoscan_isp as isp
from holoscan_isp import _abi

import cases
import gpu_support as gs
from oracle_lib import OraclePipeline

pytestmark = pytest.mark.gpu

EXPOSURE = (10_000_000, 2.0)


@pytest.mark.parametrize("wb,stats", [(_abi.WB_GRAY_WORLD, True), (_abi.WB_WHITE_PATCH, False),
                                      (_abi.WB_MANUAL, True)])
def test_process_under_graph_capture(torch_gpu, wb, stats):
    import torch
    config = cases.camera(_abi.GRBG, cases.FORMATS[5], 200, 120)
    frame = isp.render_synthetic(config)
    params = isp.synthetic_params(config, 5)
    params.white_balance, params.sharpen = wb, (0.5 if stats else 0.0)
    out_format = _abi.OUT_RGBA8
    gpu, cpu = isp.IspPipeline(params), OraclePipeline(params)
    d_raw = gs.to_device(frame)
    pitch = frame.width * 4
    out = torch.zeros(pitch * frame.height, dtype=torch.uint8, device="cuda")
    hist = torch.zeros(4 * 256, dtype=torch.int32, device="cuda") if stats else None
    vec = torch.zeros(12, dtype=torch.float32, device="cuda") if stats else None
    side = torch.cuda.Stream()

    def check(name):
        expect = cpu.process(frame, out_format, stats, EXPOSURE)
        got = gs.Result(out.cpu().numpy(), hist.cpu().numpy().view(np.uint32) if stats else None,
                        vec.cpu().numpy() if stats else None)
        assert gs.mismatches(got, expect) == 0, f"{name}: {gs.mismatches(got, expect)} bytes differ"
        if stats:
            assert np.array_equal(got.histogram, expect.histogram), f"{name}: histograms differ"
            assert gs.same_stats(got, expect), f"{name}: stats differ"

    def process():
        gpu.process(frame.view(d_raw), out, pitch, out_format, hist, vec, EXPOSURE, side.cuda_stream)

    # An eager frame on the side stream first, so the captured call queues behind it on its own
    # stream (no cross-stream wait inside the capture).
    torch.cuda.synchronize()
    with torch.cuda.stream(side):
        process()
    side.synchronize()
    check("eager")
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=side):
        process()
    torch.cuda.synchronize()
    for k in range(3):
        out.zero_()
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        check(f"replay {k}")
        torch.ones(1 << 20, device="cuda").sum().item()  # eager work between replays
