"""The batched form (isp_batch_*): many streams of mixed raw formats, sizes,
patterns and settings in one persistent launch. Every frame's image, histogram
and statistics must equal, bit for bit, what the stream's own pipeline gives
per frame (isp_process) and what the CPU oracle gives, over several ticks (the
white-balance smoothing and the device exposure state carry across ticks),
under every schedule the autotuner can choose, and on hipGraph replays with
other GPU work between them."""
import os

import numpy as np
import pytest

import holoscan_isp as isp
from holoscan_isp import _abi
from holoscan_isp.params import output_channel_bytes, output_channels

import cases
import gpu_support as gs
from oracle_lib import OraclePipeline

pytestmark = pytest.mark.gpu

EXPOSURE = (10_000_000, 2.0)


def mixed_specs():
    """Small frames covering every raw format, pattern, demosaic, sharpening, white balance mode,
    output format, both histogram sizes and odd sizes."""
    specs = []
    rows = [
        # (w, h, fmt, bits, pattern, black, demosaic, sharpen, wb, out, bins, tone)
        (320, 180, _abi.RAW_RAW10, 10, _abi.BGGR, 64.0, _abi.BILINEAR, 0.0, _abi.WB_GRAY_WORLD, _abi.OUT_RGB8, 256, _abi.TONE_SRGB),
        (256, 160, _abi.RAW_RAW12, 12, _abi.RGGB, 240.0, _abi.MALVAR, 0.5, _abi.WB_WHITE_PATCH, _abi.OUT_RGBA8, 256, _abi.TONE_SRGB),
        (200, 122, _abi.RAW_U16, 14, _abi.GBRG, 1024.0, _abi.MALVAR, 0.0, _abi.WB_MANUAL, _abi.OUT_RGB16, 1024, _abi.TONE_GAMMA),
        (130, 66, _abi.RAW_U8, 8, _abi.GRBG, 16.0, _abi.BILINEAR, 1.0, _abi.WB_GRAY_WORLD, _abi.OUT_RGB32F, 256, _abi.TONE_LINEAR),
        (67, 35, _abi.RAW_RAW12, 12, _abi.GRBG, 200.0, _abi.MALVAR, 1.5, _abi.WB_GRAY_WORLD, _abi.OUT_RGB8, 256, _abi.TONE_SRGB),
        (33, 17, _abi.RAW_RAW10, 10, _abi.RGGB, 60.0, _abi.BILINEAR, 0.0, _abi.WB_MANUAL, _abi.OUT_RGBA8, 256, _abi.TONE_SRGB),
        (2, 2, _abi.RAW_U16, 12, _abi.BGGR, 256.0, _abi.MALVAR, 0.5, _abi.WB_GRAY_WORLD, _abi.OUT_RGB8, 256, _abi.TONE_SRGB),
        (450, 260, _abi.RAW_U16, 12, _abi.RGGB, 256.0, _abi.MALVAR, 0.5, _abi.WB_GRAY_WORLD, _abi.OUT_RGBA8, 256, _abi.TONE_SRGB),
    ]
    for i, (w, h, fmt, bits, pattern, black, dem, sharpen, wb, out, bins, tone) in enumerate(rows):
        sc = cases.baseline_scene(w, h, fmt, bits, pattern, black, 900 + i)
        p = isp.IspParams(black_level=black, lens_shading=isp.scene_lens_shading(sc, 9, 7) if i % 3 != 2 else None,
                          white_balance=wb, wb_gains=(1.8, 1.0, 1.5), demosaic=dem, sharpen=sharpen,
                          ccm=cases.stream_ccm(i), tone_curve=tone, histogram_bins=bins)
        specs.append((sc, p, out))
    return specs


class Slot:
    """One stream's device buffers: raw in, image, histogram and statistics out."""

    def __init__(self, scene, params, out_format):
        import torch
        self.scene, self.params, self.out_format = scene, params, out_format
        self.frame = isp.render_scene(scene)
        self.d_raw = torch.from_numpy(self.frame.data).to("cuda")
        self.pitch = scene.width * output_channels(out_format) * output_channel_bytes(out_format)
        bins = params.histogram_bins or 256
        self.out = torch.zeros(self.pitch * scene.height, dtype=torch.uint8, device="cuda")
        self.hist = torch.zeros(4 * bins, dtype=torch.int32, device="cuda")
        self.stats = torch.zeros(12, dtype=torch.float32, device="cuda")

    def load(self, tick):
        """Renders the stream's frame for `tick` (seed + 1000 * tick) into the raw buffer."""
        import torch
        sc = isp.scene_config(**{k: getattr(self.scene, k) for k, _ in _abi.SceneConfig._fields_
                                 if k != "illumination"})
        sc.illumination[:] = list(self.scene.illumination)
        sc.seed = self.scene.seed + 1000 * tick
        self.frame = isp.render_scene(sc)
        self.d_raw.copy_(torch.from_numpy(self.frame.data))

    def entry(self, pipe, exposure=EXPOSURE):
        return (pipe, self.frame.view(self.d_raw), self.out, self.pitch, self.out_format, self.hist,
                self.stats, exposure)

    def result(self):
        return gs.Result(self.out.cpu().numpy(), self.hist.cpu().numpy().view(np.uint32), self.stats.cpu().numpy())


def compare(name, g, h):
    assert gs.mismatches(g, h) == 0, f"{name}: {gs.mismatches(g, h)} bytes differ"
    assert np.array_equal(g.histogram, h.histogram), f"{name}: histograms differ"
    assert gs.same_stats(g, h), f"{name}: stats {g.stats} vs {h.stats}"


def run_batch(specs, schedule=None, ticks=1, between=None, feedback=None):
    """Runs `ticks` launches of one batch over fresh pipelines; returns [tick][stream] results."""
    import torch
    slots = [Slot(*s) for s in specs]
    pipes = [isp.IspPipeline(s.params) for s in slots]
    if feedback is not None:
        for p in pipes:
            p.set_exposure_feedback(feedback)
    batch = isp.IspBatch(schedule)
    stream = torch.cuda.current_stream().cuda_stream
    results = []
    for t in range(ticks):
        for s in slots:
            s.load(t)
        if t == 0:
            batch.set_frames([s.entry(p) for s, p in zip(slots, pipes)])
        batch.launch(stream)
        if between is not None:
            between()
        torch.cuda.synchronize()
        results.append([s.result() for s in slots])
    batch.synchronize()
    return results


def run_single(specs, ticks=1, feedback=None):
    """The same streams, one isp_process per frame on each stream's own pipeline."""
    slots = [Slot(*s) for s in specs]
    pipes = [isp.IspPipeline(s.params) for s in slots]
    if feedback is not None:
        for p in pipes:
            p.set_exposure_feedback(feedback)
    results = []
    for t in range(ticks):
        row = []
        for s, p in zip(slots, pipes):
            s.load(t)
            row.append(gs.run(p, s.frame, s.out_format, True, EXPOSURE, d_raw=s.d_raw))
        results.append(row)
    return results


def run_oracle(specs, ticks=1):
    slots = [Slot(*s) for s in specs]
    orcs = [OraclePipeline(s.params) for s in slots]
    results = []
    for t in range(ticks):
        row = []
        for s, o in zip(slots, orcs):
            s.load(t)
            row.append(o.process(s.frame, s.out_format, True, EXPOSURE, threads=8))
        results.append(row)
    return results


def test_batch_matches_per_frame_and_oracle(torch_gpu):
    specs = mixed_specs()
    batched = run_batch(specs, ticks=3)
    single = run_single(specs, ticks=3)
    oracle = run_oracle(specs, ticks=3)
    for t in range(3):
        for i in range(len(specs)):
            compare(f"tick {t} stream {i} vs per-frame", batched[t][i], single[t][i])
            compare(f"tick {t} stream {i} vs oracle", batched[t][i], oracle[t][i])


SCHEDULES = [
    dict(graph=0),
    dict(queue=_abi.BATCH_QUEUE_STATIC),
    dict(queue=_abi.BATCH_QUEUE_STATIC_XCD, tiles_per_item=1, tile16=True),
    dict(tile16=True),
    dict(tiles_per_item=1, awb_tiles_per_item=1),
    dict(tiles_per_item=32, awb_tiles_per_item=64, order=_abi.BATCH_ORDER_AS_GIVEN),
    dict(blocks_per_cu=1, tile16=True, graph=0),
    dict(per_frame=1),
    dict(per_frame=1, graph=0),
]


def make_schedule(tile16=False, tiles_per_item=None, **kw):
    s = isp.default_schedule()
    for f in range(4):
        for d in range(2):
            if tile16:
                s.tile_height[f][d] = 16
            if tiles_per_item is not None:
                s.tiles_per_item[f][d] = tiles_per_item
    for k, v in kw.items():
        setattr(s, k, v)
    return s


@pytest.mark.parametrize("options", SCHEDULES, ids=lambda o: ",".join(f"{k}={v}" for k, v in o.items()))
def test_every_schedule_gives_the_same_bits(torch_gpu, options):
    specs = mixed_specs()
    got = run_batch(specs, make_schedule(**options), ticks=2)
    want = run_single(specs, ticks=2)
    for t in range(2):
        for i in range(len(specs)):
            compare(f"tick {t} stream {i}", got[t][i], want[t][i])


def one_variant_specs(fmt, bits, dem, sharpen):
    """Frames that all run one tile body (raw format, demosaic, sharpening, 32-row tiles), at
    several sizes, patterns and settings: the batch then runs the kernel of that body alone."""
    rows = [(320, 180, _abi.RGGB, _abi.WB_GRAY_WORLD, _abi.OUT_RGBA8), (130, 66, _abi.GBRG, _abi.WB_WHITE_PATCH, _abi.OUT_RGB8),
            (67, 35, _abi.GRBG, _abi.WB_MANUAL, _abi.OUT_RGB16), (258, 100, _abi.BGGR, _abi.WB_GRAY_WORLD, _abi.OUT_RGB32F)]
    specs = []
    for i, (w, h, pattern, wb, out) in enumerate(rows):
        black = 16.0 if bits == 8 else 64.0
        sc = cases.baseline_scene(w, h, fmt, bits, pattern, black, 700 + i)
        p = isp.IspParams(black_level=black, lens_shading=isp.scene_lens_shading(sc, 9, 7) if i % 2 == 0 else None,
                          white_balance=wb, wb_gains=(1.8, 1.0, 1.5), demosaic=dem, sharpen=sharpen,
                          ccm=cases.stream_ccm(i), histogram_bins=1024 if i == 2 else 256)
        specs.append((sc, p, out))
    return specs


@pytest.mark.parametrize("fmt,bits,dem,sharpen", [(_abi.RAW_U16, 12, _abi.MALVAR, 0.5), (_abi.RAW_RAW12, 12, _abi.MALVAR, 0.5),
                                                  (_abi.RAW_RAW10, 10, _abi.BILINEAR, 0.0), (_abi.RAW_U8, 8, _abi.BILINEAR, 0.5),
                                                  (_abi.RAW_U16, 14, _abi.BILINEAR, 0.0)],
                         ids=["u16-malvar-sharpen", "raw12-malvar-sharpen", "raw10-bilinear", "u8-bilinear-sharpen",
                              "u16-bilinear"])
def test_one_variant_batch_matches_per_frame_and_oracle(torch_gpu, fmt, bits, dem, sharpen):
    """A batch of one tile body runs that body's own kernel; results equal the per-frame runs and
    the oracle over two ticks, eager and replayed."""
    specs = one_variant_specs(fmt, bits, dem, sharpen)
    single = run_single(specs, ticks=2)
    oracle = run_oracle(specs, ticks=2)
    for graph in (0, 1):
        batched = run_batch(specs, make_schedule(graph=graph), ticks=2)
        for t in range(2):
            for i in range(len(specs)):
                compare(f"graph {graph} tick {t} stream {i} vs per-frame", batched[t][i], single[t][i])
                compare(f"graph {graph} tick {t} stream {i} vs oracle", batched[t][i], oracle[t][i])


def test_graph_replays_with_other_work_between(torch_gpu):
    """Replays of the captured launch, with eager GPU work and host synchronisations between
    them, each equal to the eager (uncaptured) launch's results."""
    import torch
    specs = mixed_specs()[:5]
    scratch = torch.ones(1 << 20, device="cuda")

    def between():
        scratch.mul_(1.0001).add_(0.5)
        torch.cuda.synchronize()
        scratch.sum().item()

    replayed = run_batch(specs, make_schedule(graph=1), ticks=4, between=between)
    eager = run_batch(specs, make_schedule(graph=0), ticks=4)
    for t in range(4):
        for i in range(len(specs)):
            compare(f"replay {t} stream {i}", replayed[t][i], eager[t][i])


def test_device_exposure_feedback_alone_and_batched(torch_gpu):
    """Each stream's exposure state lives on the device: a frame runs at its predecessor's
    suggestion. Batched, per-frame, and a host loop that feeds the statistics back by hand give
    the same bits."""
    specs = mixed_specs()[:4]
    start = (8_000_000, 1.5)
    batched = run_batch(specs, ticks=3, feedback=start)
    single = run_single(specs, ticks=3, feedback=start)
    # The host loop: exposure_ns = round(stats[6]), gain = stats[7].
    slots = [Slot(*s) for s in specs]
    for i, s in enumerate(slots):
        pipe = isp.IspPipeline(s.params)
        exposure = start
        for t in range(3):
            s.load(t)
            r = gs.run(pipe, s.frame, s.out_format, True, exposure, d_raw=s.d_raw)
            compare(f"tick {t} stream {i} batched", batched[t][i], r)
            compare(f"tick {t} stream {i} per-frame", single[t][i], r)
            e = float(r.stats[_abi.STAT_SUGGESTED_EXPOSURE_NS])
            exposure = (int(np.floor(np.float64(e) + 0.5)), float(r.stats[_abi.STAT_SUGGESTED_GAIN]))


def test_large_lens_shading_grids_batched(torch_gpu):
    """The large lens shading grids (device-memory reads, tile tables and per-sample gains) in one
    batch: the batch kernels against the oracle over two ticks."""
    specs = [cases.large_grid(row, 700 + i) for i, row in enumerate(cases.LARGE_GRIDS)]
    batched = run_batch(specs, ticks=2)
    oracle = run_oracle(specs, ticks=2)
    for t in range(2):
        for i in range(len(specs)):
            compare(f"tick {t} grid {cases.LARGE_GRIDS[i][:3]}", batched[t][i], oracle[t][i])


def test_feedback_switched_after_the_frames_are_bound(torch_gpu):
    """Exposure feedback switched on, off and on again between isp_batch_set_frames and the
    launches: each launch runs with the switch as it stands then (the bound table keeps the
    pipelines' exposure state, which lives as long as the pipeline), equal to per-frame processing
    with the same switches."""
    import torch
    specs = mixed_specs()[:4]
    slots = [Slot(*s) for s in specs]
    pipes = [isp.IspPipeline(s.params) for s in slots]
    singles = [isp.IspPipeline(s.params) for s in slots]
    batch = isp.IspBatch(make_schedule(graph=0))
    stream = torch.cuda.current_stream().cuda_stream
    batch.set_frames([s.entry(p) for s, p in zip(slots, pipes)])
    switches = [(8_000_000, 1.5), "keep", None, (20_000_000, 3.0), "keep", None]
    for t, switch in enumerate(switches):
        if switch != "keep":
            batch.synchronize()
            for p in pipes + singles:
                p.set_exposure_feedback(switch)
        for s in slots:
            s.load(t)
        batch.launch(stream)
        torch.cuda.synchronize()
        got = [s.result() for s in slots]
        for i, (s, p) in enumerate(zip(slots, singles)):
            want = gs.run(p, s.frame, s.out_format, True, EXPOSURE, d_raw=s.d_raw)
            compare(f"tick {t} stream {i}", got[i], want)
        if switch is None:
            with pytest.raises(ValueError, match="not enabled"):
                pipes[0].exposure_state()
        else:
            assert pipes[0].exposure_state() == singles[0].exposure_state()


def zero_sign_spec(seed):
    """A stream whose tone table starts at -0.0, with the signed zeros it can carry: many samples
    below black, a colour matrix row of negative weights, RGB32F output."""
    sc = cases.baseline_scene(96, 70, _abi.RAW_U16, 12, _abi.GRBG, 600.0, seed)
    lut = np.linspace(0.0, 1.0, 257, dtype=np.float32)
    lut[0] = np.float32(-0.0)
    p = isp.IspParams(black_level=600.0, wb_gains=(1.0, 1.0, 1.0), demosaic=_abi.MALVAR, sharpen=1.5,
                      ccm=(1.0, 0.0, 0.0, 0.0, 1.0, 0.0, -0.25, -0.5, -0.25), tone_curve=_abi.TONE_LUT,
                      lut=[float(v) for v in lut], white_balance=_abi.WB_GRAY_WORLD)
    return sc, p, _abi.OUT_RGB32F


def test_zero_sign_frames_in_a_batch(torch_gpu):
    """Streams whose tone table starts at -0.0 run the per-frame kernels inside the batch's launch
    sequence: the batch equals per-frame processing and the oracle, bit for bit (signed zeros of
    the RGB32F output included), on eager and captured launches."""
    specs = mixed_specs()[:3] + [zero_sign_spec(41), zero_sign_spec(42)]
    oracle = run_oracle(specs, ticks=2)
    single = run_single(specs, ticks=2)
    for graph in (0, 1):
        batched = run_batch(specs, make_schedule(graph=graph), ticks=2)
        for t in range(2):
            for i in range(len(specs)):
                compare(f"graph {graph} tick {t} stream {i} vs oracle", batched[t][i], oracle[t][i])
                compare(f"graph {graph} tick {t} stream {i} vs per-frame", batched[t][i], single[t][i])
    negative = [np.count_nonzero(r.image.view(np.uint32) == 0x80000000) for r in oracle[0][3:]]
    assert min(negative) > 0  # the outputs hold -0.0


def test_autotuner_cache_and_state(torch_gpu, tmp_path):
    """The tuner returns a valid schedule, stores it, reuses it without timing, and leaves the
    streams' white balance state as it found it."""
    import torch
    specs = mixed_specs()[:6]
    slots = [Slot(*s) for s in specs]
    pipes = [isp.IspPipeline(s.params) for s in slots]
    batch = isp.IspBatch()
    stream = torch.cuda.current_stream().cuda_stream
    batch.set_frames([s.entry(p) for s, p in zip(slots, pipes)])
    cache = str(tmp_path / "tune.txt")
    chosen, ms = batch.tune(stream, cache, launches=2)
    assert ms > 0.0
    assert all(h in (16, 32) for row in chosen.tile_height for h in row)
    assert all(1 <= n <= 32 for row in chosen.tiles_per_item for n in row)
    lines = open(cache).read().splitlines()
    assert len(lines) == 1 and "\t" in lines[0]
    again, ms2 = batch.tune(stream, cache, launches=2)
    assert ms2 == 0.0
    assert bytes(again) == bytes(chosen)
    # After tuning, the first launch equals a fresh stream's first frame.
    batch.launch(stream)
    torch.cuda.synchronize()
    want = run_single(specs, ticks=1)[0]
    for i, s in enumerate(slots):
        compare(f"stream {i} after tuning", s.result(), want[i])


def test_config3_batch_reduced(torch_gpu):
    """The 16-stream mixed workload at 1/8 size against per-frame runs and the oracle."""
    specs = cases.config3(divisor=8)
    batched = run_batch(specs, ticks=2)
    single = run_single(specs, ticks=2)
    oracle = run_oracle(specs, ticks=2)
    for t in range(2):
        for i in range(len(specs)):
            compare(f"tick {t} stream {i}", batched[t][i], single[t][i])
            compare(f"tick {t} stream {i} oracle", batched[t][i], oracle[t][i])


def test_config3_batch_full_size_matches_per_frame(torch_gpu):
    """The 16-stream mixed workload at full size (55.3 Mpx a tick) equals per-frame processing."""
    specs = cases.config3()
    batched = run_batch(specs, ticks=1)
    single = run_single(specs, ticks=1)
    for i in range(len(specs)):
        compare(f"stream {i}", batched[0][i], single[0][i])


def test_batch_validation(torch_gpu):
    specs = mixed_specs()[:2]
    slots = [Slot(*s) for s in specs]
    pipe = isp.IspPipeline(slots[0].params)
    batch = isp.IspBatch()
    with pytest.raises(ValueError, match="already has a frame"):
        batch.set_frames([slots[0].entry(pipe), slots[0].entry(pipe)])
    with pytest.raises(ValueError, match="count"):
        batch.set_frames([])
    with pytest.raises(ValueError, match="no frames"):
        batch.launch()
    bad = make_schedule(tiles_per_item=0)  # every format
    with pytest.raises(ValueError, match="tiles_per_item"):
        batch.set_schedule(bad)
    bad = make_schedule()
    bad.tile_height[1][0] = 24
    with pytest.raises(ValueError, match="tile_height"):
        batch.set_schedule(bad)
    short = slots[1].entry(isp.IspPipeline(slots[1].params))
    short = short[:3] + (16,) + short[4:]  # out pitch shorter than a row
    with pytest.raises(ValueError, match="batch frame 0"):
        batch.set_frames([short])
