"""The served form (isp_server_*): frames of many cameras submitted as they
arrive, admitted against an HBM budget and run in continuous batches by the
native scheduling thread. Each camera's results (images, histograms,
statistics) must be the CPU oracle's bit for bit, with the device exposure
state feeding each frame from its predecessor's suggestion as the oracle chain
does it by hand (as the reference checks each frame against its CPU backend,
modules/isp/tests/pipeline_test.cpp:164-235), alone and in the mix, whatever
the budget, the launch slots, the batch size and whether frames come from host
or device memory; and at the served workload's own size: 32 cameras of 720p to
4K RAW10 / RAW12 with white patch balance."""
import random

import numpy as np
import pytest

import holoscan_isp as isp
from holoscan_isp import _abi
from holoscan_isp.params import output_channel_bytes, output_channels

import cases
import gpu_support as gs
from oracle_lib import OraclePipeline

pytestmark = pytest.mark.gpu

INITIAL = (10_000_000, 2.0)


def cameras():
    """(scene, params, out_format) per camera: RAW10 / RAW12 at several sizes, white patch at the
    0.99 percentile, Malvar or bilinear, some sharpened."""
    rows = [(320, 180, _abi.RAW_RAW10, _abi.BGGR, _abi.MALVAR, 0.5),
            (256, 144, _abi.RAW_RAW12, _abi.RGGB, _abi.BILINEAR, 0.0),
            (384, 240, _abi.RAW_RAW12, _abi.GRBG, _abi.MALVAR, 0.0),
            (130, 66, _abi.RAW_RAW10, _abi.GBRG, _abi.BILINEAR, 1.0),
            (512, 288, _abi.RAW_RAW12, _abi.GBRG, _abi.MALVAR, 0.5),
            (66, 34, _abi.RAW_RAW12, _abi.RGGB, _abi.MALVAR, 0.0)]
    out = []
    for i, (w, h, fmt, pattern, dem, sharpen) in enumerate(rows):
        bits = 10 if fmt == _abi.RAW_RAW10 else 12
        black = 64.0 if bits == 10 else 240.0
        sc = cases.baseline_scene(w, h, fmt, bits, pattern, black, 4000 + i)
        p = isp.IspParams(black_level=black, lens_shading=isp.scene_lens_shading(sc, 17, 13),
                          white_balance=_abi.WB_WHITE_PATCH, white_patch_percentile=0.99, demosaic=dem,
                          sharpen=sharpen, ccm=cases.stream_ccm(i), tone_curve=_abi.TONE_SRGB)
        out.append((sc, p, _abi.OUT_RGBA8 if i % 2 else _abi.OUT_RGB8))
    return out


def frames_of(scene, n):
    """n frames of a camera: its scene with seeds seed, seed + 1000, ..."""
    frames = []
    for k in range(n):
        sc = isp.scene_config(**{f: getattr(scene, f) for f, _ in _abi.SceneConfig._fields_ if f != "illumination"})
        sc.illumination[:] = list(scene.illumination)
        sc.seed = scene.seed + 1000 * k
        frames.append(isp.render_scene(sc))
    return frames


def pinned(frame):
    import torch
    t = torch.empty(frame.data.size, dtype=torch.uint8, pin_memory=True)
    t.numpy()[:] = frame.data
    return t


class Got:
    def __init__(self, image, histogram, stats):
        self.image, self.histogram, self.stats = image, histogram, stats


def collect(server, n, specs):
    """Polls n results; returns {(camera, tag): Got} with the image rows unpadded."""
    got = {}
    while len(got) < n:
        r = server.poll(2_000_000)
        assert r is not None, "no result within 2 s"
        assert r.status == _abi.OK, isp.last_error()
        sc, p, of = specs[r.camera]
        row = sc.width * output_channels(of) * output_channel_bytes(of)
        img = gs.device_bytes(r.out, r.out_pitch * sc.height).reshape(sc.height, r.out_pitch)[:, :row]
        hist = gs.device_bytes(r.histogram, 4 * 256 * 4).view(np.uint32)
        stats = gs.device_bytes(r.stats, 48).view(np.float32)
        got[(r.camera, r.tag)] = Got(np.ascontiguousarray(img).reshape(-1), hist, stats)
        server.release(r.request)
    return got


def serve(specs, n, budget=256 << 20, max_batch=32, slots=2, per_frame=0, device_source=False, seed=7):
    """Submits n frames of every camera, interleaved at random (each camera's in order)."""
    import torch
    server = isp.IspServer(hbm_budget=budget, max_batch=max_batch, launch_slots=slots, per_frame=per_frame)
    ids = [server.add_camera(p, sc.width, sc.height, sc.format, sc.bit_depth, sc.pattern, of, True, INITIAL)
           for sc, p, of in specs]
    bufs = []
    for cam, (sc, p, of) in zip(ids, specs):
        fr = frames_of(sc, n)
        bufs.append([(f, torch.from_numpy(f.data).to("cuda") if device_source else pinned(f)) for f in fr])
    torch.cuda.synchronize()
    order = [c for c in range(len(specs)) for _ in range(n)]
    random.Random(seed).shuffle(order)
    nxt = [0] * len(specs)
    for c in order:
        f, buf = bufs[c][nxt[c]]
        server.submit(ids[c], buf, f.pitch, _abi.LOCATION_DEVICE if device_source else _abi.LOCATION_HOST, tag=nxt[c])
        nxt[c] += 1
    got = collect(server, n * len(specs), specs)
    counters = server.counters()
    server.close()
    return got, counters


def alone(specs, n):
    """Each camera by itself: its own pipeline, exposure feedback on, one isp_process per frame."""
    want = {}
    for c, (sc, p, of) in enumerate(specs):
        pipe = isp.IspPipeline(p)
        pipe.set_exposure_feedback(INITIAL)
        for k, f in enumerate(frames_of(sc, n)):
            r = gs.run(pipe, f, of, True, (0, 0.0))
            want[(c, k)] = r
    return want


def oracle_chain(specs, n, threads=16):
    """Each camera through the CPU oracle, frame after frame, its exposure carried by hand as the
    device state carries it: the first frame at INITIAL, each later one at its predecessor's
    suggestion, exposure_ns = (long long)(double(stats[6]) + 0.5) and gain = stats[7], whenever
    stats[6] > 0."""
    want = {}
    for c, (sc, p, of) in enumerate(specs):
        cpu = OraclePipeline(p)
        exposure = INITIAL
        for k, f in enumerate(frames_of(sc, n)):
            h = cpu.process(f, of, True, exposure, threads=threads)
            want[(c, k)] = h
            suggested = float(np.float64(h.stats[_abi.STAT_SUGGESTED_EXPOSURE_NS]))
            if suggested > 0.0:
                exposure = (int(suggested + 0.5), float(h.stats[_abi.STAT_SUGGESTED_GAIN]))
    return want


def compare(got, want):
    assert got.keys() == want.keys()
    for key in sorted(want):
        g, w = got[key], want[key]
        assert np.array_equal(g.image, w.image.reshape(-1)), f"camera {key[0]} frame {key[1]}: image differs"
        assert np.array_equal(g.histogram, w.histogram), f"camera {key[0]} frame {key[1]}: histogram differs"
        assert np.array_equal(g.stats.view(np.uint32), w.stats.view(np.uint32)), \
            f"camera {key[0]} frame {key[1]}: stats {g.stats} vs {w.stats}"


def test_same_alone_and_in_the_mix(torch_gpu):
    specs = cameras()
    got, counters = serve(specs, 3)
    compare(got, oracle_chain(specs, 3))
    compare(got, alone(specs, 3))
    assert counters.completed == 3 * len(specs)
    assert counters.budget_used == 0


def test_served_workload_full_size_matches_the_oracle(torch_gpu):
    """The 32-camera served mix at its own sizes (720p, 1080p, 1920x1200 and 4K; RAW10 and RAW12;
    white patch at the 0.99 percentile; device exposure feedback), three frames a camera
    submitted at once in a shuffled order, so that the loop admits them in batches under a 1 GiB
    budget: every image, histogram and statistics vector equals the oracle's, frame by frame."""
    specs = cases.config4()
    got, counters = serve(specs, 3, budget=1 << 30)
    assert counters.completed == 3 * len(specs)
    assert counters.budget_peak <= 1 << 30
    assert counters.launches < counters.completed  # frames shared launches
    for c in range(len(specs)):
        want = oracle_chain(specs[c:c + 1], 3)
        compare({(0, k): got.pop((c, k)) for k in range(3)}, want)


def test_tight_budget_admits_in_turn(torch_gpu):
    """A budget that holds about two frames: requests wait for memory, results are unchanged,
    and admitted memory never exceeds the budget."""
    specs = cameras()
    budget = 2 * (512 * 288 * 4 + 512 * 288 * 3 // 2) + (1 << 16)
    got, counters = serve(specs, 2, budget=budget)
    compare(got, oracle_chain(specs, 2))
    assert counters.budget_peak <= budget
    assert counters.launches >= 4


@pytest.mark.parametrize("options", [dict(per_frame=1), dict(slots=1, max_batch=2), dict(slots=4, max_batch=3),
                                     dict(device_source=True)],
                         ids=["per-frame", "1 slot batch 2", "4 slots batch 3", "device source"])
def test_modes_give_the_same_bits(torch_gpu, options):
    specs = cameras()[:4]
    got, _ = serve(specs, 2, **options)
    compare(got, oracle_chain(specs, 2))


def test_admission_errors(torch_gpu):
    specs = cameras()
    server = isp.IspServer(hbm_budget=1 << 20)
    sc, p, of = specs[4]
    with pytest.raises(MemoryError, match="budget"):
        server.add_camera(p, 1920, 1080, _abi.RAW_RAW12, 12, _abi.RGGB, _abi.OUT_RGBA8)
    cam = server.add_camera(p, sc.width, sc.height, sc.format, sc.bit_depth, sc.pattern, of)
    with pytest.raises(ValueError, match="camera"):
        server.submit(cam + 5, np.zeros(16, np.uint8), 16)
    with pytest.raises(ValueError, match="pitch"):
        server.submit(cam, np.zeros(16, np.uint8), 3)
    assert server.poll(0) is None
    server.close()
