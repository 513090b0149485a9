"""Benchmark: the 4K RAW12 Malvar ISP pipeline on one MI355X.

Workload (one "step" = one frame): 3840x2160 GRBG MIPI RAW12 (packed, 5760 B rows)
seeded noisy scene, black level 240, 17x13 lens shading, gray-world white
balance estimated on the frame, Malvar-He-Cutler demosaic, unsharp mask 0.5,
fixed CCM, sRGB tone table, RGBA8 output, 4x256 histograms, statistics and an
auto-exposure suggestion. Inputs are resident in HBM before the timed region:
eight distinct frames and eight output buffers (365 MB, more than the 256 MB
last-level cache) taken in rotation, so no frame is served from that cache.

Prints one JSON line: Gpix/s (value), HBM GB/s at the compulsory 5.5 B/px, per-
frame p50/p99 GPU time, a bit-exact check of the timed sequence's last frame
against the CPU oracle (which replays the white balance state before it), the
dominant kernel's roofline (HIP events on the launch stream), and the reference
CPU backend timed on this host's cores on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "holoscan-zozo-modules_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

HBM_PEAK_GBS = 8000.0
RAW_BYTES_PER_PX = 1.5  # RAW12
OUT_BYTES_PER_PX = 4.0  # RGBA8


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cpu-frames", type=int, default=6, help="frames of the single-thread CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--batch-ticks", type=int, default=100, help="timed ticks of the 16-stream batch (0: skip)")
    ap.add_argument("--large-ticks", type=int, default=3, help="timed batches of 8 x 8192x6144 frames (0: skip)")
    ap.add_argument("--served-seconds", type=float, default=2.0,
                    help="real-time length of the 32-camera serving trace (0: skip)")
    ap.add_argument("--source-launches", type=int, default=50,
                    help="timed launches of the device raw source and the frame checker (0: skip)")
    ap.add_argument("--color-launches", type=int, default=100,
                    help="timed launches of each 4K colour conversion after the ISP (0: skip)")
    ap.add_argument("--config1-frames", type=int, default=500,
                    help="timed frames of the 1080p u16 workload (0: skip)")
    ap.add_argument("--reference-iterations", type=int, default=100,
                    help="timed frames of each of the reference benchmark's eight cases (0: skip)")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="only run a plain device copy of this many MiB (10 times): the reference "
                         "point that scales the FETCH_SIZE / WRITE_SIZE counters to bytes")
    return ap.parse_args()


def copy_calibration(mib):
    """A plain device-to-device copy of known size; under rocprofv3 --pmc its counters
    per dispatch give the KiB the counter reports for mib MiB read and mib MiB written."""
    import torch
    torch.cuda.set_device(0)
    n = mib << 20
    src = torch.ones(n, dtype=torch.uint8, device="cuda")
    dst = torch.empty(n, dtype=torch.uint8, device="cuda")
    for _ in range(10):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": n, "launches": 10}))


def traffic_from_counters():
    """HBM bytes per main-kernel launch from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    readings, each scaled by what the same counter reads for a 256 MiB device copy. Also the
    fetch scaled by the white-balance kernel, which reads the packed frame exactly once with the
    same vector loads."""
    path = os.path.join(ROOT, "profiles", "main_kernel_4k_raw12_malvar_counters.json")
    if not os.path.exists(path):
        return None, None
    with open(path) as fh:
        c = json.load(fh)
    fetch_factor = c["copy_bytes"] / 1024.0 / c["copy_fetch_kib"]
    write_factor = c["copy_bytes"] / 1024.0 / c["copy_write_kib"]
    copy_scaled = (c["fetch_kib_per_dispatch"] * fetch_factor + c["write_kib_per_dispatch"] * write_factor) * 1024.0
    awb_factor = c["awb_algorithmic_bytes"] / 1024.0 / c["awb_fetch_kib_per_dispatch"]
    awb_scaled = (c["fetch_kib_per_dispatch"] * awb_factor + c["write_kib_per_dispatch"] * write_factor) * 1024.0
    detail = {"fetch_factor_copy": round(fetch_factor, 4), "write_factor_copy": round(write_factor, 4),
              "fetch_factor_single_read_kernel": round(awb_factor, 4),
              "traffic_single_read_scaled": round(awb_scaled), "source": os.path.relpath(path, ROOT)}
    return round(copy_scaled), detail


VALU_PEAK_TLANE_OPS = 256 * 64 * 2.4e9 / 1e12  # 256 CUs x 64 lanes x 2.4 GHz, unpacked fp32


def valu_from_counters(main_ms, pixels):
    """The main kernel against its arithmetic ceiling: lane-operations per launch from the
    committed rocprofv3 SQ_INSTS_VALU reading (wave instructions x 64 lanes) over the live
    average launch time, against 256 CUs x 64 lanes x 2.4 GHz; and the VALU-busy share of the
    profiled run (SQ_ACTIVE_INST_VALU x 4 cycles over 4 SIMDs x 256 CUs x the kernel's cycles
    per XCD)."""
    path = os.path.join(ROOT, "profiles", "main_kernel_4k_raw12_malvar_counters.json")
    if not os.path.exists(path) or main_ms <= 0:
        return None
    with open(path) as fh:
        c = json.load(fh)
    lane_ops = c["sq_insts_valu_per_dispatch"] * 64.0
    achieved = lane_ops / (main_ms * 1e-3) / 1e12
    busy = c["sq_active_inst_valu_per_dispatch"] * 4.0 / (
        4 * 256 * c["grbm_gui_active_cycles_per_dispatch"] / c["xcds"])
    return {"bound": "valu", "achieved": round(achieved, 2), "peak": round(VALU_PEAK_TLANE_OPS, 2),
            "unit": "T lane-ops/s", "frac": round(achieved / VALU_PEAK_TLANE_OPS, 4),
            "lane_ops_per_px": round(lane_ops / pixels, 1), "busy_profiled": round(busy, 3),
            "source": os.path.relpath(path, ROOT)}


ROTATION = 8  # distinct 4K frames and output buffers the timed sequence cycles through


def render_frames(isp, scene, count):
    """`count` frames of one scene, seeds seed, seed + 1000, ...: rendered on host threads."""
    from concurrent.futures import ThreadPoolExecutor
    from holoscan_isp import _abi
    configs = []
    for k in range(count):
        c = isp.scene_config(**{f: getattr(scene, f) for f, _ in _abi.SceneConfig._fields_ if f != "illumination"})
        c.illumination[:] = list(scene.illumination)
        c.seed = scene.seed + 1000 * k
        configs.append(c)
    with ThreadPoolExecutor(max_workers=min(count, os.cpu_count() or 1)) as pool:
        return list(pool.map(isp.render_scene, configs))


def percentile(values, q):
    s = sorted(values)
    if not s:
        return None
    k = min(len(s) - 1, max(0, int(round(q / 100.0 * (len(s) - 1)))))
    return s[k]


def cpu_model():
    """The host CPU's model name, as /proc/cpuinfo gives it."""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(frame, params, out_format, frames):
    """The reference CPU backend (compiled from its own sources) when present, else the
    oracle; 1 thread, one warm-up call then the mean of `frames` calls
    (modules/isp/tests/benchmark.cpp:147-165). Also the oracle on all cores."""
    import oracle_lib
    pixels = frame.width * frame.height
    out = {}
    if oracle_lib.ref_available():
        pipe, kind = oracle_lib.RefPipeline(params), "reference"
        run = lambda: pipe.process(frame, out_format, True, (10_000_000, 2.0))  # noqa: E731
    else:
        pipe, kind = oracle_lib.OraclePipeline(params), "port"
        run = lambda: pipe.process(frame, out_format, True, (10_000_000, 2.0), threads=1)  # noqa: E731
    run()
    t0 = time.perf_counter()
    for _ in range(frames):
        run()
    dt = (time.perf_counter() - t0) / frames
    out["single"] = dict(value=pixels / dt / 1e9, unit="Gpix/s", cores=1, kind=kind,
                         sample=f"{frames} frames of the 3840x2160 RAW12 workload after 1 warm-up, "
                                f"{dt * 1e3:.1f} ms/frame", cpu=cpu_model(), nproc=os.cpu_count())
    threads = min(16, os.cpu_count() or 1)
    orc = oracle_lib.OraclePipeline(params)
    orc.process(frame, out_format, True, (10_000_000, 2.0), threads=threads)
    n = max(2, frames)
    t0 = time.perf_counter()
    for _ in range(n):
        orc.process(frame, out_format, True, (10_000_000, 2.0), threads=threads)
    dt2 = (time.perf_counter() - t0) / n
    out["all_cores"] = dict(value=pixels / dt2 / 1e9, unit="Gpix/s", cores=threads, kind="port",
                            ms_per_frame=dt2 * 1e3)
    return out


class StreamBuffers:
    """One stream's frame resident in HBM and its output buffers."""

    def __init__(self, isp, torch, scene, params, out_format):
        from holoscan_isp.params import output_channel_bytes, output_channels
        self.scene, self.params, self.out_format = scene, params, out_format
        self.frame = isp.render_scene(scene)
        self.d_raw = torch.from_numpy(self.frame.data).to("cuda")
        self.view = self.frame.view(self.d_raw)
        self.pitch = scene.width * output_channels(out_format) * output_channel_bytes(out_format)
        self.out = torch.empty(self.pitch * scene.height, dtype=torch.uint8, device="cuda")
        self.hist = torch.empty(4 * 256, dtype=torch.int32, device="cuda")
        self.stats = torch.empty(12, dtype=torch.float32, device="cuda")


def batched_vs_per_frame(ticks):
    """The 16-stream mixed tick (4 x 720p RAW10, 8 x 1920x1200 RAW12, 4 x 4K u16 14-bit): 16
    per-frame launches against one autotuned persistent launch replayed as a hipGraph. Per-tick
    GPU time from HIP events on the launch stream."""
    import torch
    import holoscan_isp as isp
    import cases
    specs = cases.config3()
    bufs = [StreamBuffers(isp, torch, *s) for s in specs]
    pixels = sum(b.scene.width * b.scene.height for b in bufs)
    stream = torch.cuda.current_stream()
    sh = stream.cuda_stream
    exposure = (10_000_000, 2.0)

    def timed(step):
        for _ in range(3):
            step()
        torch.cuda.synchronize()
        evs = []
        for _ in range(ticks):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            step()
            e1.record(stream)
            evs.append((e0, e1))
        torch.cuda.synchronize()
        ms = [a.elapsed_time(b) for a, b in evs]
        mean = sum(ms) / len(ms)
        return {"ms_per_tick": round(mean, 4), "gpix_s": round(pixels / mean / 1e6, 3),
                "tick_ms_p50": round(percentile(ms, 50), 4), "tick_ms_p99": round(percentile(ms, 99), 4)}

    pipes = [isp.IspPipeline(b.params) for b in bufs]

    def per_frame():
        for b, p in zip(bufs, pipes):
            p.process(b.view, b.out, b.pitch, b.out_format, b.hist, b.stats, exposure, sh)

    single = timed(per_frame)
    bpipes = [isp.IspPipeline(b.params) for b in bufs]
    batch = isp.IspBatch()
    batch.set_frames([(p, b.view, b.out, b.pitch, b.out_format, b.hist, b.stats, exposure)
                      for b, p in zip(bufs, bpipes)])
    chosen, _ = batch.tune(sh, None, 5)
    batched = timed(lambda: batch.launch(sh))
    batch.synchronize()
    batched["schedule"] = {"tile_height": [list(r) for r in chosen.tile_height],
                           "tiles_per_item": [list(r) for r in chosen.tiles_per_item],
                           "awb_tiles_per_item": chosen.awb_tiles_per_item, "blocks_per_cu": chosen.blocks_per_cu,
                           "order": chosen.order, "queue": chosen.queue, "graph": chosen.graph,
                           "per_frame_kernels": chosen.per_frame}
    return {"workload": "16 streams: 4 x 1280x720 BGGR RAW10, 8 x 1920x1200 RGGB RAW12, 4 x 3840x2160 GBRG u16 "
                        "14-bit; gray-world, bilinear/Malvar alternating, RGB8 + histograms + stats",
            "pixels_per_tick": pixels, "ticks": ticks, "per_frame_launches": single, "batched": batched,
            "speedup": round(single["ms_per_tick"] / batched["ms_per_tick"], 3)}


def large_frames(frames, ticks):
    """The large-frame batch: `frames` streams of 8192x6144 RGGB u16 at 16 bits (BL 1024, white
    patch, Malvar, sharpening, 65536-entry tone table, RGB16 + 1024-bin histograms), as per-frame
    launches and as one batch. One scene is rendered and copied to every stream's buffer."""
    import torch
    import holoscan_isp as isp
    import cases
    sc, params, out_format = cases.config5()
    frame = isp.render_scene(sc)
    W, H = frame.width, frame.height
    d_raw = [torch.from_numpy(frame.data).to("cuda") for _ in range(frames)]
    pitch = W * 6
    outs = [torch.empty(pitch * H, dtype=torch.uint8, device="cuda") for _ in range(frames)]
    hists = [torch.empty(4 * 1024, dtype=torch.int32, device="cuda") for _ in range(frames)]
    stats = [torch.empty(12, dtype=torch.float32, device="cuda") for _ in range(frames)]
    stream = torch.cuda.current_stream()
    sh = stream.cuda_stream
    exposure = (10_000_000, 2.0)
    pixels = frames * W * H
    algo = frames * W * H * (2 + 6)

    def timed(step):
        step()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(ticks):
            step()
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / ticks
        return {"ms_per_batch": round(ms, 3), "gpix_s": round(pixels / ms / 1e6, 3),
                "hbm_gbs_algorithmic": round(algo / ms / 1e6, 1)}

    pipes = [isp.IspPipeline(params) for _ in range(frames)]
    views = [frame.view(d) for d in d_raw]

    def per_frame():
        for i in range(frames):
            pipes[i].process(views[i], outs[i], pitch, out_format, hists[i], stats[i], exposure, sh)

    single = timed(per_frame)
    bpipes = [isp.IspPipeline(params) for _ in range(frames)]
    batch = isp.IspBatch()
    batch.set_frames([(bpipes[i], views[i], outs[i], pitch, out_format, hists[i], stats[i], exposure)
                      for i in range(frames)])
    chosen, _ = batch.tune(sh, None, 2)
    batched = timed(lambda: batch.launch(sh))
    batch.synchronize()
    return {"workload": f"{frames} x 8192x6144 RGGB u16 16-bit, BL 1024, white patch, Malvar, sharpen 0.5, CCM, "
                        "65536-entry sRGB table, RGB16 + 4x1024 histograms + stats",
            "algorithmic_bytes_per_batch": algo, "per_frame_launches": single, "batched": batched,
            "batched_queue": chosen.queue, "batched_per_frame_kernels": chosen.per_frame}


def served(seconds, per_frame, device_source, saturate=False, fps=30.0):
    """32 cameras at 30 fps with Poisson arrivals (resolutions uniform over 720p / 1080p / 1920x1200
    / 4K, RAW10 or RAW12, white patch 0.99, device exposure feedback), submitted in real time to
    the serving loop; latency = submit to the loop seeing the launch finish. With `saturate`,
    every frame of the trace is submitted back to back before any result is collected, for
    the loop's capacity."""
    import torch
    import holoscan_isp as isp
    from holoscan_isp import _abi
    import cases
    specs = cases.config4()
    server = isp.IspServer(hbm_budget=4 << 30, max_batch=32, launch_slots=2, per_frame=1 if per_frame else 0)
    ids = []
    sources = []
    for sc, p, of in specs:
        ids.append(server.add_camera(p, sc.width, sc.height, sc.format, sc.bit_depth, sc.pattern, of, True,
                                     (10_000_000, 2.0)))
        f = isp.render_scene(sc)
        if device_source:
            buf = torch.from_numpy(f.data).to("cuda")
        else:
            buf = torch.empty(f.data.size, dtype=torch.uint8, pin_memory=True)
            buf.numpy()[:] = f.data
        sources.append((buf, f.pitch, sc.width * sc.height))
    torch.cuda.synchronize()
    location = _abi.LOCATION_DEVICE if device_source else _abi.LOCATION_HOST
    events = cases.poisson_arrivals(len(specs), fps, seconds)
    lat = []
    pixels = 0
    first = None
    last = None
    pending = 0

    def drain(timeout_us=0):
        nonlocal pending, pixels, last
        while True:
            r = server.poll(timeout_us)
            if r is None:
                return
            if r.status != _abi.OK:
                raise RuntimeError(isp.last_error())
            lat.append((r.done_ns - r.submit_ns) / 1e6)
            pixels += sources[r.camera][2]
            last = r.done_ns if last is None else max(last, r.done_ns)
            server.release(r.request)
            pending -= 1
            timeout_us = 0

    t0 = isp.now_ns()
    for t, c, k in events:
        if not saturate:
            due = t0 + int(t * 1e9)
            while isp.now_ns() < due:
                drain()
        buf, pitch, _ = sources[c]
        server.submit(ids[c], buf, pitch, location, tag=k)
        first = isp.now_ns() if first is None else first
        pending += 1
        if not saturate:
            drain()
    while pending > 0:
        drain(1_000_000)
    counters = server.counters()
    server.close()
    span = (last - first) / 1e9
    return {"frames": len(lat), "latency_ms_p50": round(percentile(lat, 50), 4),
            "latency_ms_p99": round(percentile(lat, 99), 4), "latency_ms_max": round(max(lat), 4),
            "gpix_s": round(pixels / span / 1e9, 4), "launches": int(counters.launches),
            "frames_per_launch": round(len(lat) / max(1, counters.launches), 2),
            "max_launches_in_flight": int(counters.max_in_flight),
            "budget_peak_mib": round(counters.budget_peak / 2 ** 20, 1)}


def served_report(seconds):
    out = {"workload": "32 cameras, 30 fps Poisson arrivals, 720p/1080p/1920x1200/4K RAW10/RAW12, white patch "
                       "0.99, device exposure feedback, RGB8 + histograms + stats; 4 GiB budget, 2 launch slots, "
                       "32 frames per launch", "seconds": seconds}
    for source in ("device", "host_pinned"):
        dev = source == "device"
        out[source] = {"served": served(seconds, False, dev), "per_frame_launches": served(seconds, True, dev)}
    cap_seconds = min(seconds, 1.0)
    for scale in (4, 10):
        sec = min(seconds, 0.5)
        out[f"device_{scale}x_rate"] = {"fps": 30 * scale, "seconds": sec,
                                       "served": served(sec, False, True, fps=30.0 * scale),
                                       "per_frame_launches": served(sec, True, True, fps=30.0 * scale)}
    out["capacity_device"] = {"served": served(cap_seconds, False, True, saturate=True),
                              "per_frame_launches": served(cap_seconds, True, True, saturate=True)}
    return out


def source_and_check(launches, cpu=True):
    """The device raw source (isp_synthetic_render_device) and the GPU frame checker
    (isp_check_frame_device) on the reference demo's chart camera: each kernel's average
    launch time from HIP events over `launches` launches on one stream, at 4K RAW12 and 1080p
    u16, with the checked frame's PSNR and mismatched bins. With `cpu`, the CPU paths beside
    them, one thread, one call each: the reference's host renderer compiled from its sources
    (synthetic_raw.cpp) and the oracle's checker (a restatement of frame_check.cu)."""
    import torch
    import holoscan_isp as isp
    from holoscan_isp import _abi
    import cases
    stream = torch.cuda.current_stream()
    out = {"workload": "demo chart camera (vignetting 0.4, per-channel black), Malvar, manual white balance, "
                       "RGB8; render then check", "launches": launches}
    for name, w, h, packing in (("3840x2160_raw12", 3840, 2160, "raw12"), ("1920x1080_u16", 1920, 1080, "none")):
        cam = cases.demo_camera(width=w, height=h, packing=packing, frame=3)
        params = cases.demo_params(cam)
        pitch = isp.lib().isp_raw_row_bytes(cam.format, w)
        raw = torch.empty(pitch * h, dtype=torch.uint8, device="cuda")
        rgb = torch.empty(3 * w * h, dtype=torch.uint8, device="cuda")
        hist = torch.empty(1024, dtype=torch.int32, device="cuda")
        stats = torch.empty(12, dtype=torch.float32, device="cuda")
        pipe = isp.IspPipeline(params)
        checker = isp.FrameChecker(params)
        view = _abi.RawView(raw.data_ptr(), w, h, pitch, cam.format, cam.bit_depth, cam.pattern)

        def timed(body):
            for _ in range(3):
                body()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            for _ in range(launches):
                body()
            e1.record(stream)
            e1.synchronize()
            return e0.elapsed_time(e1) / launches * 1e3

        render_us = timed(lambda: isp.render_synthetic_device(cam, raw, pitch, stream.cuda_stream))
        pipe.process(view, rgb, 3 * w, _abi.OUT_RGB8, hist, stats, cases.DEMO_EXPOSURE, stream.cuda_stream)
        check_us = timed(lambda: checker.check(cam, rgb, hist, stats, 0, stream.cuda_stream))
        r = checker.result()
        out[name] = {"render_us": round(render_us, 2), "render_gpix_s": round(w * h / render_us / 1e3, 2),
                     "check_us": round(check_us, 2), "check_gpix_s": round(w * h / check_us / 1e3, 2),
                     "psnr_db": round(isp.frame_psnr(r, w, h), 3), "histogram_mismatches": r.histogram_mismatches}
        if cpu:
            out[name]["cpu"] = cpu_source_and_check(cam, params, rgb, hist, pitch)
        pipe.close()
    return out


def cpu_source_and_check(cam, params, rgb, hist, pitch):
    """One call each of the host renderer (the reference's when oracle/_ref is built, else
    this library's host render) and the oracle's frame check, on the frame just checked."""
    import ctypes as C
    import numpy as np
    import holoscan_isp as isp
    import oracle_lib
    w, h = cam.width, cam.height
    buf = np.zeros(pitch * h, dtype=np.uint8)
    t0 = time.perf_counter()
    if oracle_lib.ref_available():
        kind = "reference"
        if oracle_lib.ref().ref_synthetic_render(C.byref(cam), buf.ctypes.data, pitch) != 0:
            raise RuntimeError("ref_synthetic_render failed")
    else:
        kind = "port"
        isp.render_synthetic(cam)
    render_s = time.perf_counter() - t0
    truth, lut = isp.synthetic_truth(cam), isp.tone_lut(params)
    host_rgb, host_hist = rgb.cpu().numpy().reshape(h, w, 3), hist.cpu().numpy().view(np.uint32)
    t0 = time.perf_counter()
    oracle_lib.frame_check(truth, lut, host_rgb, host_hist)
    check_s = time.perf_counter() - t0
    return {"render_ms": round(render_s * 1e3, 1), "render_kind": kind, "check_ms": round(check_s * 1e3, 1),
            "check_kind": "port", "cores": 1}


def config1_frames(frames, cpu_frames):
    """The 1080p u16 workload (1920x1080 RGGB, 12 bits in 16, bilinear, manual white balance,
    lens shading, RGB8 + histograms + statistics): per-frame launches back to back, Gpix/s and
    HBM GB/s at its compulsory 5 B/px, per-frame p50/p99 from HIP events, a bit-exact check
    of a fresh pipeline's frame against the oracle, and the reference CPU backend (oracle/_ref)
    on one thread beside it."""
    import numpy as np
    import torch
    import holoscan_isp as isp
    import cases
    import gpu_support as gs
    import oracle_lib
    scene, params, out_format = cases.config1()
    frame = isp.render_scene(scene)
    w, h = frame.width, frame.height
    d_raw = torch.from_numpy(frame.data).to("cuda")
    out = torch.empty(3 * w * h, dtype=torch.uint8, device="cuda")
    hist = torch.empty(4 * 256, dtype=torch.int32, device="cuda")
    stats = torch.empty(12, dtype=torch.float32, device="cuda")
    pipe = isp.IspPipeline(params)
    view = frame.view(d_raw)
    stream = torch.cuda.current_stream()
    exposure = (10_000_000, 2.0)

    def step():
        pipe.process(view, out, 3 * w, out_format, hist, stats, exposure, stream.cuda_stream)

    for _ in range(20):
        step()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(frames):
        step()
    e1.record(stream)
    e1.synchronize()
    ms = e0.elapsed_time(e1) / frames
    per = []
    for _ in range(frames):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        step()
        b.record(stream)
        per.append((a, b))
    torch.cuda.synchronize()
    per = [a.elapsed_time(b) for a, b in per]
    pipe.close()
    g = gs.run(isp.IspPipeline(params), frame, out_format, True, exposure, d_raw=d_raw)
    o = oracle_lib.OraclePipeline(params).process(frame, out_format, True, exposure,
                                                 threads=min(16, os.cpu_count() or 1))
    exact = bool(gs.mismatches(g, o) == 0 and np.array_equal(g.histogram, o.histogram) and gs.same_stats(g, o))
    result = {"workload": "1920x1080 RGGB u16 12-bit, bilinear, manual white balance, lens shading, CCM, "
                          "gamma 2.2, RGB8 + histograms + stats", "frames": frames,
              "ms_per_frame": round(ms, 4), "gpix_s": round(w * h / ms / 1e6, 3),
              "hbm_gbs": round(w * h * 5 / ms / 1e6, 1), "frame_ms_p50": round(percentile(per, 50), 4),
              "frame_ms_p99": round(percentile(per, 99), 4), "bit_exact": exact}
    if cpu_frames > 0 and oracle_lib.ref_available():
        ref = oracle_lib.RefPipeline(params)
        ref.process(frame, out_format, True, exposure)
        t0 = time.perf_counter()
        for _ in range(cpu_frames):
            ref.process(frame, out_format, True, exposure)
        dt = (time.perf_counter() - t0) / cpu_frames
        result["cpu_baseline"] = {"value": round(w * h / dt / 1e9, 5), "unit": "Gpix/s", "cores": 1,
                                  "kind": "reference", "ms_per_frame": round(dt * 1e3, 1),
                                  "sample": f"{cpu_frames} frames after 1 warm-up"}
    return result


# The reference's own benchmark (modules/isp/tests/benchmark.cpp:89-160) and its published
# RTX 5090 times, ms per frame (modules/isp/README.md:285-292).
REFERENCE_CASES_RTX5090_MS = {
    ("1920x1080", "bilinear", "plain"): 0.017, ("1920x1080", "bilinear", "full"): 0.100,
    ("1920x1080", "malvar", "plain"): 0.019, ("1920x1080", "malvar", "full"): 0.093,
    ("3840x2160", "bilinear", "plain"): 0.067, ("3840x2160", "bilinear", "full"): 0.242,
    ("3840x2160", "malvar", "plain"): 0.074, ("3840x2160", "malvar", "full"): 0.251,
}


def reference_benchmark(iterations):
    """The reference's isp_benchmark cases on this library: the demo camera's default raw
    (12-bit u16 RGGB, vignetting 0.3) rendered on the device; "plain" = black level, manual
    white balance, demosaic, colour matrix, sRGB; "full" adds lens shading, gray world,
    sharpening 0.5, histograms and statistics; RGB8 out; 5 warm-up frames, then the mean of
    `iterations` frames between two HIP events on one stream, as benchmark.cpp:136-143 times
    them. The reference's published RTX 5090 times stand beside each."""
    import torch
    import holoscan_isp as isp
    from holoscan_isp import _abi
    stream = torch.cuda.current_stream()
    out = {"workload": "the reference benchmark's cases (modules/isp/tests/benchmark.cpp)", "iterations": iterations,
           "rtx5090_source": "modules/isp/README.md:285-292"}
    for w, h in ((1920, 1080), (3840, 2160)):
        cfg = isp.synthetic_config(width=w, height=h, vignetting=0.3)
        raw = torch.empty(isp.lib().isp_raw_row_bytes(cfg.format, w) * h + 2 * h, dtype=torch.uint8, device="cuda")
        pitch = isp.render_synthetic_device(cfg, raw, None, stream.cuda_stream)
        rgb = torch.empty(3 * w * h, dtype=torch.uint8, device="cuda")
        hist = torch.empty(1024, dtype=torch.int32, device="cuda")
        stats = torch.empty(12, dtype=torch.float32, device="cuda")
        view = _abi.RawView(raw.data_ptr(), w, h, pitch, cfg.format, cfg.bit_depth, cfg.pattern)
        for method, mname in ((_abi.BILINEAR, "bilinear"), (_abi.MALVAR, "malvar")):
            for full in (False, True):
                params = isp.synthetic_params(cfg)
                params.demosaic = method
                if full:
                    params.white_balance, params.sharpen = _abi.WB_GRAY_WORLD, 0.5
                else:
                    params.lens_shading = None
                pipe = isp.IspPipeline(params)
                h_out, s_out = (hist, stats) if full else (None, None)
                for _ in range(5):
                    pipe.process(view, rgb, 3 * w, _abi.OUT_RGB8, h_out, s_out, (0, 0.0), stream.cuda_stream)
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                for _ in range(iterations):
                    pipe.process(view, rgb, 3 * w, _abi.OUT_RGB8, h_out, s_out, (0, 0.0), stream.cuda_stream)
                e1.record(stream)
                e1.synchronize()
                ms = e0.elapsed_time(e1) / iterations
                pipe.close()
                key = (f"{w}x{h}", mname, "full" if full else "plain")
                ref = REFERENCE_CASES_RTX5090_MS[key]
                out["_".join(key)] = {"ms": round(ms, 4), "gpix_s": round(w * h / ms / 1e6, 2),
                                      "rtx5090_ms": ref, "speedup_vs_rtx5090": round(ref / ms, 2)}
    return out


def color_conversion(launches, cpu=True):
    """The conversion after the ISP (isp_color_convert_device) on a 4K frame: each kernel's
    average launch time from HIP events over `launches` launches, and its rate against the
    bytes it must move (read the RGB frame once, write the YCbCr planes once). With `cpu`, the
    colour module's own CPU converter (compiled from its sources) beside it, one thread, one
    call each."""
    import torch
    import numpy as np
    from holoscan_isp import _abi, color
    stream = torch.cuda.current_stream()
    out = {"workload": "3840x2160, BT.709 limited range, seeded RGB, device buffers", "launches": launches}
    w, h = 3840, 2160
    for name, src_fmt, dst_fmt in (("rgb8_to_nv12", _abi.PIX_RGB8, _abi.PIX_NV12),
                                   ("rgba8_to_nv12", _abi.PIX_RGBA8, _abi.PIX_NV12),
                                   ("rgba8_to_i420", _abi.PIX_RGBA8, _abi.PIX_I420),
                                   ("rgb8_to_yuyv", _abi.PIX_RGB8, _abi.PIX_YUYV)):
        src = color.Picture(src_fmt, w, h, device=True)
        src.buf.copy_(torch.from_numpy(np.random.default_rng(1).integers(0, 256, src.nbytes, dtype=np.uint8)))
        dst = color.Picture(dst_fmt, w, h, device=True)
        sv, dv = src.view(), dst.view()

        def run():
            color.convert_device(sv, dv, _abi.MATRIX_BT709, _abi.RANGE_LIMITED, stream.cuda_stream)

        for _ in range(3):
            run()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(launches):
            run()
        e1.record(stream)
        e1.synchronize()
        us = e0.elapsed_time(e1) / launches * 1e3
        moved = src.nbytes + dst.nbytes
        out[name] = {"us": round(us, 2), "bytes": moved, "gb_s": round(moved / us / 1e3, 1),
                     "hbm_frac": round(moved / us / 1e3 / HBM_PEAK_GBS, 4)}
        if cpu:
            import oracle_lib
            if oracle_lib.cc_ref_available():
                host_src = color.Picture(src_fmt, w, h)
                host_src.buf[:] = src.buf.cpu().numpy()
                host_dst = color.Picture(dst_fmt, w, h)
                t0 = time.perf_counter()
                oracle_lib.ref_color_convert(host_src, host_dst, _abi.MATRIX_BT709, _abi.RANGE_LIMITED)
                out[name]["cpu_ms"] = round((time.perf_counter() - t0) * 1e3, 1)
                out[name]["cpu_kind"] = "reference"
                out[name]["bit_exact"] = bool(np.array_equal(host_dst.buf, dst.buf.cpu().numpy()))
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus above 1 is not supported"}))
        sys.exit(2)
    if args.copy_mib:
        copy_calibration(args.copy_mib)
        return
    import numpy as np
    import torch

    import holoscan_isp as isp
    import cases
    import gpu_support as gs
    from oracle_lib import OraclePipeline

    torch.cuda.set_device(0)
    scene, params, out_format = cases.config2()
    # ROTATION distinct frames (seeds seed, seed + 1000, ...) and as many output buffers, 365 MB
    # together, more than the 256 MB last-level cache: frame n of the sequence is pair n % ROTATION.
    frames = render_frames(isp, scene, ROTATION)
    frame = frames[0]
    W, H = frame.width, frame.height
    pixels = W * H
    pitch = W * 4
    d_raws = [torch.from_numpy(f.data).to("cuda") for f in frames]
    views = [f.view(d) for f, d in zip(frames, d_raws)]
    outs = [torch.empty(pitch * H, dtype=torch.uint8, device="cuda") for _ in range(ROTATION)]
    hist = torch.empty(4 * 256, dtype=torch.int32, device="cuda")
    stats = torch.empty(12, dtype=torch.float32, device="cuda")
    pipe = isp.IspPipeline(params)
    stream = torch.cuda.current_stream()
    sh = stream.cuda_stream
    exposure = (10_000_000, 2.0)
    count = [0]

    def step():
        k = count[0] % ROTATION
        pipe.process(views[k], outs[k], pitch, out_format, hist, stats, exposure, sh)
        count[0] += 1

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()

    # Timed region: K frames back to back.
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    start.record(stream)
    for _ in range(args.steps):
        step()
    end.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = start.elapsed_time(end)
    elapsed = max(wall, gpu_ms / 1e3)
    gpix = args.steps * pixels / elapsed / 1e9
    gbs = args.steps * pixels * (RAW_BYTES_PER_PX + OUT_BYTES_PER_PX) / elapsed / 1e9
    # The timed sequence's last frame, kept for the bit-exact check.
    checked = count[0] - 1
    last = gs.Result(outs[checked % ROTATION].cpu().numpy(), hist.cpu().numpy().view(np.uint32),
                     stats.cpu().numpy())

    # Kernel timing pass: HIP events around each launch on the launch stream.
    pipe.timing_begin()
    frame_ev = []
    for _ in range(args.steps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        step()
        e1.record(stream)
        frame_ev.append((e0, e1))
    torch.cuda.synchronize()
    awb_ms, main_ms, n = pipe.timing_end()
    per_frame = [a.elapsed_time(b) for a, b in frame_ev]
    main_avg_s = main_ms / n / 1e3
    algo_bytes = pixels * (RAW_BYTES_PER_PX + OUT_BYTES_PER_PX)
    achieved = algo_bytes / main_avg_s / 1e9

    # Bit-exact check of the timed sequence's last frame: the oracle replays the white balance
    # state of the frames before it (gray world with smoothing carries it from frame to frame),
    # then runs that frame.
    threads = min(16, os.cpu_count() or 1)
    cpu = OraclePipeline(params)
    cpu.replay_white_balance(frames, [i % ROTATION for i in range(checked)], threads=threads)
    want = cpu.process(frames[checked % ROTATION], out_format, True, exposure, threads=threads)
    bit_exact = bool(gs.mismatches(last, want) == 0 and np.array_equal(last.histogram, want.histogram)
                     and gs.same_stats(last, want))

    traffic, traffic_detail = traffic_from_counters()
    result = {
        "metric": "Gpix/s + HBM GB/s (bit-exact); p99 frame latency, batched vs per-frame",
        "value": round(gpix, 3),
        "unit": "Gpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "3840x2160 GRBG MIPI RAW12 seeded noisy scene, BL 240, 17x13 LSC, gray-world WB, "
                        "Malvar, sharpen 0.5, CCM, sRGB, RGBA8 + 4x256 histograms + stats/AE, per-frame launches",
            "width": W, "height": H, "frames": args.steps,
            "distinct_frames": ROTATION,
        },
        "hbm_gbs": round(gbs, 2),
        "compulsory_bytes_per_px": RAW_BYTES_PER_PX + OUT_BYTES_PER_PX,
        "frame_ms_p50": round(percentile(per_frame, 50), 5),
        "frame_ms_p99": round(percentile(per_frame, 99), 5),
        "kernel_ms": {"awb_avg": round(awb_ms / n, 5), "main_avg": round(main_ms / n, 5)},
        "bit_exact": bit_exact,
        "bit_exact_frame": f"frame {checked} of the sequence (the timed region's last), white balance state replayed",
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "main_kernel<malvar,sharpen>", "algorithmic_bytes_per_launch": algo_bytes,
                     "traffic_detail": traffic_detail,
                     "valu": valu_from_counters(main_ms / n, pixels)},
    }
    if args.batch_ticks > 0:
        result["batched_16_streams"] = batched_vs_per_frame(args.batch_ticks)
    if args.large_ticks > 0:
        result["large_frames_8"] = large_frames(8, args.large_ticks)
    if args.served_seconds > 0:
        result["served_32_cameras"] = served_report(args.served_seconds)
    if args.source_launches > 0:
        result["device_source_and_check"] = source_and_check(args.source_launches, cpu=not args.no_cpu)
    if args.reference_iterations > 0:
        result["reference_benchmark_cases"] = reference_benchmark(args.reference_iterations)
    if args.config1_frames > 0:
        result["config1_1080p_u16"] = config1_frames(args.config1_frames, 0 if args.no_cpu else 10)
    if args.color_launches > 0:
        result["color_conversion_4k"] = color_conversion(args.color_launches, cpu=not args.no_cpu)
    if not args.no_cpu:
        base = cpu_baseline(frame, params, out_format, args.cpu_frames)
        result["cpu_baseline"] = base["single"]
        result["cpu_all_cores"] = base["all_cores"]
    print(json.dumps(result))


if __name__ == "__main__":
    main()
