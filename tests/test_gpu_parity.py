"""The HIP pipeline against the CPU oracle, bit for bit: images, histograms and
the whole statistics vector (the auto-exposure fields included), on the
reference GPU suite's cases (modules/isp/tests/pipeline_test.cpp:164-235), a
matrix of every pattern, sample format, demosaic, sharpening, white balance
mode, tone curve and output format, odd and tiny sizes, padded rows, and the
benchmark configurations' seeded noisy scenes."""
import itertools

import numpy as np
import pytest

import holoscan_isp as isp
from holoscan_isp import _abi

import cases
import gpu_support as gs
from oracle_lib import OraclePipeline

pytestmark = pytest.mark.gpu

EXPOSURE = (10_000_000, 2.0)


def compare(name, g, h):
    assert gs.mismatches(g, h) == 0, f"{name}: {gs.mismatches(g, h)} bytes differ"
    if h.histogram is not None:
        assert np.array_equal(g.histogram, h.histogram), f"{name}: histograms differ"
        assert gs.same_stats(g, h), f"{name}: stats {g.stats} vs {h.stats}"


@pytest.mark.parametrize("case", cases.matches_cpu_cases(), ids=lambda c: c[0])
def test_matches_oracle_reference_cases(torch_gpu, case):
    """cuda_matches_cpu (pipeline_test.cpp:164-235), AE fields bit-exact too."""
    name, config, out_format, tweak = case
    frame = isp.render_synthetic(config)
    params = isp.synthetic_params(config, 9)
    tweak(params)
    gpu = isp.IspPipeline(params)
    cpu = OraclePipeline(params)
    for _ in range(2):  # two frames, so the white balance smoothing runs too
        g = gs.run(gpu, frame, out_format, True, EXPOSURE)
        h = cpu.process(frame, out_format, True, EXPOSURE)
        compare(name, g, h)


MATRIX = list(itertools.product(range(len(cases.FORMATS)), cases.PATTERNS, (_abi.BILINEAR, _abi.MALVAR),
                                 (0.0, 0.7), (_abi.WB_MANUAL, _abi.WB_GRAY_WORLD, _abi.WB_WHITE_PATCH)))


@pytest.mark.parametrize("fi,pattern,method,sharpen,wb", MATRIX)
def test_matrix(torch_gpu, fi, pattern, method, sharpen, wb):
    config = cases.camera(pattern, cases.FORMATS[fi], 70, 38)
    frame = isp.render_synthetic(config)
    params = isp.synthetic_params(config, 5)
    params.demosaic, params.sharpen, params.white_balance = method, sharpen, wb
    params.tone_curve = [_abi.TONE_SRGB, _abi.TONE_GAMMA, _abi.TONE_LINEAR][(fi + pattern) % 3]
    out_format = (fi + pattern + method) % 4
    gpu = isp.IspPipeline(params)
    cpu = OraclePipeline(params)
    for _ in range(2):
        compare("matrix", gs.run(gpu, frame, out_format, True, EXPOSURE),
                cpu.process(frame, out_format, True, EXPOSURE))


def flat(width, height, r, g, b, pattern=_abi.RGGB):
    """A flat 16-bit frame (pipeline_test.cpp:459-482)."""
    cfa = [[0, 1, 2, 3], [3, 2, 1, 0], [1, 0, 3, 2], [2, 3, 0, 1]][pattern]
    img = np.zeros((height, width), dtype=np.uint16)
    for y in range(height):
        for x in range(width):
            ch = cfa[((y & 1) << 1) | (x & 1)]
            img[y, x] = r if ch == 0 else b if ch == 3 else g
    return isp.RawFrame(img.view(np.uint8).ravel(), width, height, 2 * width, _abi.RAW_U16, 16, pattern)


@pytest.mark.parametrize("size", [(2, 2), (3, 5), (33, 17), (5, 2), (2, 9), (131, 3)])
@pytest.mark.parametrize("pattern", cases.PATTERNS)
@pytest.mark.parametrize("method", [_abi.BILINEAR, _abi.MALVAR])
def test_tiny_sizes(torch_gpu, size, pattern, method):
    """edge_cases (pipeline_test.cpp:637-661): down to 2x2, every pattern, sharpened."""
    w, h = size
    frame = flat(w, h, 9000, 21000, 15000, pattern)
    params = isp.IspParams(tone_curve=_abi.TONE_LINEAR, white_level=65535.0, sharpen=1.0, demosaic=method)
    g = gs.run(isp.IspPipeline(params), frame, _abi.OUT_RGB8, True)
    c = OraclePipeline(params).process(frame, _abi.OUT_RGB8, True)
    compare(f"{w}x{h}", g, c)


PLAIN_SIZES = [(16, 16), (33, 17), (70, 38), (257, 66), (1000, 17), (96, 130), (130, 47)]
PLAIN_TINY = [(2, 2), (5, 3), (131, 3), (3, 37), (66, 2)]


def plain_params(params, method, wb, k):
    params.demosaic, params.sharpen, params.white_balance = method, 0.0, wb
    if k % 2:
        params.lens_shading = None
    return params


@pytest.mark.parametrize("fi", range(len(cases.FORMATS)))
@pytest.mark.parametrize("method", [_abi.BILINEAR, _abi.MALVAR])
@pytest.mark.parametrize("wb", [_abi.WB_MANUAL, _abi.WB_WHITE_PATCH])
def test_plain_frames(torch_gpu, fi, method, wb):
    """Frames with nothing to finish (no histograms or statistics, manual or white-patch
    white balance, no sharpening) run the plain kernel (16-row tiles, 256-thread blocks);
    every sample format, the patterns in turn, lens shading on and off, odd sizes, two
    frames each (white patch smooths across them), then a frame with statistics on the
    same pipeline (the full kernel after the plain one)."""
    for k, (w, h) in enumerate(PLAIN_SIZES):
        pattern = cases.PATTERNS[(fi + k) % len(cases.PATTERNS)]
        config = cases.camera(pattern, cases.FORMATS[fi], w, h)
        frame = isp.render_synthetic(config)
        params = plain_params(isp.synthetic_params(config, 5), method, wb, k)
        params.tone_curve = [_abi.TONE_SRGB, _abi.TONE_GAMMA, _abi.TONE_LINEAR][(fi + k) % 3]
        out_format = (fi + k + method) % 4
        gpu, cpu = isp.IspPipeline(params), OraclePipeline(params)
        for with_stats in (False, False, True):
            compare(f"{w}x{h} stats {with_stats}", gs.run(gpu, frame, out_format, with_stats, EXPOSURE),
                    cpu.process(frame, out_format, with_stats, EXPOSURE))


@pytest.mark.parametrize("fi", range(len(cases.FORMATS)))
def test_plain_large_frames(torch_gpu, fi):
    """Plain frames with more 64 x 16 tiles than the six-wave plain kernel keeps resident
    (2048 x 1200: 2,400 tiles), which run the eight-wave kernel (RAW10: the six-wave one),
    both demosaics, lens shading on and off, white patch on the second."""
    pattern = cases.PATTERNS[fi % len(cases.PATTERNS)]
    config = cases.camera(pattern, cases.FORMATS[fi], 2048, 1200)
    frame = isp.render_synthetic(config)
    for k, method in enumerate((_abi.BILINEAR, _abi.MALVAR)):
        params = plain_params(isp.synthetic_params(config, 5), method, (_abi.WB_MANUAL, _abi.WB_WHITE_PATCH)[k], k)
        out_format = (fi + k) % 4
        gpu, cpu = isp.IspPipeline(params), OraclePipeline(params)
        for _ in range(2):
            compare(f"2048x1200 {method}", gs.run(gpu, frame, out_format, False, EXPOSURE),
                    cpu.process(frame, out_format, False, EXPOSURE, threads=8))


@pytest.mark.parametrize("fmt", [_abi.RAW_U8, _abi.RAW_U16])
@pytest.mark.parametrize("method", [_abi.BILINEAR, _abi.MALVAR])
def test_plain_tiny_frames(torch_gpu, fmt, method):
    """The plain kernel on frames smaller than a tile in either direction (random samples, every
    pattern in turn, white patch on every other size)."""
    rng = np.random.default_rng(77 + fmt + 2 * method)
    for k, (w, h) in enumerate(PLAIN_TINY):
        pattern = cases.PATTERNS[k % len(cases.PATTERNS)]
        if fmt == _abi.RAW_U8:
            data = rng.integers(0, 256, size=w * h, dtype=np.uint8)
            frame = isp.RawFrame(data, w, h, w, fmt, 8, pattern)
        else:
            img = rng.integers(0, 4096, size=(h, w), dtype=np.uint16)
            frame = isp.RawFrame(img.view(np.uint8).ravel(), w, h, 2 * w, fmt, 12, pattern)
        wb = _abi.WB_WHITE_PATCH if k % 2 else _abi.WB_MANUAL
        params = isp.IspParams(black_level=16.0, wb_gains=(1.5, 1.0, 1.25), demosaic=method, white_balance=wb)
        out_format = (k + method) % 4
        gpu, cpu = isp.IspPipeline(params), OraclePipeline(params)
        for _ in range(2):
            compare(f"{w}x{h}", gs.run(gpu, frame, out_format, False), cpu.process(frame, out_format, False))


@pytest.mark.parametrize("case", cases.float_edge_cases(), ids=lambda c: c[0])
def test_float_edges(torch_gpu, case):
    """Subnormal samples (tiny white balance gains) and signed zeros reaching an RGB32F output,
    bit for bit."""
    name, frame, params, out_format = case
    gpu, cpu = isp.IspPipeline(params), OraclePipeline(params)
    for _ in range(2):
        compare(name, gs.run(gpu, frame, out_format, True, EXPOSURE), cpu.process(frame, out_format, True, EXPOSURE))
    # Without statistics (the plain kernel, where nothing else needs finishing).
    gpu, cpu = isp.IspPipeline(params), OraclePipeline(params)
    for _ in range(2):
        compare(name + " plain", gs.run(gpu, frame, out_format, False), cpu.process(frame, out_format, False))


@pytest.mark.parametrize("row", cases.LARGE_GRIDS, ids=lambda r: f"{r[0]}x{r[1]}x{r[2]} on {r[3]}x{r[4]}")
def test_large_lens_shading_grids(torch_gpu, row):
    """Grids of more than 1,024 values (read from device memory, through the tile tables and per
    sample) and the 1,024-value boundary, in both passes (white balance and main), two frames."""
    sc, params, out_format = cases.large_grid(row, 600 + row[0] + row[3])
    frame = isp.render_scene(sc)
    gpu, cpu = isp.IspPipeline(params), OraclePipeline(params)
    for _ in range(2):
        compare(f"grid {row[:3]}", gs.run(gpu, frame, out_format, True, EXPOSURE),
                cpu.process(frame, out_format, True, EXPOSURE, threads=8))


def test_random_mosaics(torch_gpu):
    """LCG mosaics (cpu_reference_test.cpp:57-66) at odd sizes, every option on."""
    rng = np.random.default_rng(12345)
    for w, h in [(13, 9), (6, 5), (97, 41), (257, 66), (64, 32), (65, 33)]:
        img = rng.integers(0, 4096, size=(h, w), dtype=np.uint16)
        frame = isp.RawFrame(img.view(np.uint8).ravel(), w, h, 2 * w, _abi.RAW_U16, 12, _abi.GBRG)
        for method, sharpen, wb, out in [(0, 0.0, 0, 0), (1, 2.0, 1, 1), (1, 0.3, 2, 2), (0, 4.0, 1, 3)]:
            params = isp.IspParams(channel_black_levels=(64.0, 60.0, 70.0, 50.0), wb_gains=(1.5, 1.0, 1.25),
                                   demosaic=method, sharpen=sharpen, white_balance=wb,
                                   ccm=(0.9, 0.2, -0.1, 0.05, 0.8, 0.15, -0.2, 0.1, 1.1))
            compare(f"{w}x{h}", gs.run(isp.IspPipeline(params), frame, out, True, EXPOSURE),
                    OraclePipeline(params).process(frame, out, True, EXPOSURE))


def test_padded_rows(torch_gpu):
    """padded_rows (pipeline_test.cpp:384-401): padding beyond the samples changes nothing."""
    for fi in (0, 2, 5):
        config = cases.camera(_abi.GBRG, cases.FORMATS[fi], 75, 41)
        packed = isp.render_synthetic(config)
        pitch = packed.pitch + 38
        data = np.full(pitch * config.height, 0xAB, dtype=np.uint8)
        for y in range(config.height):
            data[y * pitch:y * pitch + packed.pitch] = packed.data[y * packed.pitch:(y + 1) * packed.pitch]
        padded = isp.RawFrame(data, packed.width, packed.height, pitch, packed.format, packed.bit_depth,
                              packed.pattern)
        pipeline = isp.IspPipeline(isp.synthetic_params(config, 5))
        a = gs.run(pipeline, packed, _abi.OUT_RGB8)
        b = gs.run(pipeline, padded, _abi.OUT_RGB8)
        compare("padded", a, b)
        compare("padded vs oracle", b, OraclePipeline(isp.synthetic_params(config, 5)).process(padded))


@pytest.mark.parametrize("size", [(1920, 1080), (3840, 2160)])
def test_reference_benchmark_cases(torch_gpu, size):
    """The reference benchmark's eight cases (modules/isp/tests/benchmark.cpp:89-160) as the
    bench times them: the demo camera's default raw with vignetting 0.3, plain (no lens
    shading) and full (gray world, sharpening 0.5, histograms and statistics), bilinear and
    Malvar, RGB8; three frames each, so the gray-world state runs too."""
    w, h = size
    cfg = isp.synthetic_config(width=w, height=h, vignetting=0.3)
    frame = isp.render_synthetic(cfg)
    for method in (_abi.BILINEAR, _abi.MALVAR):
        for full in (False, True):
            print(f"reference case {w}x{h} method {method} full {full}", flush=True)
            params = isp.synthetic_params(cfg)
            params.demosaic = method
            if full:
                params.white_balance, params.sharpen = _abi.WB_GRAY_WORLD, 0.5
            else:
                params.lens_shading = None
            gpu, cpu = isp.IspPipeline(params), OraclePipeline(params)
            for _ in range(3):
                compare(f"{w}x{h} {method} {full}", gs.run(gpu, frame, _abi.OUT_RGB8, full, (0, 0.0)),
                        cpu.process(frame, _abi.OUT_RGB8, full, (0, 0.0), threads=8))


@pytest.mark.parametrize("which", ["config1", "config2", "config5"])
def test_baseline_configs_reduced(torch_gpu, which):
    """The benchmark configurations' pipelines on their seeded noisy scenes at reduced sizes."""
    make = getattr(cases, which)
    for w, h in [(320, 180), (258, 130)]:
        sc, params, out_format = make(w, h)
        frame = isp.render_scene(sc)
        gpu = isp.IspPipeline(params)
        cpu = OraclePipeline(params)
        for _ in range(2):
            compare(which, gs.run(gpu, frame, out_format, True, EXPOSURE),
                    cpu.process(frame, out_format, True, EXPOSURE, threads=8))


@pytest.mark.parametrize("which", ["config1", "config2"])
def test_baseline_configs_full_size(torch_gpu, which):
    """Configs 1 and 2 at full size (1920x1080 u16, 3840x2160 RAW12), two frames."""
    sc, params, out_format = getattr(cases, which)()
    frame = isp.render_scene(sc)
    gpu = isp.IspPipeline(params)
    cpu = OraclePipeline(params)
    for _ in range(2):
        compare(which, gs.run(gpu, frame, out_format, True, EXPOSURE),
                cpu.process(frame, out_format, True, EXPOSURE, threads=16))


def test_config5_full_size(torch_gpu):
    """The large-frame configuration at full size: 8192x6144 u16 at 16 bits, white patch, Malvar,
    sharpening, a 65536-entry tone table (read from global memory), RGB16 and 1024-bin
    histograms; one frame against the oracle, and the histogram rebuilt from the image."""
    from test_oracle import rebuild_histogram
    sc, params, out_format = cases.config5()
    frame = isp.render_scene(sc)
    g = gs.run(isp.IspPipeline(params), frame, out_format, True, EXPOSURE)
    compare("config5", g, OraclePipeline(params).process(frame, out_format, True, EXPOSURE, threads=16))
    assert np.array_equal(g.histogram, rebuild_histogram(g.image, out_format, frame.width, frame.height, 1024))
    assert all(int(row.sum()) == frame.width * frame.height for row in g.histogram.reshape(4, 1024))


def test_determinism(torch_gpu):
    """determinism (pipeline_test.cpp:621-634): same bytes on every run and instance."""
    config = cases.camera(_abi.RGGB, cases.FORMATS[2], 640, 480)
    frame = isp.render_synthetic(config)
    params = isp.synthetic_params(config)
    params.white_balance, params.sharpen = _abi.WB_GRAY_WORLD, 1.0
    a = gs.run(isp.IspPipeline(params), frame)
    for _ in range(3):
        compare("again", gs.run(isp.IspPipeline(params), frame), a)


GOLDEN_CASES = None


def _golden_cases():
    global GOLDEN_CASES
    if GOLDEN_CASES is None:
        import golden_cases
        GOLDEN_CASES = golden_cases.cases()
    return GOLDEN_CASES


def test_golden_fixtures(torch_gpu):
    """The HIP path reproduces the hashes the reference's compiled CPU backend recorded
    (tests/golden/golden.json); 1024-bin histograms (no reference counterpart) are
    checked by rebuilding them from the written image."""
    import hashlib
    import json
    import os
    from test_oracle import rebuild_histogram
    with open(os.path.join(os.path.dirname(__file__), "golden", "golden.json")) as f:
        golden = json.load(f)["cases"]
    for name, frame, params, out_format, exposure in _golden_cases():
        g = golden[name]
        assert hashlib.sha256(frame.data.tobytes()).hexdigest() == g["raw"], name
        pipe = isp.IspPipeline(params)
        for expected in g["frames"]:
            r = gs.run(pipe, frame, out_format, True, exposure)
            assert hashlib.sha256(r.image.tobytes()).hexdigest() == expected["image"], name
            assert hashlib.sha256(r.stats.tobytes()).hexdigest() == expected["stats"], name
            if (params.histogram_bins or 256) == 256:
                assert hashlib.sha256(r.histogram.tobytes()).hexdigest() == expected["histogram"], name
            else:
                assert np.array_equal(r.histogram, rebuild_histogram(r.image, out_format, frame.width,
                                                                     frame.height, params.histogram_bins))
