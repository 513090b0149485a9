"""The CPU oracle pinned: against the reference's own CPU backend compiled from
its sources (byte for byte), against the golden fixtures that backend wrote,
and against closed forms the mathematics fixes (the reference's host tests,
modules/isp/tests/host/cpu_reference_test.cpp, and the closed-form checks of
its GPU suite, modules/isp/tests/pipeline_test.cpp)."""
import ctypes as C
import dataclasses
import hashlib
import itertools
import json
import math
import os

import numpy as np
import pytest

import holoscan_isp as isp
from holoscan_isp import _abi

import cases
import golden_cases
import oracle_lib
from oracle_lib import OraclePipeline

HERE = os.path.dirname(os.path.abspath(__file__))
needs_ref = pytest.mark.skipif(not oracle_lib.ref_available(),
                               reason="the reference CPU backend is built only where the reference is mounted")


def sha(a):
    return hashlib.sha256(a.tobytes()).hexdigest()


def golden():
    with open(os.path.join(HERE, "golden", "golden.json")) as f:
        return json.load(f)["cases"]


GOLDEN_CASES = golden_cases.cases()


@pytest.mark.parametrize("case", GOLDEN_CASES, ids=lambda c: c[0])
def test_oracle_matches_golden(case):
    """The oracle reproduces the reference backend's recorded hashes; the inputs hash the same."""
    name, frame, params, out_format, exposure = case
    g = golden()[name]
    assert sha(frame.data) == g["raw"], f"{name}: the input generator changed"
    o = OraclePipeline(params)
    for expected in g["frames"]:
        r = o.process(frame, out_format, True, exposure)
        assert sha(r.image) == expected["image"], name
        assert sha(r.stats) == expected["stats"], name
        if (params.histogram_bins or 256) == 256:
            assert sha(r.histogram) == expected["histogram"], name


MATRIX = list(itertools.product(range(len(cases.FORMATS)), cases.PATTERNS, (_abi.BILINEAR, _abi.MALVAR),
                                (0.0, 0.7), (_abi.WB_MANUAL, _abi.WB_GRAY_WORLD, _abi.WB_WHITE_PATCH)))


@needs_ref
@pytest.mark.parametrize("fi,pattern,method,sharpen,wb", MATRIX[::5])
def test_oracle_matches_reference_matrix(fi, pattern, method, sharpen, wb):
    config = cases.camera(pattern, cases.FORMATS[fi], 41, 23)
    frame = isp.render_synthetic(config)
    params = isp.synthetic_params(config, 5)
    params.demosaic, params.sharpen, params.white_balance = method, sharpen, wb
    params.tone_curve = [_abi.TONE_SRGB, _abi.TONE_GAMMA, _abi.TONE_LINEAR][(fi + pattern) % 3]
    out_format = (fi + pattern + method) % 4
    o, r = OraclePipeline(params), oracle_lib.RefPipeline(params)
    for _ in range(2):
        a = o.process(frame, out_format, True, (10_000_000, 2.0))
        b = r.process(frame, out_format, True, (10_000_000, 2.0))
        assert np.array_equal(a.image, b.image)
        assert np.array_equal(a.histogram, b.histogram)
        assert np.array_equal(a.stats.view(np.uint32), b.stats.view(np.uint32))


@needs_ref
@pytest.mark.parametrize("case", cases.float_edge_cases(), ids=lambda c: c[0])
def test_oracle_matches_reference_at_float_edges(case):
    """Subnormal samples and signed zeros: the oracle's bits are the compiled reference's."""
    name, frame, params, out_format = case
    a = OraclePipeline(params).process(frame, out_format, True, (10_000_000, 2.0))
    b = oracle_lib.RefPipeline(params).process(frame, out_format, True, (10_000_000, 2.0))
    assert np.array_equal(a.image, b.image), name
    assert np.array_equal(a.histogram, b.histogram), name
    assert np.array_equal(a.stats.view(np.uint32), b.stats.view(np.uint32)), name


def test_float_edge_cases_reach_the_edges():
    """The cases do what they are for: subnormal linear values, and -0.0 in an RGB32F output."""
    edge = {name: (frame, params, fmt) for name, frame, params, fmt in cases.float_edge_cases()}
    frame, params, fmt = edge["subnormal_malvar"]
    img = OraclePipeline(params).process(frame, fmt, True, (10_000_000, 2.0)).image.view(np.float32)
    nonzero = np.abs(img[img != 0])
    assert nonzero.size > 0 and nonzero.max() < np.float32(1.2e-38)
    frame, params, fmt = edge["signed_zero_1_0.0"]
    img = OraclePipeline(params).process(frame, fmt, True, (10_000_000, 2.0)).image.view(np.float32)
    assert np.any((img == 0) & np.signbit(img))
    frame, params, fmt = edge["table_top_0_0.0"]
    img = OraclePipeline(params).process(frame, fmt, True, (10_000_000, 2.0)).image.view(np.float32)
    assert np.count_nonzero((img == 0) & np.signbit(img)) > 100  # the saturated band


@needs_ref
def test_oracle_threads_match_reference():
    """Row-band threads change nothing: the sums are integers."""
    scene, params, out_format = cases.config2(384, 216, seed=3)
    frame = isp.render_scene(scene)
    b = oracle_lib.RefPipeline(params).process(frame, out_format, True, (10_000_000, 2.0))
    for threads in (1, 3, 8):
        a = OraclePipeline(params).process(frame, out_format, True, (10_000_000, 2.0), threads=threads)
        assert np.array_equal(a.image, b.image) and np.array_equal(a.histogram, b.histogram)
        assert np.array_equal(a.stats.view(np.uint32), b.stats.view(np.uint32))


@pytest.mark.parametrize("wb", [_abi.WB_GRAY_WORLD, _abi.WB_WHITE_PATCH])
def test_white_balance_replay_matches_full_frames(wb):
    """The replay of a frame sequence's white balance (statistics measured once per distinct frame,
    then applied in order) leaves the state that running every frame leaves, bit for bit, and the
    next frame then processes identically: smoothing 0.5 over the sequence 0 1 2 0 1 2 0."""
    scene, params, out_format = cases.config2(192, 108, seed=5)
    params = dataclasses.replace(params, white_balance=wb, awb_smoothing=0.5)
    frames = []
    for k in range(3):
        scene.seed = 5 + 1000 * k
        frames.append(isp.render_scene(scene))
    sequence = [0, 1, 2, 0, 1, 2, 0]
    full = OraclePipeline(params)
    for i in sequence:
        full.process(frames[i], out_format, True, (10_000_000, 2.0))
    replay = OraclePipeline(params)
    replay.replay_white_balance(frames, sequence, threads=2)
    g_full, v_full = full.awb_state()
    g_replay, v_replay = replay.awb_state()
    assert v_full == v_replay == 1
    assert np.array_equal(np.float32(g_full).view(np.uint32), np.float32(g_replay).view(np.uint32))
    assert g_full != OraclePipeline(params).awb_state()[0]  # the sequence moved the state
    a = full.process(frames[1], out_format, True, (10_000_000, 2.0))
    b = replay.process(frames[1], out_format, True, (10_000_000, 2.0))
    assert np.array_equal(a.image, b.image) and np.array_equal(a.histogram, b.histogram)
    assert np.array_equal(a.stats.view(np.uint32), b.stats.view(np.uint32))
    if oracle_lib.ref_available():
        ref = oracle_lib.RefPipeline(params)
        for i in sequence:
            ref.process(frames[i], out_format, True, (10_000_000, 2.0))
        g = (C.c_float * 3)()
        v = C.c_int()
        oracle_lib.ref().ref_awb_state(ref._h, g, C.byref(v))
        assert np.array_equal(np.float32(list(g)).view(np.uint32), np.float32(g_replay).view(np.uint32))


def test_mirror_brute_force():
    """mirror() is reflection about the first and last pixel (isp_math.hpp:82-91)."""
    o = oracle_lib.oracle()
    for n in (2, 3, 5, 16):
        seq = list(range(n)) + list(range(n - 2, 0, -1))  # one period
        for v in range(-3 * n, 3 * n):
            assert o.oracle_mirror(v, n) == seq[v % len(seq)]


def test_load_raw_bit_layout():
    """RAW10: four 8-bit MSBs then a byte of 2-bit LSBs; RAW12: two MSB bytes then the nibbles."""
    rng = np.random.default_rng(3)
    o = oracle_lib.oracle()
    for fmt, bits, per, nbytes in ((_abi.RAW_RAW10, 10, 4, 5), (_abi.RAW_RAW12, 12, 2, 3)):
        w = 12
        samples = rng.integers(0, 1 << bits, size=w)
        row = bytearray()
        for u in range(w // per):
            group = samples[u * per:(u + 1) * per]
            lsb = 0
            for k, s in enumerate(group):
                row.append(int(s) >> (bits - 8))
                lsb |= (int(s) & ((1 << (bits - 8)) - 1)) << ((bits - 8) * k)
            row.append(lsb)
        data = np.frombuffer(bytes(row), dtype=np.uint8).copy()
        view = _abi.RawView(data.ctypes.data, w, 1, len(row), fmt, bits, 0)
        got = [o.oracle_load_raw(C.byref(view), x, 0) for x in range(w)]
        assert got == [int(s) for s in samples]


def _process(params, frame, threads=1):
    return OraclePipeline(params).process(frame, _abi.OUT_RGB32F, False, threads=threads).image.view(np.float32)


def color_at(pattern, x, y):
    name = ["RGGB", "BGGR", "GRBG", "GBRG"][pattern]
    c = name[(y & 1) * 2 + (x & 1)]
    return 0 if c == "R" else 1 if c == "G" else 2


def test_bilinear_against_independent_reference():
    """cpu_reference_test.cpp:57-119: each missing colour the mean of its 3x3 neighbours."""
    for pattern in cases.PATTERNS:
        for w, h in ((13, 9), (2, 2), (6, 5)):
            raw = np.zeros(w * h, dtype=np.uint16)
            seed = 12345
            for i in range(w * h):
                seed = (seed * 1664525 + 1013904223) & 0xFFFFFFFF
                raw[i] = 64 + (seed >> 16) % 4000
            params = isp.IspParams(demosaic=_abi.BILINEAR, tone_curve=_abi.TONE_LINEAR,
                                   channel_black_levels=(64.0, 60.0, 70.0, 50.0), wb_gains=(1.5, 1.0, 1.25))
            frame = isp.RawFrame(raw.view(np.uint8), w, h, 2 * w, _abi.RAW_U16, 12, pattern)
            got = _process(params, frame).reshape(h, w, 3)

            def reflect(v, n):
                while v < 0 or v >= n:
                    v = -v if v < 0 else 2 * (n - 1) - v
                return v

            def sample(x, y):
                c = color_at(pattern, x, y)
                red_row = color_at(pattern, 0, y) == 0 or color_at(pattern, 1, y) == 0
                ch = 0 if c == 0 else 3 if c == 2 else 1 if red_row else 2
                black = params.channel_black_levels[ch]
                v = (float(raw[y * w + x]) - black) / (4095.0 - black)
                return min(max(v, 0.0) * params.wb_gains[c], 1.0)

            worst = 0.0
            for y in range(h):
                for x in range(w):
                    for c in range(3):
                        if color_at(pattern, x, y) == c:
                            want = sample(x, y)
                        else:
                            vals = [sample(reflect(x + dx, w), reflect(y + dy, h))
                                    for dy in (-1, 0, 1) for dx in (-1, 0, 1)
                                    if (dx or dy) and color_at(pattern, x + dx, y + dy) == c]
                            want = sum(vals) / len(vals)
                        worst = max(worst, abs(float(got[y, x, c]) - want))
            assert worst < 1e-6, (pattern, w, h, worst)


def test_malvar_exact_on_gradients():
    """cpu_reference_test.cpp:122-157: Malvar-He-Cutler reproduces linear gradients, < 2e-6."""
    w, h = 24, 20
    a, bx, by = (9000.0, 20000.0, 14000.0), (700.0, -300.0, 450.0), (-250.0, 400.0, 600.0)
    for pattern in cases.PATTERNS:
        raw = np.zeros((h, w), dtype=np.uint16)
        for y in range(h):
            for x in range(w):
                c = color_at(pattern, x, y)
                raw[y, x] = int(a[c] + bx[c] * x + by[c] * y)
        params = isp.IspParams(demosaic=_abi.MALVAR, tone_curve=_abi.TONE_LINEAR)
        frame = isp.RawFrame(raw.view(np.uint8).ravel(), w, h, 2 * w, _abi.RAW_U16, 16, pattern)
        got = _process(params, frame).reshape(h, w, 3)
        worst = max(abs(float(got[y, x, c]) - (a[c] + bx[c] * x + by[c] * y) / 65535.0)
                    for y in range(2, h - 2) for x in range(2, w - 2) for c in range(3))
        assert worst < 2e-6, (pattern, worst)


def test_packed_formats_equal_unpacked():
    """cpu_reference_test.cpp:167-187: RAW10/RAW12 give the unpacked samples' image."""
    for fmt, bits in ((_abi.RAW_RAW10, 10), (_abi.RAW_RAW12, 12)):
        kw = dict(width=70, height=30, bit_depth=bits, black_levels=(64.0, 64.0, 64.0, 64.0))
        unpacked = isp.render_synthetic(isp.synthetic_config(format=_abi.RAW_U16, **kw))
        packed = isp.render_synthetic(isp.synthetic_config(format=fmt, **kw))
        params = isp.synthetic_params(isp.synthetic_config(format=fmt, **kw))
        assert np.array_equal(_process(params, unpacked), _process(params, packed))


def test_synthetic_truth_psnr():
    """cpu_reference_test.cpp:190-221: > 32 dB Malvar, > 26.5 dB bilinear, linear output."""
    config = isp.synthetic_config(width=320, height=240, vignetting=0.4,
                                  black_levels=(256.0, 260.0, 252.0, 258.0))
    frame = isp.render_synthetic(config)
    truth = isp.synthetic_truth(config)
    for method, floor in ((_abi.BILINEAR, 26.5), (_abi.MALVAR, 32.0)):
        params = isp.synthetic_params(config)
        params.demosaic, params.tone_curve = method, _abi.TONE_LINEAR
        got = _process(params, frame).reshape(240, 320, 3).astype(np.float64)
        mse = float(np.mean((got - truth) ** 2))
        assert 10 * math.log10(1.0 / mse) > floor


def _flat(w, h, r, g, b, pattern=_abi.RGGB):
    return golden_cases._flat(w, h, r, g, b, pattern)


def test_color_matrix_closed_form():
    """pipeline_test.cpp:492-514: a flat colour through the CCM, < 2e-6."""
    params = isp.IspParams(tone_curve=_abi.TONE_LINEAR, white_level=65535.0, wb_gains=(1.25, 1.0, 0.8),
                           ccm=(0.9, 0.2, -0.1, 0.05, 0.8, 0.15, -0.2, 0.1, 1.1))
    got = _process(params, _flat(32, 24, 20000, 30000, 40000, _abi.GBRG)).reshape(24, 32, 3)
    cam = (20000.0 / 65535.0 * 1.25, 30000.0 / 65535.0, 40000.0 / 65535.0 * 0.8)
    for k in range(3):
        want = sum(params.ccm[3 * k + j] * cam[j] for j in range(3))
        assert np.max(np.abs(got[:, :, k] - want)) < 2e-6


def test_tone_curves_closed_form():
    """pipeline_test.cpp:517-565: sRGB, gamma, linear and a table against their formulas, < 3e-5."""
    w = 256
    raw = np.tile((np.arange(w) * 257).astype(np.uint16), (4, 1))
    frame = isp.RawFrame(raw.view(np.uint8).ravel(), w, 4, 2 * w, _abi.RAW_U16, 16, _abi.RGGB)
    curves = [(_abi.TONE_SRGB, (), lambda x: 12.92 * x if x <= 0.0031308 else 1.055 * x ** (1 / 2.4) - 0.055),
              (_abi.TONE_GAMMA, (), lambda x: x ** (1 / 2.2)), (_abi.TONE_LINEAR, (), lambda x: x),
              (_abi.TONE_LUT, (1.0, 0.0), lambda x: 1.0 - x)]
    for curve, lut, formula in curves:
        params = isp.IspParams(tone_curve=curve, lut=lut, white_level=65535.0, demosaic=_abi.BILINEAR)
        got = _process(params, frame).reshape(4, w, 3)
        worst = max(abs(float(got[1, x, k]) - formula(x / 255.0)) for x in range(1, w - 1) for k in range(3))
        assert worst < 3e-5, (curve, worst)


def test_statistics_and_exposure_closed_form():
    """pipeline_test.cpp:568-618: flat mid-gray one stop under 0.18: luma bin 23, error 1 EV,
    scale 2, 30 ms wanted -> 20 ms at gain 1.5; no suggestion without the frame's exposure."""
    level = int(0.09 * 65535.0 + 0.5)
    frame = _flat(64, 48, level, level, level)
    params = isp.IspParams(tone_curve=_abi.TONE_LINEAR, white_level=65535.0,
                           auto_exposure=isp.AutoExposureParams(target=0.18, damping=1.0,
                                                                max_exposure_ns=20_000_000))
    o = OraclePipeline(params)
    r = o.process(frame, _abi.OUT_RGB8, True, (15_000_000, 1.0))
    assert all(r.histogram[row * 256:(row + 1) * 256].sum() == 64 * 48 for row in range(4))
    assert r.histogram[23] == 64 * 48
    mean = np.float32(level) / np.float32(65535.0)
    assert abs(r.stats[_abi.STAT_MEAN_R] - mean) < 1e-6
    assert abs(r.stats[_abi.STAT_AE_ERROR_EV] - 1.0) < 1e-3
    assert abs(r.stats[_abi.STAT_EXPOSURE_SCALE] - 2.0) < 2e-3
    assert abs(r.stats[_abi.STAT_SUGGESTED_EXPOSURE_NS] - 20e6) < 1.0
    assert abs(r.stats[_abi.STAT_SUGGESTED_GAIN] - 1.5) < 2e-3
    assert r.stats[_abi.STAT_SATURATED_FRACTION] == 0.0
    assert o.process(frame, _abi.OUT_RGB8, True).stats[_abi.STAT_SUGGESTED_EXPOSURE_NS] == 0.0


def test_black_frame():
    """pipeline_test.cpp:687-703: black in, black out, gains kept, AE error at its bound 8."""
    config = cases.camera(_abi.RGGB, cases.FORMATS[2], 64, 48, exposure=1e-6)
    params = isp.synthetic_params(config)
    params.white_balance = _abi.WB_GRAY_WORLD
    r = OraclePipeline(params).process(isp.render_synthetic(config), _abi.OUT_RGB8, True)
    assert not r.image.any()
    assert r.stats[_abi.STAT_WB_GAIN_R] == np.float32(params.wb_gains[0])
    assert r.stats[_abi.STAT_MEAN_LUMA] == 0.0 and r.stats[_abi.STAT_AE_ERROR_EV] == 8.0
    assert r.histogram[0] == 64 * 48


def rebuild_histogram(image, out_format, width, height, bins):
    """The histograms from the written image alone (pipeline_test.cpp:239-265), with the
    10-bit codes of the 1024-bin extension for RGB16 / RGB32F."""
    top = bins - 1
    if out_format in (_abi.OUT_RGB8, _abi.OUT_RGBA8):
        ch = 3 if out_format == _abi.OUT_RGB8 else 4
        q = image.reshape(height, width, ch)[:, :, :3].astype(np.int64)
    elif out_format == _abi.OUT_RGB16:
        q = (image.view(np.uint16).reshape(height, width, 3).astype(np.int64) * top + 32767) // 65535
    else:
        v = np.clip(image.view(np.float32).reshape(height, width, 3), 0.0, 1.0)
        q = (v * np.float32(top) + np.float32(0.5)).astype(np.int64)
    luma = (2126 * q[:, :, 0] + 7152 * q[:, :, 1] + 722 * q[:, :, 2] + 5000) // 10000
    return np.concatenate([np.bincount(a.ravel(), minlength=bins) for a in (luma, q[:, :, 0], q[:, :, 1], q[:, :, 2])])


@pytest.mark.parametrize("out_format", [_abi.OUT_RGB8, _abi.OUT_RGBA8, _abi.OUT_RGB16, _abi.OUT_RGB32F])
def test_histograms_are_the_images(out_format):
    config = cases.camera(_abi.GBRG, cases.FORMATS[2], 201, 117)
    params = isp.synthetic_params(config, 9)
    params.sharpen = 0.5
    frame = isp.render_synthetic(config)
    for bins in ((256, 1024) if out_format in (_abi.OUT_RGB16, _abi.OUT_RGB32F) else (256,)):
        params.histogram_bins = bins
        r = OraclePipeline(params).process(frame, out_format, True)
        assert np.array_equal(r.histogram, rebuild_histogram(r.image, out_format, 201, 117, bins))


def test_tone_lut_matches_reference_sources():
    """tone_lut: the product's host table, the oracle's and (where built) the reference's agree."""
    o = oracle_lib.oracle()
    for curve, gamma in ((_abi.TONE_SRGB, 2.2), (_abi.TONE_GAMMA, 2.2), (_abi.TONE_GAMMA, 1.7),
                         (_abi.TONE_LINEAR, 2.2)):
        params = isp.IspParams(tone_curve=curve, gamma=gamma)
        mine = isp.tone_lut(params)
        c, keep = params.to_c()
        buf = np.zeros(4097, dtype=np.float32)
        assert o.oracle_tone_lut(C.byref(c), buf.ctypes.data_as(C.POINTER(C.c_float)), 4097) == 4097
        assert np.array_equal(mine.view(np.uint32), buf.view(np.uint32))
        if oracle_lib.ref_available():
            ref = np.zeros(4097, dtype=np.float32)
            oracle_lib.ref().ref_tone_lut(C.byref(c), ref.ctypes.data_as(C.POINTER(C.c_float)), 4097)
            assert np.array_equal(mine.view(np.uint32), ref.view(np.uint32))
