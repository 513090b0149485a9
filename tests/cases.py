"""Shared test cases: synthetic-camera frames and settings, as the reference's
GPU suite builds them (modules/isp/tests/pipeline_test.cpp:45-88, :164-235),
plus the seeded noisy scenes of the benchmark configurations."""
import numpy as np

import holoscan_isp as isp
from holoscan_isp import _abi

# SampleFormat table (pipeline_test.cpp:50-57): format, bits, black.
FORMATS = [
    (_abi.RAW_U8, 8, 16.0),
    (_abi.RAW_U16, 10, 64.0),
    (_abi.RAW_U16, 12, 256.0),
    (_abi.RAW_U16, 16, 4096.0),
    (_abi.RAW_RAW10, 10, 64.0),
    (_abi.RAW_RAW12, 12, 256.0),
]
PATTERNS = [_abi.RGGB, _abi.BGGR, _abi.GRBG, _abi.GBRG]


def camera(pattern, fmt, width=480, height=360, **kw):
    """camera() (pipeline_test.cpp:75-88): per-channel black levels, WB gains, CCM, vignetting 0.4, frame 3."""
    f, bits, black = fmt
    step = black / 16.0
    cfg = dict(width=width, height=height, pattern=pattern, format=f, bit_depth=bits,
               black_levels=(black, black + step, black + 2 * step, black + 3 * step),
               vignetting=0.4, frame=3)
    cfg.update(kw)
    return isp.synthetic_config(**cfg)


def four_channel(params):
    """The 4-channel lens shading variant of cuda_matches_cpu (pipeline_test.cpp:200-206)."""
    s = params.lens_shading
    four = []
    for g in s.gains:
        for c in range(4):
            factor = _f32(1.0) + _f32(0.01) * _f32(c)  # float arithmetic, as 1.0F + 0.01F * c
            four.append(float(_f32(g) * factor))
    params.lens_shading = isp.LensShading(s.width, s.height, 4, four)
    return params


def _f32(v):
    import numpy as np
    return np.float32(v)


def matches_cpu_cases():
    """The cases of cuda_matches_cpu (pipeline_test.cpp:164-208): (name, config, out_format, tweak)."""
    u12 = FORMATS[2]
    cases = []
    for method in (_abi.BILINEAR, _abi.MALVAR):
        m = "bilinear" if method == _abi.BILINEAR else "malvar"

        def plain(p, method=method):
            p.demosaic = method

        def sharpened(p, method=method):
            p.demosaic = method
            p.sharpen = 1.5

        def odd(p, method=method):
            p.demosaic = method
            p.sharpen = 0.5

        def gray(p, method=method):
            p.demosaic = method
            p.white_balance = _abi.WB_GRAY_WORLD
            p.tone_curve = _abi.TONE_GAMMA

        def patch(p, method=method):
            p.demosaic = method
            p.white_balance = _abi.WB_WHITE_PATCH
            p.tone_curve = _abi.TONE_LUT
            p.lut = [0.0, 0.6, 0.85, 1.0]
            four_channel(p)

        cases.append((m, camera(_abi.RGGB, u12, 200, 150), _abi.OUT_RGB8, plain))
        cases.append((m + " sharpened", camera(_abi.GBRG, u12, 200, 150), _abi.OUT_RGBA8, sharpened))
        cases.append((m + " odd size", camera(_abi.GRBG, FORMATS[4], 67, 35), _abi.OUT_RGB16, odd))
        cases.append((m + " gray world", camera(_abi.BGGR, FORMATS[5], 160, 120), _abi.OUT_RGB32F, gray))
        cases.append((m + " white patch, 4-channel shading, table", camera(_abi.RGGB, FORMATS[0], 160, 120),
                      _abi.OUT_RGB8, patch))
    return cases


def float_edge_cases():
    """Frames whose float path reaches the edges of IEEE arithmetic, with settings validation
    accepts: (name, frame, params, out_format).
    - subnormal samples: white balance gains of about 1e-39 (validation asks only > 0) make
      every linear sample subnormal, through demosaic, sharpening and the colour matrix;
    - signed zeros: a random mosaic with many samples below black, a colour matrix whose
      last row is all negative (so black pixels give -0 there, which clamp01 keeps), and a
      user tone table whose first entry is -0.0 (validation accepts it as 0), so a -0 can
      reach the RGB32F output;
    - the table's top entry: x = 1 on saturated pixels, through a table whose last entry is
      -0.0."""
    out = []
    cfg = camera(_abi.GRBG, FORMATS[2], 67, 37)
    frame = isp.render_synthetic(cfg)
    p = isp.synthetic_params(cfg, 5)
    p.wb_gains = (3e-39, 1e-39, 2e-39)
    p.demosaic, p.sharpen = _abi.MALVAR, 0.7
    p.ccm = (0.9, 0.2, -0.1, 0.05, 0.8, 0.15, -0.2, 0.1, 1.1)
    p.tone_curve = _abi.TONE_LINEAR
    out.append(("subnormal_malvar", frame, p, _abi.OUT_RGB32F))
    q = isp.synthetic_params(cfg, 5)
    q.wb_gains = (5e-39, 2e-39, 1e-39)
    q.demosaic, q.sharpen = _abi.BILINEAR, 2.0
    q.tone_curve = _abi.TONE_GAMMA
    out.append(("subnormal_bilinear", frame, q, _abi.OUT_RGB16))
    rng = np.random.default_rng(77)
    img = rng.integers(0, 600, size=(29, 45), dtype=np.uint16)
    mosaic = isp.RawFrame(img.view(np.uint8).ravel(), 45, 29, 90, _abi.RAW_U16, 12, _abi.RGGB)
    lut = np.linspace(0.0, 1.0, 257, dtype=np.float32)
    lut[0] = np.float32(-0.0)
    for method, sharpen in ((_abi.BILINEAR, 0.0), (_abi.MALVAR, 0.0), (_abi.MALVAR, 1.5)):
        r = isp.IspParams(black_level=400.0, wb_gains=(1.0, 1.0, 1.0), demosaic=method, sharpen=sharpen,
                          ccm=(1.0, 0.0, 0.0, 0.0, 1.0, 0.0, -0.25, -0.5, -0.25),
                          tone_curve=_abi.TONE_LUT, lut=[float(v) for v in lut])
        out.append((f"signed_zero_{method}_{sharpen}", mosaic, r, _abi.OUT_RGB32F))
    # The table's top entry: a decreasing table ending in -0.0, reached at x = 1 by a frame
    # with a saturated band (the reference returns lut[n - 1] itself there).
    bright = rng.integers(0, 4096, size=(29, 45), dtype=np.uint16)
    bright[8:20, :] = 4095
    top = isp.RawFrame(bright.view(np.uint8).ravel(), 45, 29, 90, _abi.RAW_U16, 12, _abi.GBRG)
    falling = np.linspace(1.0, 0.0, 4097, dtype=np.float32)
    falling[-1] = np.float32(-0.0)
    for method, sharpen in ((_abi.BILINEAR, 0.0), (_abi.MALVAR, 0.8)):
        r = isp.IspParams(black_level=64.0, wb_gains=(1.0, 1.0, 1.0), demosaic=method, sharpen=sharpen,
                          tone_curve=_abi.TONE_LUT, lut=[float(v) for v in falling])
        out.append((f"table_top_{method}_{sharpen}", top, r, _abi.OUT_RGB32F))
    return out


def baseline_scene(width, height, fmt, bits, pattern, black, seed, **kw):
    return isp.scene_config(width=width, height=height, format=fmt, bit_depth=bits, pattern=pattern,
                            black_level=black, seed=seed, **kw)


BASELINE_CCM = (1.6, -0.4, -0.2, -0.3, 1.5, -0.2, -0.1, -0.5, 1.6)


def config1(width=1920, height=1080, seed=1):
    """1920x1080 RGGB u16 12-bit, BL 256, 17x13 LSC = 1/falloff, manual WB (2, 1, 1.6),
    bilinear, no sharpen, CCM, gamma 2.2, RGB8 + histograms."""
    sc = baseline_scene(width, height, _abi.RAW_U16, 12, _abi.RGGB, 256.0, seed)
    p = isp.IspParams(black_level=256.0, lens_shading=isp.scene_lens_shading(sc, 17, 13),
                      white_balance=_abi.WB_MANUAL, wb_gains=(2.0, 1.0, 1.6), demosaic=_abi.BILINEAR,
                      ccm=BASELINE_CCM, tone_curve=_abi.TONE_GAMMA, gamma=2.2)
    return sc, p, _abi.OUT_RGB8


def config2(width=3840, height=2160, seed=2):
    """3840x2160 GRBG RAW12, BL 240, 17x13 LSC, gray world, Malvar, sharpen 0.5, CCM, sRGB,
    RGBA8, histograms + AE."""
    sc = baseline_scene(width, height, _abi.RAW_RAW12, 12, _abi.GRBG, 240.0, seed)
    p = isp.IspParams(black_level=240.0, lens_shading=isp.scene_lens_shading(sc, 17, 13),
                      white_balance=_abi.WB_GRAY_WORLD, demosaic=_abi.MALVAR, sharpen=0.5,
                      ccm=BASELINE_CCM, tone_curve=_abi.TONE_SRGB)
    return sc, p, _abi.OUT_RGBA8


def srgb16_lut():
    """The 16-bit tone curve: a 65536-entry sRGB table (ToneCurve::kLut)."""
    import numpy as np
    x = np.arange(65536, dtype=np.float64) / 65535.0
    y = np.where(x <= 0.0031308, 12.92 * x, 1.055 * np.power(x, 1.0 / 2.4) - 0.055)
    return np.clip(y, 0.0, 1.0).astype(np.float32)


def config5(width=8192, height=6144, seed=5):
    """8192x6144 RGGB u16 16-bit, BL 1024, white patch, Malvar, sharpen, CCM, 16-bit tone
    curve, RGB16 + 1024-bin histograms."""
    sc = baseline_scene(width, height, _abi.RAW_U16, 16, _abi.RGGB, 1024.0, seed)
    p = isp.IspParams(black_level=1024.0, lens_shading=isp.scene_lens_shading(sc, 17, 13),
                      white_balance=_abi.WB_WHITE_PATCH, demosaic=_abi.MALVAR, sharpen=0.5,
                      ccm=BASELINE_CCM, tone_curve=_abi.TONE_LUT, lut=srgb16_lut(),
                      histogram_bins=1024)
    return sc, p, _abi.OUT_RGB16


def stream_ccm(i):
    """A per-stream colour matrix with rows summing to 1: the benchmark CCM with stream-dependent
    saturation."""
    s = 1.0 + 0.05 * (i % 5)
    base = BASELINE_CCM
    rows = []
    for r in range(3):
        row = [base[3 * r + k] * s for k in range(3)]
        row[r] += 1.0 - sum(row)
        rows.extend(row)
    return tuple(rows)


def config3(divisor=1, seed=300):
    """The 16-stream mixed batch: 4 x 1280x720 BGGR RAW10, 8 x 1920x1200 RGGB RAW12, 4 x 3840x2160
    GBRG u16 at 14 bits; per-stream seed, black level, 17x13 lens shading and CCM; bilinear and
    Malvar alternating by stream index; gray-world white balance; RGB8 + histograms + stats.
    `divisor` scales the frame sizes down (even sizes kept) for checks against the oracle.
    Returns [(scene, params, out_format)]."""
    groups = [(4, 1280, 720, _abi.RAW_RAW10, 10, _abi.BGGR, 64.0, 4.0),
              (8, 1920, 1200, _abi.RAW_RAW12, 12, _abi.RGGB, 240.0, 8.0),
              (4, 3840, 2160, _abi.RAW_U16, 14, _abi.GBRG, 1024.0, 32.0)]
    out = []
    i = 0
    for count, w, h, fmt, bits, pattern, black, step in groups:
        for j in range(count):
            width = max(2, (w // divisor) & ~1)
            height = max(2, (h // divisor) & ~1)
            bl = black + step * j
            sc = baseline_scene(width, height, fmt, bits, pattern, bl, seed + i)
            p = isp.IspParams(black_level=bl, lens_shading=isp.scene_lens_shading(sc, 17, 13),
                              white_balance=_abi.WB_GRAY_WORLD,
                              demosaic=_abi.BILINEAR if i % 2 == 0 else _abi.MALVAR,
                              ccm=stream_ccm(i), tone_curve=_abi.TONE_SRGB)
            out.append((sc, p, _abi.OUT_RGB8))
            i += 1
    return out


# Lens shading grids past the 1,024 values a block stages in LDS (validation accepts up to
# 256 x 256 x 4, isp_params.cpp:102-112): the kernels then read the grid from device memory,
# through the tile's table where a tile's rows span at most two grid rows and per sample where
# they span more. Rows: grid width, height, channels; frame width, height; raw format, bits,
# pattern; demosaic, sharpening, white balance.
LARGE_GRIDS = [
    (17, 17, 4, 320, 240, _abi.RAW_RAW12, 12, _abi.GRBG, _abi.MALVAR, 0.5, _abi.WB_GRAY_WORLD),
    (41, 25, 1, 300, 200, _abi.RAW_RAW10, 10, _abi.RGGB, _abi.BILINEAR, 0.0, _abi.WB_WHITE_PATCH),
    (32, 32, 1, 300, 200, _abi.RAW_U16, 12, _abi.BGGR, _abi.MALVAR, 0.0, _abi.WB_GRAY_WORLD),
    (64, 48, 4, 400, 300, _abi.RAW_U16, 14, _abi.GBRG, _abi.MALVAR, 0.7, _abi.WB_WHITE_PATCH),
    (64, 64, 4, 1920, 1080, _abi.RAW_RAW10, 10, _abi.BGGR, _abi.BILINEAR, 0.0, _abi.WB_GRAY_WORLD),
    (256, 256, 1, 640, 480, _abi.RAW_RAW12, 12, _abi.RGGB, _abi.MALVAR, 0.5, _abi.WB_GRAY_WORLD),
    (256, 256, 4, 512, 256, _abi.RAW_U8, 8, _abi.RGGB, _abi.MALVAR, 1.0, _abi.WB_WHITE_PATCH),
]
BLACK_OF_BITS = {8: 16.0, 10: 64.0, 12: 240.0, 14: 1024.0, 16: 4096.0}


def large_grid(row, seed):
    """(scene, params, out_format) of a LARGE_GRIDS row: seeded gains uniform in [0.5, 2.5)."""
    gw, gh, ch, w, h, fmt, bits, pattern, dem, sharpen, wb = row
    black = BLACK_OF_BITS[bits]
    sc = baseline_scene(w, h, fmt, bits, pattern, black, seed)
    gains = np.random.default_rng(seed).uniform(0.5, 2.5, gw * gh * ch).astype(np.float32)
    p = isp.IspParams(black_level=black, lens_shading=isp.LensShading(gw, gh, ch, [float(g) for g in gains]),
                      white_balance=wb, wb_gains=(1.7, 1.0, 1.4), demosaic=dem, sharpen=sharpen, ccm=BASELINE_CCM,
                      tone_curve=_abi.TONE_SRGB)
    return sc, p, _abi.OUT_RGBA8


CONFIG4_SIZES = [(1280, 720), (1920, 1080), (1920, 1200), (3840, 2160)]


def config4(cameras=32, seed=400):
    """The served mix: `cameras` cameras at resolutions drawn uniformly from 720p, 1080p, 1920x1200
    and 4K, RAW10 or RAW12 and a Bayer pattern per camera (seeded), white patch white balance at the
    0.99 percentile, bilinear / Malvar alternating, 17x13 lens shading, per-camera CCM, sRGB, RGB8
    with histograms and statistics. Returns [(scene, params, out_format)]."""
    import random
    rng = random.Random(seed)
    out = []
    for i in range(cameras):
        w, h = CONFIG4_SIZES[rng.randrange(4)]
        fmt = _abi.RAW_RAW10 if rng.randrange(2) == 0 else _abi.RAW_RAW12
        bits = 10 if fmt == _abi.RAW_RAW10 else 12
        pattern = rng.randrange(4)
        black = 64.0 if bits == 10 else 240.0
        sc = baseline_scene(w, h, fmt, bits, pattern, black, seed + i)
        p = isp.IspParams(black_level=black, lens_shading=isp.scene_lens_shading(sc, 17, 13),
                          white_balance=_abi.WB_WHITE_PATCH, white_patch_percentile=0.99,
                          demosaic=_abi.BILINEAR if i % 2 == 0 else _abi.MALVAR,
                          ccm=stream_ccm(i), tone_curve=_abi.TONE_SRGB)
        out.append((sc, p, _abi.OUT_RGB8))
    return out


def poisson_arrivals(cameras, fps, duration, seed=401):
    """Per camera, arrival times (s) of a Poisson process at `fps` over [0, duration); returns
    [(t, camera, k)] sorted by time, k the camera's frame index."""
    import random
    rng = random.Random(seed)
    events = []
    for c in range(cameras):
        t = rng.expovariate(fps)
        k = 0
        while t < duration:
            events.append((t, c, k))
            k += 1
            t += rng.expovariate(fps)
    events.sort()
    return events


# ---- the reference demo's camera (applications/isp/src/main.cpp) -------------

def demo_camera(width=640, height=480, pattern=_abi.RGGB, bits=12, packing="none", scene=_abi.SCENE_CHART,
                frame=0):
    """camera_of (main.cpp:214-228) after the option rules of main.cpp:177-180: packed RAW10/RAW12
    force 10/12 bits, 8 bits or fewer travel as uint8; black about 6 % of full scale, a little
    higher on the second and fourth channel; vignetting 0.4."""
    import math
    fmt = {"raw10": _abi.RAW_RAW10, "raw12": _abi.RAW_RAW12}.get(packing, _abi.RAW_U16)
    if fmt == _abi.RAW_RAW10:
        bits = 10
    if fmt == _abi.RAW_RAW12:
        bits = 12
    if fmt == _abi.RAW_U16 and bits <= 8:
        fmt = _abi.RAW_U8
    black = float(math.floor(float(np.float32(0.0625) * np.float32((1 << bits) - 1)) + 0.5))  # std::round, >= 0
    step = float(math.floor(black / 32.0))
    return isp.synthetic_config(width=width, height=height, pattern=pattern, format=fmt, bit_depth=bits,
                                black_levels=(black, black + step, black, black + 2.0 * step), vignetting=0.4,
                                kind=scene, frame=frame)


def demo_params(camera, demosaic=_abi.MALVAR, white_balance=_abi.WB_MANUAL, sharpen=0.0):
    """The demo's IspOp params (main.cpp:254-261): the camera's settings, the chosen demosaic,
    white balance and sharpening; an estimating white balance starts from neutral gains."""
    p = isp.synthetic_params(camera, 17)
    p.demosaic = demosaic
    p.white_balance = white_balance
    p.sharpen = sharpen
    if white_balance != _abi.WB_MANUAL:
        p.wb_gains = (1.0, 1.0, 1.0)
    return p


def demo_min_psnr(bits, demosaic):
    """The bound every frame must beat (main.cpp:397-400)."""
    return 24.0 if bits <= 8 else 28.0 if demosaic == _abi.MALVAR else 26.0


# The demo's exposure descriptor (raw_source_op.hpp:28-29).
DEMO_EXPOSURE = (10_000_000, 1.0)
