"""The frame checker's oracle (oracle_frame_check, the host path of the reference
demo's checker, check_op.cpp:44-72) without a GPU.

Pins: the reference demo's published ci run (applications/isp/README.md:89: 30
frames of 640x480 RGGB 12-bit, Malvar, manual white balance, chart: mean PSNR
30.16, worst 30.12, white balance gains 1.900,1.000,1.600, mean luma 0.275, AE
error -0.613 EV), reproduced end to end by the oracle ISP and the oracle
checker; and closed forms of the check itself."""
import numpy as np
import pytest

import holoscan_isp as isp
from holoscan_isp import _abi

import cases
import oracle_lib
from oracle_lib import OraclePipeline


def run_demo_on_oracle(frames, **camera_kw):
    """The demo's source -> IspOp -> checker chain on the host: per frame the PSNR, the
    histogram mismatches and the statistics."""
    out = []
    pipe = None
    for frame in range(frames):
        cam = cases.demo_camera(frame=frame, **camera_kw)
        if pipe is None:
            params = cases.demo_params(cam)
            pipe = OraclePipeline(params)
            lut = isp.tone_lut(params)
        res = pipe.process(isp.render_synthetic(cam), _abi.OUT_RGB8, True, cases.DEMO_EXPOSURE, threads=8)
        r = oracle_lib.frame_check(isp.synthetic_truth(cam), lut,
                                   res.image.reshape(cam.height, cam.width, 3), res.histogram, res.stats)
        out.append((oracle_lib.frame_psnr(r, cam.width, cam.height), r.histogram_mismatches, list(r.stats)))
    return out


def test_demo_ci_run_reproduces_the_published_numbers():
    run = run_demo_on_oracle(30)
    psnr = [p for p, _, _ in run]
    assert all(m == 0 for _, m, _ in run)
    assert f"{np.mean(psnr):.2f}" == "30.16"
    assert f"{min(psnr):.2f}" == "30.12"
    stats = run[-1][2]  # the report keeps the last frame's statistics (check_op.cpp:230-234)
    gains = ",".join(f"{stats[k]:.3f}" for k in (_abi.STAT_WB_GAIN_R, _abi.STAT_WB_GAIN_G, _abi.STAT_WB_GAIN_B))
    assert gains == "1.900,1.000,1.600"
    assert f"{stats[_abi.STAT_MEAN_LUMA]:.3f}" == "0.275"
    assert f"{stats[_abi.STAT_AE_ERROR_EV]:.3f}" == "-0.613"


def truth_codes(truth, lut):
    """truth_codes (frame_check.hpp:54-62) in numpy float32: the interpolated table, then
    round(255 v) half away from zero."""
    size = lut.size
    f = truth.astype(np.float32) * np.float32(size - 1)
    i = f.astype(np.int32)
    top = i >= size - 1
    j = np.minimum(i, size - 2)
    v = np.where(top, lut[size - 1], lut[j] + (f - i.astype(np.float32)) * (lut[j + 1] - lut[j]))
    return np.floor(255.0 * v.astype(np.float64) + 0.5).astype(np.int32)


def histograms(rgb):
    r, g, b = (rgb[..., k].astype(np.int64).ravel() for k in range(3))
    luma = (2126 * r + 7152 * g + 722 * b + 5000) // 10000
    return np.concatenate([np.bincount(c, minlength=256) for c in (luma, r, g, b)]).astype(np.uint32)


@pytest.fixture(scope="module")
def small_frame():
    cam = cases.demo_camera(width=97, height=41, frame=5)
    params = cases.demo_params(cam)
    return cam, isp.synthetic_truth(cam), isp.tone_lut(params)


def test_a_perfect_image_scores_zero_error(small_frame):
    cam, truth, lut = small_frame
    rgb = truth_codes(truth, lut).astype(np.uint8)
    stats = np.arange(12, dtype=np.float32)
    r = oracle_lib.frame_check(truth, lut, rgb, histograms(rgb), stats)
    assert r.squared_error == 0 and r.histogram_mismatches == 0
    assert oracle_lib.frame_psnr(r, cam.width, cam.height) == 99.0
    assert list(r.stats) == list(stats)


def test_errors_add_their_squares_and_bins_count_once(small_frame):
    cam, truth, lut = small_frame
    codes = truth_codes(truth, lut)
    rgb = codes.astype(np.uint8)
    rng = np.random.default_rng(3)
    for _ in range(40):
        y, x, k = rng.integers(cam.height), rng.integers(cam.width), rng.integers(3)
        rgb[y, x, k] = rng.integers(256)
    expected = int(((rgb.astype(np.int64) - codes) ** 2).sum())
    hist = histograms(rgb)
    r = oracle_lib.frame_check(truth, lut, rgb, hist)
    assert r.squared_error == expected and r.histogram_mismatches == 0
    assert oracle_lib.frame_psnr(r, cam.width, cam.height) == pytest.approx(
        10 * np.log10(255.0 ** 2 * 3 * cam.width * cam.height / expected))
    hist[3 * 256 + 7] += 1
    hist[5] += 2
    assert oracle_lib.frame_check(truth, lut, rgb, hist).histogram_mismatches == 2
    assert list(oracle_lib.frame_check(truth, lut, rgb, hist).stats) == [0.0] * 12


def test_the_luma_row_uses_rec709_weights_rounded():
    truth = np.zeros((1, 3, 3), dtype=np.float32)
    lut = np.linspace(0, 1, 4097, dtype=np.float32)
    rgb = np.array([[[255, 0, 0], [0, 255, 0], [10, 20, 30]]], dtype=np.uint8)
    hist = np.zeros(1024, dtype=np.uint32)
    for q in rgb[0]:
        r, g, b = (int(v) for v in q)
        # 0.2126 R + 0.7152 G + 0.0722 B, rounded half up
        luma = int(np.floor(0.2126 * r + 0.7152 * g + 0.0722 * b + 0.5 + 1e-9))
        for row, v in enumerate((luma, r, g, b)):
            hist[row * 256 + v] += 1
    assert oracle_lib.frame_check(truth, lut, rgb, hist).histogram_mismatches == 0
