"""The cases of the golden fixtures (tests/golden/golden.json): inputs and
settings only. The expected outputs come from the reference's compiled CPU
backend (tests/golden/make_golden.py)."""
import numpy as np

import holoscan_isp as isp
from holoscan_isp import _abi

import cases as test_cases

EXPOSURE = (10_000_000, 2.0)


def _flat(width, height, r, g, b, pattern):
    cfa = [[0, 1, 2, 3], [3, 2, 1, 0], [1, 0, 3, 2], [2, 3, 0, 1]][pattern]
    img = np.zeros((height, width), dtype=np.uint16)
    for y in range(height):
        for x in range(width):
            ch = cfa[((y & 1) << 1) | (x & 1)]
            img[y, x] = r if ch == 0 else b if ch == 3 else g
    return isp.RawFrame(img.view(np.uint8).ravel(), width, height, 2 * width, _abi.RAW_U16, 16, pattern)


def cases():
    """(name, frame, params, out_format, exposure) for every golden case."""
    out = []
    for name, config, out_format, tweak in test_cases.matches_cpu_cases():
        params = isp.synthetic_params(config, 9)
        tweak(params)
        out.append(("reference/" + name, isp.render_synthetic(config), params, out_format, EXPOSURE))
    for which in ("config1", "config2", "config5"):
        scene, params, out_format = getattr(test_cases, which)(194, 110, seed=11)
        out.append(("scene/" + which, isp.render_scene(scene), params, out_format, EXPOSURE))
    for pattern in test_cases.PATTERNS:
        for method in (_abi.BILINEAR, _abi.MALVAR):
            params = isp.IspParams(tone_curve=_abi.TONE_LINEAR, white_level=65535.0, sharpen=1.0,
                                   demosaic=method)
            out.append((f"tiny/5x3/{pattern}/{method}", _flat(5, 3, 9000, 21000, 15000, pattern), params,
                        _abi.OUT_RGB8, EXPOSURE))
    # statistics (pipeline_test.cpp:568-618): flat mid-gray, one stop under the target
    level = int(0.09 * 65535.0 + 0.5)
    params = isp.IspParams(tone_curve=_abi.TONE_LINEAR, white_level=65535.0,
                           auto_exposure=isp.AutoExposureParams(target=0.18, damping=1.0,
                                                                max_exposure_ns=20_000_000))
    out.append(("statistics/flat", _flat(64, 48, level, level, level, _abi.RGGB), params, _abi.OUT_RGB8,
                (15_000_000, 1.0)))
    return out
