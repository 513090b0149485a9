"""Test and benchmark inputs: the synthetic camera renders, settings and truth
equal the reference's own (compiled from its sources, where built), and the
seeded noisy scene is deterministic and pinned by hashes."""
import ctypes as C
import hashlib

import numpy as np
import pytest

import holoscan_isp as isp
from holoscan_isp import _abi

import cases
import oracle_lib

needs_ref = pytest.mark.skipif(not oracle_lib.ref_available(), reason="reference backend not built here")


@needs_ref
@pytest.mark.parametrize("fi", range(len(cases.FORMATS)))
@pytest.mark.parametrize("pattern", cases.PATTERNS)
def test_chart_render_matches_reference(fi, pattern):
    r = oracle_lib.ref()
    for kind, frame_no, exposure in ((_abi.SCENE_CHART, 3, 1.0), (_abi.SCENE_GRAY, 17, 1.0),
                                     (_abi.SCENE_CHART, 40, 4.0)):
        c = cases.camera(pattern, cases.FORMATS[fi], 101, 37, kind=kind, frame=frame_no, exposure=exposure)
        mine = isp.render_synthetic(c)
        theirs = np.zeros_like(mine.data)
        assert r.ref_synthetic_render(C.byref(c), theirs.ctypes.data, mine.pitch) == 0
        assert np.array_equal(mine.data, theirs)
        p = isp.IspParams().to_c()[0]
        lsc = (C.c_float * 81)()
        assert r.ref_synthetic_params(C.byref(c), 9, C.byref(p), lsc) == 0
        assert isp.synthetic_params(c, 9) == isp.IspParams.from_c(p)
        truth = isp.synthetic_truth(c)
        for x, y in ((0, 0), (50, 20), (100, 36), (7, 30)):
            t = (C.c_float * 3)()
            r.ref_synthetic_truth(C.byref(c), x, y, t)
            assert list(t) == list(truth[y, x])


def test_scene_is_deterministic_and_pinned():
    sc = isp.scene_config(width=257, height=131, format=_abi.RAW_RAW12, bit_depth=12, pattern=_abi.GRBG,
                          black_level=240.0, seed=42)
    a = isp.render_scene(sc)
    b = isp.render_scene(sc)
    assert np.array_equal(a.data, b.data)
    sc2 = isp.scene_config(width=257, height=131, format=_abi.RAW_RAW12, bit_depth=12, pattern=_abi.GRBG,
                           black_level=240.0, seed=43)
    assert not np.array_equal(a.data, isp.render_scene(sc2).data)
    # unpacked samples stay within [0, 4095] and the black level shows in the dark
    sc3 = isp.scene_config(width=64, height=64, seed=1)
    u = isp.render_scene(sc3).data.view(np.uint16)
    assert u.max() <= 4095 and 200 < np.median(u) < 4095


def test_scene_lens_shading_inverts_the_falloff():
    sc = isp.scene_config(width=1921, height=1081)
    g = isp.scene_lens_shading(sc, 17, 13).gains.reshape(13, 17)
    assert abs(g[6, 8] - 1.0) < 1e-3  # centre
    assert abs(g[0, 0] - 1 / 0.6) < 1e-4 and abs(g[12, 16] - 1 / 0.6) < 1e-4  # corners: cos^4 = 0.6
