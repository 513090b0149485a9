"""The C ABI without a GPU: the library loads, exports every function
include/*.h declares, and the ctypes mirror's struct layouts equal the C
compiler's."""
import os
import re
import subprocess

import ctypes as C

import holoscan_isp as isp
from holoscan_isp import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", h) for h in sorted(os.listdir(os.path.join(ROOT, "include")))
           if h.endswith(".h")]


def declared_functions():
    names = set()
    for path in HEADERS:
        text = open(path).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b(isp_\w+)\s*\(", text, flags=re.M):
            if not text[m.start():m.end()].lstrip().startswith(("typedef", "#")):
                names.add(m.group(1))
    return names


def test_library_loads_and_exports_every_declared_function():
    lib = isp.lib()
    names = declared_functions()
    assert len(names) > 30
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    bound = {p[0] for p in _abi.PROTOTYPES}
    assert names <= bound, sorted(names - bound)


STRUCTS = {"isp_params": _abi.Params, "isp_lens_shading": _abi.LensShading,
           "isp_auto_exposure": _abi.AutoExposure, "isp_raw_view": _abi.RawView, "isp_rgb_view": _abi.RgbView,
           "isp_stats_out": _abi.StatsOut, "isp_frame_exposure": _abi.FrameExposure,
           "isp_synthetic_config": _abi.SyntheticConfig, "isp_scene_config": _abi.SceneConfig,
           "isp_batch_frame": _abi.BatchFrame, "isp_batch_schedule": _abi.BatchSchedule,
           "isp_server_config": _abi.ServerConfig, "isp_camera_config": _abi.CameraConfig,
           "isp_server_result": _abi.ServerResult, "isp_server_counters": _abi.ServerCounters,
           "isp_tensor": _abi.Tensor, "isp_image": _abi.Image, "isp_sample_metadata": _abi.SampleMetadata,
           "isp_op_config": _abi.OpConfig, "isp_op_port": _abi.OpPort, "isp_op_statistics_t": _abi.OpStatistics,
           "isp_op_runtime": _abi.OpRuntime, "isp_frame_check_result": _abi.FrameCheckResult,
           "isp_epilogue_case": _abi.EpilogueCase}


def test_struct_layouts_match_the_c_compiler(tmp_path):
    src = ['#include <stdio.h>', '#include <stddef.h>', '#include "isp_synthetic.h"', '#include "isp_op.h"',
           '#include "isp_source.h"',
           'int main(void) {']
    for cname, py in STRUCTS.items():
        src.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for field, _ in py._fields_:
            src.append(f'printf("{cname}.{field} %zu\\n", offsetof({cname}, {field}));')
    src.append("return 0; }")
    (tmp_path / "layout.c").write_text("\n".join(src))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(tmp_path / "layout.c"), "-o",
                    str(exe)], check=True)
    got = dict(line.split() for line in subprocess.run([str(exe)], check=True, capture_output=True,
                                                        text=True).stdout.splitlines())
    for cname, py in STRUCTS.items():
        assert int(got[cname]) == C.sizeof(py), cname
        for field, _ in py._fields_:
            assert int(got[f"{cname}.{field}"]) == getattr(py, field).offset, (cname, field)


def test_batch_schedule_defaults_without_a_device():
    s = isp.default_schedule()
    assert [list(r) for r in s.tile_height] == [[32, 32]] * 4
    assert [list(r) for r in s.tiles_per_item] == [[4, 4]] * 4
    assert (s.awb_tiles_per_item, s.blocks_per_cu, s.order, s.queue, s.graph, s.per_frame) == (8, 0, 0, 0, 1, 0)


def test_batch_create_without_a_device_fails_cleanly():
    if isp.lib().isp_probe_platform():
        return
    h = C.c_void_p()
    assert isp.lib().isp_batch_create(C.byref(h)) == _abi.FAILURE
    assert b"hip" in isp.lib().isp_last_error()


def test_server_config_defaults_without_a_device():
    c = isp.server_config()
    assert (c.hbm_budget, c.max_batch, c.launch_slots, c.per_frame) == (256 << 20, 32, 2, 0)
    assert c.schedule.graph == 0 and c.schedule.tiles_per_item[3][1] == 4
