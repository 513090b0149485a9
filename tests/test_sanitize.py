"""The host code under sanitizers: tests/sanitize/host_check.cpp (settings
validation, tone tables, the chart camera and noisy scene renderers on threads,
the statistics epilogue's host evaluation, and the CPU oracle on threads over
every format, layout, demosaic, white balance and output at small and odd sizes)
built with g++ under AddressSanitizer + UndefinedBehaviorSanitizer and under
ThreadSanitizer, run, and required to finish without a report. This is the host
half of the reference's sanitizer tier (compute-sanitizer memcheck, racecheck,
initcheck and synccheck over its pipeline test, modules/isp/tests/CMakeLists.txt:37-50)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sanitize")

pytestmark = pytest.mark.skipif(shutil.which("g++") is None or not os.path.isdir(HERE),
                                reason="needs g++ and tests/sanitize")


@pytest.fixture(scope="module")
def built():
    subprocess.run(["make", "-j", "4", "-C", HERE], check=True, capture_output=True, timeout=600)
    return os.path.join(HERE, "_build")


def symbols(path):
    out = subprocess.run(["nm", path], check=True, capture_output=True, text=True).stdout
    return {line.split()[-1] for line in out.splitlines() if line.strip()}


@pytest.mark.parametrize("kind,runtime", [("asan", "__asan_init"), ("tsan", "__tsan_init")])
def test_host_code_is_clean_under_sanitizers(built, kind, runtime):
    exe = os.path.join(built, f"host_check_{kind}")
    # The binary really is instrumented.
    assert any(s.startswith(runtime) for s in symbols(exe)), f"{exe} lacks {runtime}"
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1", TSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600, env=env)
    report = r.stderr[-4000:]
    assert r.returncode == 0, f"host_check_{kind} exited {r.returncode}:\n{report}"
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, report
    assert "WARNING: ThreadSanitizer" not in r.stderr, report
    assert "0 failures" in r.stdout, r.stdout + report


def test_a_planted_overrun_is_reported(tmp_path):
    """The same flags catch a planted heap overrun, so a clean run above means something."""
    src = tmp_path / "overrun.cpp"
    src.write_text("#include <vector>\nint main(int argc, char**) {\n  std::vector<int> v(4);\n"
                   "  return v.data()[3 + argc];\n}\n")
    exe = tmp_path / "overrun"
    subprocess.run(["g++", "-g", "-O1", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                    str(src), "-o", str(exe)], check=True, timeout=120)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "heap-buffer-overflow" in r.stderr
