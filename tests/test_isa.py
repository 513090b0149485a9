"""The kernels' gfx950 assembly keeps the reference's float semantics: no
contracted (fused) single-precision operation outside the IEEE division
expansion (v_div_scale .. v_div_fixup), and no packed fp32 arithmetic."""
import os
import re
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "holoscan-zozo-modules_amd")

_BUILT = {}
FILES = ("isp_kernels.s", "isp_source.s", "isp_color.s")


def asm_text(name="isp_kernels.s"):
    """A kernel file's gfx950 assembly; all are built once per test session."""
    if not _BUILT:
        subprocess.run(["make", "-s", "-C", PKG, "asm"], check=True)
        for f in FILES:
            _BUILT[f] = open(os.path.join(PKG, "_build", f)).read()
    return _BUILT[name]


FUSED = re.compile(r"\bv_(fma|fmac|mac|mad|fma_mix|fma_mixlo|fma_mixhi|pk_fma)_f32\b|\bv_fma_mix")
PACKED = re.compile(r"\bv_pk_(add|mul|fma)_f32\b")


def in_division(lines, i):
    """A fused op that belongs to a division: the IEEE float division expansion (v_div_scale
    before, v_div_fmas / v_div_fixup after), or the exact remainder step of the compiler's
    small-integer division through a float reciprocal (v_rcp_iflag)."""
    before = "\n".join(lines[max(0, i - 16):i + 1])
    after = "\n".join(lines[i:i + 24])
    if "v_div_scale_f32" in before and ("v_div_fmas_f32" in after or "v_div_fixup_f32" in after):
        return True
    return "v_rcp_iflag_f32" in "\n".join(lines[max(0, i - 8):i])


def test_no_contraction_in_kernels():
    for name in FILES:
        lines = asm_text(name).splitlines()
        for i, line in enumerate(lines):
            assert not PACKED.search(line), f"{name}:{i + 1}: {line.strip()}"
            if FUSED.search(line):
                assert in_division(lines, i), f"{name}:{i + 1}: fused op outside a division: {line.strip()}"


def test_kernels_make_no_calls_and_use_no_scratch():
    """Every kernel is one function (no calls: a call spills the caller's state to
    private memory) with no private segment (no register spills), but for the
    512-thread batch kernels' few bytes."""
    kernels = []
    for f in FILES:
        text = asm_text(f)
        assert "s_swappc_b64" not in text, f  # (s_setpc_b64 alone is a long branch inside a large kernel)
        kernels += re.findall(r"\.amdhsa_kernel (\S+)\n(.*?)\.end_amdhsa_kernel", text, re.S)
    assert len(kernels) >= 24
    for name, body in kernels:
        scratch = int(re.search(r"\.amdhsa_private_segment_fixed_size (\d+)", body).group(1))
        # The 512-thread batch kernels hold every tile variant in one function at 128
        # registers; a few per-thread constants spill once per launch (at most 64 bytes).
        limit = 64 if "_wide" in name else 0
        assert scratch <= limit, f"{name}: {scratch} bytes of scratch"
