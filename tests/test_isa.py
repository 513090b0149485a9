"""The kernels' gfx950 assembly keeps the reference's float semantics: no
contracted (fused) single-precision operation outside the IEEE division
expansion (v_div_scale .. v_div_fixup), and no packed fp32 arithmetic."""
import os
import re
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "holoscan-zozo-modules_amd")

_BUILT = []


def asm_text():
    """The kernels' gfx950 assembly, built once per test session."""
    if not _BUILT:
        subprocess.run(["make", "-s", "-C", PKG, "asm"], check=True)
        _BUILT.append(open(os.path.join(PKG, "_build", "isp_kernels.s")).read())
    return _BUILT[0]


FUSED = re.compile(r"\bv_(fma|fmac|mac|mad|fma_mix|fma_mixlo|fma_mixhi|pk_fma)_f32\b|\bv_fma_mix")
PACKED = re.compile(r"\bv_pk_(add|mul|fma)_f32\b")


def test_no_contraction_in_kernels():
    lines = asm_text().splitlines()
    for i, line in enumerate(lines):
        assert not PACKED.search(line), f"isp_kernels.s:{i + 1}: {line.strip()}"
        if FUSED.search(line):
            window = "\n".join(lines[max(0, i - 12):i + 12])
            assert "v_div_scale_f32" in window and "v_div_fixup_f32" in window, \
                f"isp_kernels.s:{i + 1}: fused op outside a division: {line.strip()}"


def test_kernels_make_no_calls_and_use_no_scratch():
    """Every kernel is one function (no calls: a call spills the caller's state to
    private memory) with no private segment (no register spills)."""
    text = asm_text()
    assert "s_swappc_b64" not in text  # (s_setpc_b64 alone is a long branch inside a large kernel)
    kernels = re.findall(r"\.amdhsa_kernel (\S+)\n(.*?)\.end_amdhsa_kernel", text, re.S)
    assert len(kernels) >= 20
    for name, body in kernels:
        scratch = int(re.search(r"\.amdhsa_private_segment_fixed_size (\d+)", body).group(1))
        assert scratch == 0, f"{name}: {scratch} bytes of scratch"
