"""The kernels' gfx950 assembly keeps the reference's float semantics: no
contracted (fused) single-precision operation outside the IEEE division
expansion (v_div_scale .. v_div_fixup), and no packed fp32 arithmetic."""
import os
import re
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "holoscan-zozo-modules_amd")

FUSED = re.compile(r"\bv_(fma|fmac|mac|mad|fma_mix|fma_mixlo|fma_mixhi|pk_fma)_f32\b|\bv_fma_mix")
PACKED = re.compile(r"\bv_pk_(add|mul|fma)_f32\b")


def test_no_contraction_in_kernels():
    subprocess.run(["make", "-s", "-C", PKG, "asm"], check=True)
    asm_files = [os.path.join(PKG, "_build", f) for f in os.listdir(os.path.join(PKG, "_build")) if f.endswith(".s")]
    assert asm_files
    for path in asm_files:
        lines = open(path).read().splitlines()
        for i, line in enumerate(lines):
            assert not PACKED.search(line), f"{os.path.basename(path)}:{i + 1}: {line.strip()}"
            if FUSED.search(line):
                window = "\n".join(lines[max(0, i - 12):i + 12])
                assert "v_div_scale_f32" in window and "v_div_fixup_f32" in window, \
                    f"{os.path.basename(path)}:{i + 1}: fused op outside a division: {line.strip()}"


def test_kernels_make_no_calls_and_use_no_scratch():
    """Every kernel is one function (no calls: a call spills the caller's state to
    private memory) with no private segment (no register spills)."""
    subprocess.run(["make", "-s", "-C", PKG, "asm"], check=True)
    text = open(os.path.join(PKG, "_build", "isp_kernels.s")).read()
    assert "s_swappc_b64" not in text  # (s_setpc_b64 alone is a long branch inside a large kernel)
    kernels = re.findall(r"\.amdhsa_kernel (\S+)\n(.*?)\.end_amdhsa_kernel", text, re.S)
    assert len(kernels) >= 20
    for name, body in kernels:
        scratch = int(re.search(r"\.amdhsa_private_segment_fixed_size (\d+)", body).group(1))
        assert scratch == 0, f"{name}: {scratch} bytes of scratch"
