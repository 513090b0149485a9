"""The kernels' gfx950 assembly keeps the reference's float semantics: no
contracted (fused) single-precision operation outside the IEEE division
expansion (v_div_scale .. v_div_fixup) but those whose product is exact for
every float the kernels see (one factor a power of two of magnitude 1 or more,
which the kernels fold into a sum on purpose: scaling up is exact short of
overflow, subnormals included, while halving a subnormal with its last bit set
rounds), and no packed fp32 arithmetic."""
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


FUSED = re.compile(r"\bv_(fma|fmac|mac|mad|fmamk|fmaak|madmk|madak|fma_mix|fma_mixlo|fma_mixhi|pk_fma)_f32\b|\bv_fma_mix")
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


# Inline constants that are powers of two of magnitude 1 or more (0.5 and -0.5 are not: a
# product by them rounds a subnormal with its last bit set).
INLINE_POW2 = {"1.0", "-1.0", "2.0", "-2.0", "4.0", "-4.0"}


def pow2_bits(text):
    """Whether a 32-bit constant's float value is plus or minus a power of two of magnitude 1
    or more (a biased exponent of 127 or more, a zero fraction, finite)."""
    try:
        bits = int(text, 0) & 0xFFFFFFFF
    except ValueError:
        return False
    exponent = (bits >> 23) & 0xFF
    return (bits & 0x7FFFFF) == 0 and 127 <= exponent < 255


def exact_product(lines, i):
    """A fused op one of whose factors is a power of two: an inline constant, a literal, or an
    SGPR whose latest s_mov_b32 above (in the same function) set one. Its product is exact,
    so it rounds as the unfused sum does."""
    mnemonic, _, rest = lines[i].strip().partition(" ")
    ops = [o.strip() for o in rest.split(";")[0].split(",")]
    base = re.sub(r"_e(32|64)$", "", mnemonic)
    if base in ("v_fmamk_f32", "v_madmk_f32"):
        factors = [ops[1], ops[2]]
    elif len(ops) >= 3:
        factors = [ops[1], ops[2]]
    else:
        return False
    for f in factors:
        f = f.lstrip("-").strip("|")
        if f in INLINE_POW2 or (f.startswith("0x") and pow2_bits(f)):
            return True
        if re.fullmatch(r"s\d+", f):
            for j in range(i - 1, -1, -1):
                line = lines[j].strip()
                if line.startswith(".Lfunc_end") or line.endswith(":") and not line.startswith("."):
                    if not line.startswith(".LBB"):
                        break
                m = re.match(r"s_mov_b32 (s\d+), (\S+)", line)
                if m and m.group(1) == f:
                    if m.group(2) in INLINE_POW2 or pow2_bits(m.group(2)):
                        return True
                    break
                if re.match(rf"s_\w+ {f},", line):
                    break
    return False


def test_no_contraction_in_kernels():
    exact = 0
    for name in FILES:
        lines = asm_text(name).splitlines()
        for i, line in enumerate(lines):
            assert not PACKED.search(line), f"{name}:{i + 1}: {line.strip()}"
            if FUSED.search(line) and not in_division(lines, i):
                assert exact_product(lines, i), f"{name}:{i + 1}: fused op outside a division: {line.strip()}"
                exact += 1
    assert exact > 0  # the demosaic and sharpening fold their powers of two


def test_the_exact_product_rule():
    lines = ["s_mov_b32 s4, 0x49800000", "v_add_f32_e32 v1, v2, v3", "v_fma_f32 v0, v1, s4, 0.5",
             "v_fmac_f32_e32 v0, 2.0, v1", "v_fmac_f32_e32 v0, v3, v1", "v_fma_f32 v0, v1, 0x40400000, v2",
             "v_fmamk_f32 v0, v1, 0x4f800000, v2", "s_mov_b32 s4, 0x3f000001", "v_fma_f32 v0, v1, s4, v2",
             "v_fmac_f32_e32 v0, 0.5, v1", "v_fma_f32 v0, v1, -0.5, v2", "v_fmamk_f32 v0, v1, 0x3e000000, v2",
             "s_mov_b32 s5, 0x3d800000", "v_fma_f32 v0, v1, s5, v2", "v_fma_f32 v0, -v1, -4.0, v2"]
    assert exact_product(lines, 2) and exact_product(lines, 3) and exact_product(lines, 6)
    assert exact_product(lines, 14)  # -4.0 on a negated operand
    assert not exact_product(lines, 4)  # two registers
    assert not exact_product(lines, 5)  # 3.0
    assert not exact_product(lines, 8)  # s4 was reset to a non-power of two
    assert not exact_product(lines, 9)  # 0.5: halving rounds a subnormal
    assert not exact_product(lines, 10)  # -0.5
    assert not exact_product(lines, 11)  # 0.125 as a literal
    assert not exact_product(lines, 13)  # 0.0625 in a scalar register


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
        # The one-variant batch kernels add the queue and frame switch to a tile body
        # at the per-frame kernel's registers: a word or a few (at most 32 bytes).
        limit = 64 if "_wide" in name else 32 if "uniform" in name else 0
        assert scratch <= limit, f"{name}: {scratch} bytes of scratch"
