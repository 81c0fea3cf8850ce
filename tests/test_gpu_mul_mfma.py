"""The matrix-core mul kernels (shapes (n/64, 64): one value per wave, int8 MFMAs) at
2048, 4096, 6144, 8192 and 16384 bits.

Bar: bit-exact.  Against the oracle on config-2 operands in every rounding mode, on the
reference generator's adversarial limb alphabet, on all 9x9 class pairs and at the TINY/INF
exponent thresholds; against the radix-2^27 kernel (the default shape) on raw random limbs,
including mantissas whose top bit is clear (the recode carry is then 0), all-zero and
all-ones mantissas, and on a larger launch.
"""
import sys

import numpy as np
import pytest

import cubigfloat_amd as F
from oracle import oracle as O

pytestmark = pytest.mark.gpu

MFMA_PRECISIONS = [2048, 4096, 6144, 8192, 16384]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def mfma_variant(p):
    shapes = F.variants("mul", p)
    assert (p // 2048, 64) in shapes, shapes
    return shapes.index((p // 2048, 64))


def run_variant(p, v, ae, al, be, bl, mode="rne"):
    default = F.get_variant("mul", p)
    try:
        assert F.set_variant("mul", p, v) == 0
        return F.run_op("mul", p, ae, al, be, bl, mode=F.MODES[mode])
    finally:
        F.set_variant("mul", p, default)


def check_oracle(p, ae, al, be, bl, mode, tag):
    say(f"[gpu] mul mfma p={p} count={len(ae)} mode={mode} {tag}")
    re, rl = run_variant(p, mfma_variant(p), ae, al, be, bl, mode)
    oe, ol = O.op("mul", p, ae, al, be, bl, mode=mode)
    eq = O.results_equal(re, rl, oe, ol, p)
    assert eq.all(), f"p={p} {tag} {mode}: {int((~eq).sum())} mismatches, first {int(np.where(~eq)[0][0])}"


def check_valu(p, ae, al, be, bl, mode, tag):
    say(f"[gpu] mul mfma vs radix-2^27 p={p} count={len(ae)} mode={mode} {tag}")
    valu = next(i for i, (k, l) in enumerate(F.variants("mul", p)) if l != 64)
    re, rl = run_variant(p, mfma_variant(p), ae, al, be, bl, mode)
    we, wl = run_variant(p, valu, ae, al, be, bl, mode)
    assert np.array_equal(re, we), f"p={p} {tag}: exponent words differ at {int(np.where(re != we)[0][0])}"
    assert np.array_equal(rl, wl), f"p={p} {tag}: limbs differ in value {int(np.where(rl != wl)[0][0]) // (p // 32)}"


def count(p, base):
    return max(16, base * 4096 // p) if p >= 4096 else base


def test_variants_are_compiled_for_mul_only():
    for p in F.PRECISIONS:
        has = any(l == 64 for _, l in F.variants("mul", p))
        assert has == (p in MFMA_PRECISIONS), p
        for op in ("add", "sub", "div", "sqrt", "dot"):
            assert all(l != 64 for _, l in F.variants(op, p)), (op, p)


@pytest.mark.parametrize("p", MFMA_PRECISIONS)
@pytest.mark.parametrize("mode", ["rne", "rz", "ru", "rd", "rna"])
def test_config2_all_modes(p, mode):
    ae, al, be, bl = F.generate(2, 311, "mul", p, count(p, 2048))
    check_oracle(p, ae, al, be, bl, mode, "config-2")


@pytest.mark.parametrize("p", MFMA_PRECISIONS)
def test_adversarial_alphabet(p):
    """Limbs from {0, 1, 0x7FFFFFFF, 0x80000000, 0xFFFFFFFF} (tests.c:173-181), which
    make long runs of equal digits and carries through whole chunks."""
    rng = np.random.default_rng(17 + p)
    cnt = count(p, 512)
    N = p // 32
    alph = np.array([0, 1, 0x7FFFFFFF, 0x80000000, 0xFFFFFFFF], np.uint32)
    al = alph[rng.integers(0, 5, size=(cnt, N))]
    bl = alph[rng.integers(0, 5, size=(cnt, N))]
    al[:, -1] |= np.uint32(0x80000000)
    bl[:, -1] |= np.uint32(0x80000000)
    q = cnt // 8
    al[:2 * q] = 0xFFFFFFFF                   # all-ones mantissas: the largest product
    bl[:q] = 0xFFFFFFFF
    al[2 * q:3 * q] = 0
    al[2 * q:3 * q, -1] = 0x80000000          # 2^-1 exactly: the product is b, shifted
    e = O.BIAS + rng.integers(-40, 40, size=cnt)
    ae = (2 * e + rng.integers(0, 2, cnt)).astype(np.uint32)
    be = (2 * (O.BIAS + rng.integers(-40, 40, size=cnt)) + rng.integers(0, 2, cnt)).astype(np.uint32)
    for mode in ("rne", "rz", "ru", "rd", "rna"):
        check_oracle(p, ae, al.reshape(-1), be, bl.reshape(-1), mode, "adversarial")


@pytest.mark.parametrize("p", MFMA_PRECISIONS)
def test_special_classes(p):
    N = p // 32
    codes = [2 * (O.BIAS + 3), 2 * (O.BIAS + 3) + 1, 2, 3, 4, 5, 6, 7, 8]
    ae = np.repeat(np.array(codes, np.uint32), 9)
    be = np.tile(np.array(codes, np.uint32), 9)
    rng = np.random.default_rng(5)
    al = rng.integers(0, 2**32, size=81 * N, dtype=np.uint64).astype(np.uint32)
    bl = rng.integers(0, 2**32, size=81 * N, dtype=np.uint64).astype(np.uint32)
    al.reshape(81, N)[:, -1] |= np.uint32(0x80000000)
    bl.reshape(81, N)[:, -1] |= np.uint32(0x80000000)
    check_oracle(p, ae, al, be, bl, "rne", "classes")


@pytest.mark.parametrize("p", MFMA_PRECISIONS)
def test_exponent_range_edges(p):
    rng = np.random.default_rng(23)
    cnt = count(p, 600)
    N = p // 32
    lo, hi, B = 5, 0x7FFFFFFF, O.BIAS
    d = rng.integers(-3, 4, cnt)
    half = np.arange(cnt) < cnt // 2
    Ea = rng.integers(B - 2**29, B + 2**29, cnt)
    Eb = np.clip(np.where(half, B + lo - Ea + d, B + hi - Ea + d), lo, hi)
    ae = (2 * Ea + rng.integers(0, 2, cnt)).astype(np.uint32)
    be = (2 * Eb + rng.integers(0, 2, cnt)).astype(np.uint32)
    al = rng.integers(0, 2**32, size=cnt * N, dtype=np.uint64).astype(np.uint32)
    bl = rng.integers(0, 2**32, size=cnt * N, dtype=np.uint64).astype(np.uint32)
    al.reshape(cnt, N)[:, -1] |= np.uint32(0x80000000)
    bl.reshape(cnt, N)[:, -1] |= np.uint32(0x80000000)
    for mode in ("rne", "ru", "rd"):
        check_oracle(p, ae, al, be, bl, mode, "range")


@pytest.mark.parametrize("p", MFMA_PRECISIONS)
def test_matches_radix27_kernel_on_raw_limbs(p):
    """Random limbs with the top bit left as drawn, plus all-zero and all-ones mantissas."""
    rng = np.random.default_rng(29 + p)
    cnt = max(64, count(p, 1024))
    N = p // 32
    al = rng.integers(0, 2**32, size=(cnt, N), dtype=np.uint64).astype(np.uint32)
    bl = rng.integers(0, 2**32, size=(cnt, N), dtype=np.uint64).astype(np.uint32)
    al[:8] = 0
    bl[8:16] = 0
    al[16:24] = 0xFFFFFFFF
    bl[16:32] = 0xFFFFFFFF
    al[32:64, -1] &= np.uint32(0x7FFFFFFF)    # top bit clear: the recode carry is 0
    e = O.BIAS + rng.integers(-100, 100, size=cnt)
    ae = (2 * e + rng.integers(0, 2, cnt)).astype(np.uint32)
    be = (2 * (O.BIAS + rng.integers(-100, 100, size=cnt)) + rng.integers(0, 2, cnt)).astype(np.uint32)
    for mode in ("rne", "rz", "ru", "rd", "rna"):
        check_valu(p, ae, al.reshape(-1), be, bl.reshape(-1), mode, "raw limbs")


@pytest.mark.parametrize("p", MFMA_PRECISIONS)
def test_matches_radix27_kernel_large_launch(p):
    cnt = (1 << 28) // p               # 32 MiB of limbs per operand
    ae, al, be, bl = F.generate(2, 313, "mul", p, cnt)
    check_valu(p, ae, al, be, bl, "rne", f"{cnt} values")
