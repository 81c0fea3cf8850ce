"""The 4096-bit matrix-core mul kernel (shape K=2, L=64: one value per wave, int8 MFMAs).

Bar: bit-exact.  Against the oracle on config-2 operands in every rounding mode, on the
reference generator's adversarial limb alphabet, on all 9x9 class pairs and at the TINY/INF
exponent thresholds; against the radix-2^27 kernel (the default shape) on raw random limbs,
including mantissas whose top bit is clear (the recode carry is then 0), all-zero and
all-ones mantissas, and on a launch of 2^16 values.
"""
import sys

import numpy as np
import pytest

import cubigfloat_amd as F
from oracle import oracle as O

pytestmark = pytest.mark.gpu

P = 4096
N = P // 32


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def mfma_variant():
    shapes = F.variants("mul", P)
    assert (2, 64) in shapes, shapes
    return shapes.index((2, 64))


def run_variant(v, ae, al, be, bl, mode="rne"):
    default = F.get_variant("mul", P)
    try:
        assert F.set_variant("mul", P, v) == 0
        return F.run_op("mul", P, ae, al, be, bl, mode=F.MODES[mode])
    finally:
        F.set_variant("mul", P, default)


def check_oracle(ae, al, be, bl, mode, tag):
    say(f"[gpu] mul mfma p={P} count={len(ae)} mode={mode} {tag}")
    re, rl = run_variant(mfma_variant(), ae, al, be, bl, mode)
    oe, ol = O.op("mul", P, ae, al, be, bl, mode=mode)
    eq = O.results_equal(re, rl, oe, ol, P)
    assert eq.all(), f"{tag} {mode}: {int((~eq).sum())} mismatches, first {int(np.where(~eq)[0][0])}"


def check_valu(ae, al, be, bl, mode, tag):
    say(f"[gpu] mul mfma vs radix-2^27 p={P} count={len(ae)} mode={mode} {tag}")
    v = mfma_variant()
    valu = next(i for i, (k, l) in enumerate(F.variants("mul", P)) if l != 64)
    re, rl = run_variant(v, ae, al, be, bl, mode)
    we, wl = run_variant(valu, ae, al, be, bl, mode)
    assert np.array_equal(re, we), f"{tag}: exponent words differ at {int(np.where(re != we)[0][0])}"
    assert np.array_equal(rl, wl), f"{tag}: limbs differ in value {int(np.where(rl != wl)[0][0]) // N}"


def test_variant_is_compiled_for_mul_4096_only():
    assert (2, 64) in F.variants("mul", P)
    for op in ("add", "sub", "div", "sqrt", "dot"):
        assert (2, 64) not in F.variants(op, P), op
    for p in F.PRECISIONS:
        if p != P:
            assert (2, 64) not in F.variants("mul", p), p


@pytest.mark.parametrize("mode", ["rne", "rz", "ru", "rd", "rna"])
def test_config2_all_modes(mode):
    ae, al, be, bl = F.generate(2, 311, "mul", P, 2048)
    check_oracle(ae, al, be, bl, mode, "config-2")


def test_adversarial_alphabet():
    """Limbs from {0, 1, 0x7FFFFFFF, 0x80000000, 0xFFFFFFFF} (tests.c:173-181), which
    make long runs of equal digits and carries through whole chunks."""
    rng = np.random.default_rng(17)
    cnt = 512
    alph = np.array([0, 1, 0x7FFFFFFF, 0x80000000, 0xFFFFFFFF], np.uint32)
    al = alph[rng.integers(0, 5, size=(cnt, N))]
    bl = alph[rng.integers(0, 5, size=(cnt, N))]
    al[:, -1] |= np.uint32(0x80000000)
    bl[:, -1] |= np.uint32(0x80000000)
    al[:64] = 0xFFFFFFFF                      # all-ones mantissas: the largest product
    bl[:32] = 0xFFFFFFFF
    al[64:96] = 0
    al[64:96, -1] = 0x80000000                # 2^-1 exactly: the product is b, shifted
    e = O.BIAS + rng.integers(-40, 40, size=cnt)
    ae = (2 * e + rng.integers(0, 2, cnt)).astype(np.uint32)
    be = (2 * (O.BIAS + rng.integers(-40, 40, size=cnt)) + rng.integers(0, 2, cnt)).astype(np.uint32)
    for mode in ("rne", "rz", "ru", "rd", "rna"):
        check_oracle(ae, al.reshape(-1), be, bl.reshape(-1), mode, "adversarial")


def test_special_classes():
    codes = [2 * (O.BIAS + 3), 2 * (O.BIAS + 3) + 1, 2, 3, 4, 5, 6, 7, 8]
    ae = np.repeat(np.array(codes, np.uint32), 9)
    be = np.tile(np.array(codes, np.uint32), 9)
    rng = np.random.default_rng(5)
    al = rng.integers(0, 2**32, size=81 * N, dtype=np.uint64).astype(np.uint32)
    bl = rng.integers(0, 2**32, size=81 * N, dtype=np.uint64).astype(np.uint32)
    al.reshape(81, N)[:, -1] |= np.uint32(0x80000000)
    bl.reshape(81, N)[:, -1] |= np.uint32(0x80000000)
    check_oracle(ae, al, be, bl, "rne", "classes")


def test_exponent_range_edges():
    rng = np.random.default_rng(23)
    cnt = 600
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
        check_oracle(ae, al, be, bl, mode, "range")


def test_matches_radix27_kernel_on_raw_limbs():
    """Random limbs with the top bit left as drawn, plus all-zero and all-ones mantissas."""
    rng = np.random.default_rng(29)
    cnt = 1024
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
        check_valu(ae, al.reshape(-1), be, bl.reshape(-1), mode, "raw limbs")


def test_matches_radix27_kernel_large_launch():
    ae, al, be, bl = F.generate(2, 313, "mul", P, 1 << 16)
    check_valu(ae, al, be, bl, "rne", "2^16 values")
