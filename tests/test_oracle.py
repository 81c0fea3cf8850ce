"""CPU tests: pin the oracle (fpa_oracle.c) before trusting it.

* golden fixtures: the reference generator's cases with the reference harness's MPFR gold
  (tests.c), and MPFR gold at the extended precisions;
* MPFR (an independent implementation of correct rounding) on random, adversarial,
  near-cancellation and special-class inputs in the four MPFR rounding modes;
* ties-away-from-zero, which MPFR lacks (tests.c:69-71): brute force with Python
  integers at small precisions;
* big-integer kernels (Knuth D, integer square root) against Python integers;
* mul products on every rounding boundary (exact, ties with the kept part even and odd,
  just below and above half; product's top bit clear and set) in all five modes, against
  rounding by definition on the exact product with Python integers, at every precision;
* div and sqrt on exact results (a = u*v over v, squares of short integers) and one unit in
  the last place away from them, in all five modes, against rounding by definition with
  Python integers (long division, integer square root);
* add and sub with the smaller operand's bits below the larger one's last place exactly
  half, just below or above half, or zero, at exponent gaps 0..p+2 with both signs, in all
  five modes, against rounding by definition of the exact sum.
"""
import glob
import os
import random

import numpy as np
import pytest

from oracle import oracle as O

import addsub_boundaries as AB
import divsqrt_boundaries as DB
import mul_boundaries as MB

HERE = os.path.dirname(os.path.abspath(__file__))
OPS = ["add", "sub", "mul", "div", "sqrt"]


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(HERE, "golden", "*.npz"))),
                         ids=lambda s: os.path.basename(s))
def test_oracle_matches_golden(path):
    z = np.load(path)
    op = os.path.basename(path).split("_")[1]
    p = int(z["precision"])
    be = z["b_exps"] if "b_exps" in z else None
    bl = z["b_limbs"] if "b_limbs" in z else None
    oe, ol = O.op(op, p, z["a_exps"], z["a_limbs"], be, bl)
    assert O.results_equal(oe, ol, z["gold_exps"], z["gold_limbs"], p).all()


@pytest.mark.skipif(not O.ref_available(), reason="reference harness not built (no reference sources)")
@pytest.mark.parametrize("op", OPS)
@pytest.mark.parametrize("p", [1024, 3072, 8192])
def test_oracle_matches_reference_harness(op, p):
    g = O.ref_generate(op, 300 if p < 8192 else 60, p)
    oe, ol = O.op(op, p, g["a_exps"], g["a_limbs"], g.get("b_exps"), g.get("b_limbs"))
    assert O.results_equal(oe, ol, g["gold_exps"], g["gold_limbs"], p).all()


def _random_inputs(p, cnt, seed, alphabet=False, gap_scale=None):
    rng = np.random.default_rng(seed)
    n = p // 32
    if alphabet:
        alph = np.array([0, 1, 0x7FFFFFFF, 0x80000000, 0xFFFFFFFF], np.uint32)
        al = alph[rng.integers(0, 5, size=(cnt, n))]
        bl = alph[rng.integers(0, 5, size=(cnt, n))]
    else:
        al = rng.integers(0, 2**32, size=(cnt, n), dtype=np.uint64).astype(np.uint32)
        bl = rng.integers(0, 2**32, size=(cnt, n), dtype=np.uint64).astype(np.uint32)
    al[:, -1] |= np.uint32(0x80000000)
    bl[:, -1] |= np.uint32(0x80000000)
    g = gap_scale or p
    ea = O.BIAS + rng.integers(-g, g, cnt)
    eb = ea + rng.integers(-g - 40, g + 40, cnt)
    # equal and adjacent exponents often
    eb = np.where(rng.integers(0, 4, cnt) == 0, ea + rng.integers(-2, 3, cnt), eb)
    ae = (2 * ea + rng.integers(0, 2, cnt)).astype(np.uint32)
    be = (2 * eb + rng.integers(0, 2, cnt)).astype(np.uint32)
    return ae, al.reshape(-1), be, bl.reshape(-1)


@pytest.mark.parametrize("mode", ["rz", "ru", "rd", "rne"])
@pytest.mark.parametrize("op", OPS)
@pytest.mark.parametrize("p", [256, 1024, 4096])
@pytest.mark.parametrize("alphabet", [False, True])
def test_oracle_vs_mpfr(op, p, mode, alphabet):
    ae, al, be, bl = _random_inputs(p, 400 if p < 4096 else 80, 17 + p, alphabet)
    if op == "sqrt":
        ae &= np.uint32(0xFFFFFFFE)
        be = bl = None
    oe, ol = O.op(op, p, ae, al, be, bl, mode=mode)
    me, ml, ok = O.mpfr_gold(op, p, ae, al, be, bl, mode=mode)
    assert ok.all()
    assert O.results_equal(oe, ol, me, ml, p).all()


@pytest.mark.parametrize("op", OPS)
def test_oracle_vs_mpfr_config2_inputs(op):
    import cubigfloat_amd as F
    for p in (512, 4096, 16384):
        ae, al, be, bl = F.generate(2, 3, op, p, 200 if p < 16384 else 20, binary=op != "sqrt")
        oe, ol = O.op(op, p, ae, al, be, bl)
        me, ml, ok = O.mpfr_gold(op, p, ae, al, be, bl)
        assert ok.all()
        assert O.results_equal(oe, ol, me, ml, p).all()


@pytest.mark.parametrize("op", ["add", "sub", "mul", "div", "sqrt", "neg"])
def test_special_classes_vs_mpfr(op):
    """Classes MPFR can represent (values, zero, +-inf, nan) follow MPFR exactly."""
    p = 256
    n = p // 32
    codes = [2 * (O.BIAS + 3), 2 * (O.BIAS + 3) + 1, 2, 5, 6, 7]
    ae = np.repeat(np.array(codes, np.uint32), len(codes))
    be = np.tile(np.array(codes, np.uint32), len(codes))
    al = np.full(len(ae) * n, 0xC0000000, np.uint32)
    bl = np.full(len(ae) * n, 0xA0000000, np.uint32)
    if op in ("sqrt", "neg"):
        be = bl = None
    oe, ol = O.op(op, p, ae, al, be, bl)
    me, ml, ok = O.mpfr_gold(op, p, ae, al, be, bl)
    assert ok.all()
    assert O.results_equal(oe, ol, me, ml, p).all()


def _to_int(limbs):
    return sum(int(w) << (32 * i) for i, w in enumerate(limbs))


def _round_ties_away(num, den, p):
    """Round |num/den| > 0 to p bits, ties away from zero. Returns (mantissa, e) with
    value = mantissa * 2^(e - p)."""
    from fractions import Fraction
    q = Fraction(num, den)
    e = q.numerator.bit_length() - q.denominator.bit_length()
    while Fraction(2) ** e <= q:
        e += 1
    while Fraction(2) ** (e - 1) > q:
        e -= 1
    scaled = q * Fraction(2) ** (p - e)       # in [2^(p-1), 2^p)
    m = scaled.numerator // scaled.denominator
    frac = scaled - m
    if frac >= Fraction(1, 2):
        m += 1
    if m == 1 << p:
        m >>= 1
        e += 1
    return m, e


@pytest.mark.parametrize("op", ["add", "sub", "mul", "div", "sqrt"])
def test_ties_away_brute_force(op):
    """Mode 4 has no MPFR equivalent: exact rational arithmetic at p = 64 and 96 bits,
    with operands built so that exact ties occur."""
    rnd = random.Random(7)
    import math
    for p in (64, 96):
        n = p // 32
        cnt = 300
        ae = np.zeros(cnt, np.uint32)
        be = np.zeros(cnt, np.uint32)
        al = np.zeros(cnt * n, np.uint32)
        bl = np.zeros(cnt * n, np.uint32)
        for i in range(cnt):
            A = rnd.getrandbits(p) | (1 << (p - 1))
            Bv = rnd.getrandbits(p) | (1 << (p - 1))
            if i % 3 == 0:
                A = ((A >> 8) << 8) | (1 << (p - 1))          # trailing zeros -> add/sub ties
                Bv = (1 << (p - 1)) | rnd.getrandbits(4)
            ea_, eb_ = rnd.randint(-5, 5), rnd.randint(-5, 5)
            if i % 3 == 0:
                eb_ = ea_ - p - rnd.randint(0, 2)
            ae[i] = 2 * (O.BIAS + ea_) + rnd.getrandbits(1)
            be[i] = 2 * (O.BIAS + eb_) + rnd.getrandbits(1)
            if op == "sqrt":
                ae[i] &= 0xFFFFFFFE
            for w in range(n):
                al[i * n + w] = (A >> (32 * w)) & 0xFFFFFFFF
                bl[i * n + w] = (Bv >> (32 * w)) & 0xFFFFFFFF
        oe, ol = O.op(op, p, ae, al, None if op == "sqrt" else be, None if op == "sqrt" else bl, mode="rna")
        for i in range(cnt):
            A = _to_int(al[i * n:(i + 1) * n])
            Bv = _to_int(bl[i * n:(i + 1) * n])
            sa, sb = int(ae[i]) & 1, int(be[i]) & 1
            Ea, Eb = (int(ae[i]) >> 1) - O.BIAS - p, (int(be[i]) >> 1) - O.BIAS - p
            from fractions import Fraction
            va = Fraction(A) * Fraction(2) ** Ea * (-1 if sa else 1)
            vb = Fraction(Bv) * Fraction(2) ** Eb * (-1 if sb else 1)
            if op == "add":
                v = va + vb
            elif op == "sub":
                v = va - vb
            elif op == "mul":
                v = va * vb
            elif op == "div":
                v = va / vb
            else:
                # sqrt: floor root at 2p+8 extra bits plus a sticky bit is exact enough to round
                sc = Ea
                k = 4 * p + 16
                if (sc - k) % 2:
                    k += 1
                X = A << k
                r = math.isqrt(X)
                sticky = r * r != X
                v = (Fraction(2 * r + (1 if sticky else 0), 2)) * Fraction(2) ** ((sc - k) // 2)
            if v == 0:
                assert oe[i] == 2
                continue
            sign = 1 if v < 0 else 0
            m, e = _round_ties_away(abs(v.numerator), v.denominator, p)
            assert int(oe[i]) == 2 * (O.BIAS + e) + sign, (op, p, i)
            assert _to_int(ol[i * n:(i + 1) * n]) == m, (op, p, i)


def test_bignum_kernels_vs_python():
    import ctypes
    lib = O.lib()
    rnd = random.Random(3)
    u32 = ctypes.POINTER(ctypes.c_uint32)
    for _ in range(300):
        nu, nv = rnd.randint(2, 40), rnd.randint(1, 20)
        if nv > nu:
            nu, nv = nv, nu
        U = rnd.getrandbits(32 * nu) | 1
        V = rnd.getrandbits(32 * nv) | (1 << (32 * nv - rnd.randint(1, 32)))
        if rnd.random() < 0.3:
            V = (1 << (32 * nv - 1)) | rnd.choice([0, 1, (1 << 31) - 1])
        ua = np.array([(U >> (32 * i)) & 0xFFFFFFFF for i in range(nu)], np.uint32)
        va = np.array([(V >> (32 * i)) & 0xFFFFFFFF for i in range(nv)], np.uint32)
        q = np.zeros(nu + 1, np.uint32)
        r = np.zeros(nv, np.uint32)
        lib.orc_bn_divmod(q.ctypes.data_as(u32), r.ctypes.data_as(u32), ua.ctypes.data_as(u32), nu,
                          va.ctypes.data_as(u32), nv)
        assert _to_int(q[:nu - nv + 1]) == U // V
        assert _to_int(r) == U % V
        X = rnd.getrandbits(32 * nu)
        xa = np.array([(X >> (32 * i)) & 0xFFFFFFFF for i in range(nu)], np.uint32)
        s = np.zeros(nu // 2 + 1, np.uint32)
        inexact = lib.orc_bn_isqrt(s.ctypes.data_as(u32), xa.ctypes.data_as(u32), nu)
        import math
        assert _to_int(s) == math.isqrt(X)
        assert bool(inexact) == (math.isqrt(X) ** 2 != X)


@pytest.mark.parametrize("p", [256, 2048])
def test_oracle_dot_vs_mpfr(p):
    """Dot chains: the oracle's fold of exact mul/add equals MPFR's mpfr_mul/mpfr_add fold."""
    import cubigfloat_amd as F
    rng = np.random.default_rng(p)
    lengths = rng.integers(0, 30, 60).astype(np.uint32)
    total = int(lengths.sum())
    xe, xl, ye, yl = F.generate(2, 5, "mul", p, total)
    for mode in ("rne", "rz", "ru"):
        oe, ol = O.dot(p, lengths, xe, xl, ye, yl, mode=mode)
        me, ml = O.mpfr_dot(p, lengths, xe, xl, ye, yl, mode=mode)
        assert O.results_equal(oe, ol, me, ml, p).all()


@pytest.mark.parametrize("p", [256, 512, 1024, 2048, 3072, 4096, 5120, 6144, 7168, 8192, 16384])
def test_oracle_mul_rounding_boundaries(p):
    ae, al, be, bl = MB.operands(p, 6, 101 + p)
    for mode in MB.MODES:
        oe, ol = O.op("mul", p, ae, al, be, bl, mode=mode)
        xe, xl = MB.reference(p, ae, al, be, bl, mode)
        eq = O.results_equal(oe, ol, xe, xl, p)
        assert eq.all(), f"mode={mode}: {int((~eq).sum())} of {len(ae)} differ"


@pytest.mark.parametrize("p", [256, 512, 1024, 2048, 3072, 4096, 5120, 6144, 7168, 8192, 16384])
def test_oracle_div_sqrt_exact_boundaries(p):
    ae, al, be, bl = DB.div_operands(p, 48, 7 + p)
    xe, xl = DB.sqrt_operands(p, 48, 9 + p)
    for mode in DB.MODES:
        oe, ol = O.op("div", p, ae, al, be, bl, mode=mode)
        re, rl = DB.div_reference(p, ae, al, be, bl, mode)
        assert O.results_equal(oe, ol, re, rl, p).all(), f"div mode={mode}"
        oe, ol = O.op("sqrt", p, xe, xl, None, None, mode=mode)
        re, rl = DB.sqrt_reference(p, xe, xl, mode)
        assert O.results_equal(oe, ol, re, rl, p).all(), f"sqrt mode={mode}"


@pytest.mark.parametrize("p", [256, 512, 1024, 2048, 3072, 4096, 5120, 6144, 7168, 8192, 16384])
def test_oracle_add_sub_rounding_boundaries(p):
    ae, al, be, bl = AB.operands(p, 96, 5 + p)
    for op in ("add", "sub"):
        for mode in AB.MODES:
            oe, ol = O.op(op, p, ae, al, be, bl, mode=mode)
            re, rl = AB.reference(op, p, ae, al, be, bl, mode)
            assert O.results_equal(oe, ol, re, rl, p).all(), f"{op} mode={mode}"
