"""Pinning the oracle's set (precision conversion), sgn and cmp before they are trusted:

* the reference harness's own generators and MPFR gold (tests.c:580-699, built from the
  reference sources into oracle/_ref): set at several precision pairs, sgn, cmp;
* MPFR (mpfr_set / mpfr_cmp / mpfr_sgn) in RZ/RU/RD/RNE on random, all-ones (carry out
  of the rounding), near-tie and special-class inputs, 256..16384 bits;
* ties away from zero (no MPFR mode): exact integer rounding in Python;
* the sgn/cmp class tables against the reference's text (fpa_threaded.h:1-24).
"""
import os
import random
import re

import numpy as np
import pytest

from oracle import oracle as O

MODES = ["rz", "ru", "rd", "rne"]
REF = "/root/reference/fpa_threaded.h"


def rand_values(rng, count, p, e_range=1000, specials=True):
    n = p // 32
    limbs = rng.integers(0, 2 ** 32, size=(count, n), dtype=np.uint64).astype(np.uint32)
    limbs[:, -1] |= 0x80000000
    e = rng.integers(-e_range, e_range + 1, count)
    sign = rng.integers(0, 2, count)
    exps = (2 * (O.BIAS + e) + sign).astype(np.uint32)
    if specials and count >= 16:
        exps[:9] = np.arange(9, dtype=np.uint32)[[2, 3, 4, 5, 6, 7, 8, 2, 6]]
        limbs[:9] = 0
    return exps, limbs.reshape(-1)


def all_ones(count, p, rng):
    n = p // 32
    limbs = np.full((count, n), 0xFFFFFFFF, np.uint32)
    # vary the low words so the dropped part is all ones, a near-tie, or exactly a tie
    k = rng.integers(0, 3, count)
    limbs[k == 1, 0] = 0x7FFFFFFF
    limbs[k == 2, 0] = 0x80000000
    e = rng.integers(-50, 51, count)
    sign = rng.integers(0, 2, count)
    return (2 * (O.BIAS + e) + sign).astype(np.uint32), limbs.reshape(-1)


@pytest.mark.skipif(not O.ref_available(), reason="reference harness not built (no reference sources)")
@pytest.mark.parametrize("p,src", [(1024, 3072), (3072, 1024), (2048, 2048), (1024, 8192), (8192, 2048)])
def test_set_matches_reference_harness(p, src):
    t = O.ref_generate("set", 300, p, src)
    re_, rl = O.set_precision(p, src, t["a_exps"], t["a_limbs"], mode="rne")
    assert np.array_equal(re_, t["gold_exps"])
    assert O.results_equal(re_, rl, t["gold_exps"], t["gold_limbs"], p).all()


@pytest.mark.skipif(not O.ref_available(), reason="reference harness not built (no reference sources)")
@pytest.mark.parametrize("p", [1024, 4096, 8192])
def test_sgn_cmp_match_reference_harness(p):
    t = O.ref_generate("sgn", 400, p)
    assert np.array_equal(O.sgn(t["a_exps"]), np.sign(t["gold_exps"]))
    t = O.ref_generate("cmp", 400, p)
    got = O.cmp(p, t["a_exps"], t["a_limbs"], t["b_exps"], t["b_limbs"])
    assert np.array_equal(got, np.sign(t["gold_exps"]))


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("p,src", [(256, 512), (512, 256), (1024, 4096), (4096, 1024), (4096, 4096),
                                   (256, 16384), (16384, 256), (8192, 16384)])
def test_set_vs_mpfr(p, src, mode):
    rng = np.random.default_rng(p * 7 + src)
    for gen in ("random", "ones"):
        if gen == "random":
            xe, xl = rand_values(rng, 300, src)
        else:
            xe, xl = all_ones(200, src, rng)
        oe, ol = O.set_precision(p, src, xe, xl, mode=mode)
        me, ml, ok = O.mpfr_set(p, src, xe, xl, mode=mode)
        eq = O.results_equal(oe, ol, me, ml, p)
        assert eq[ok].all(), (gen, int((~eq[ok]).sum()))
        # TINY / SNAN (no MPFR meaning) copy through
        assert np.array_equal(oe[~ok], xe[~ok])


def test_set_overflow_to_inf_and_exact_widening():
    p, src = 256, 512
    xe, xl = all_ones(64, src, np.random.default_rng(3))
    xe[:] = (2 * 0x7FFFFFFF + (np.arange(64) & 1)).astype(np.uint32)   # largest exponent
    oe, ol = O.set_precision(p, src, xe, xl, mode="ru")
    pos = (xe & 1) == 0
    assert (oe[pos] == 6).all()                  # +all-ones rounded up past E = 0x7FFFFFFF -> +INF
    # widening is exact: the source value padded with zero limbs
    xe2, xl2 = rand_values(np.random.default_rng(4), 50, 512, specials=False)
    we, wl = O.set_precision(1024, 512, xe2, xl2)
    assert np.array_equal(we, xe2)
    w = wl.reshape(50, 32)
    assert (w[:, :16] == 0).all() and np.array_equal(w[:, 16:], xl2.reshape(50, 16))


def py_round_rna(limbs, nx, nr):
    """Narrowing nx -> nr words, ties away from zero on the magnitude: (mantissa, carried)"""
    assert nx > nr
    m = sum(int(w) << (32 * i) for i, w in enumerate(limbs))
    drop = 32 * (nx - nr)
    q, rem = m >> drop, m & ((1 << drop) - 1)
    if rem >= 1 << (drop - 1):
        q += 1
    carried = q >> (32 * nr) != 0
    if carried:
        q >>= 1
    return q, carried


@pytest.mark.parametrize("p,src", [(256, 512), (512, 1024), (256, 16384)])
def test_set_ties_away_brute_force(p, src):
    rng = np.random.default_rng(11)
    xe, xl = all_ones(100, src, rng)
    xe2, xl2 = rand_values(rng, 100, src, specials=False)
    # exact ties: the dropped part is exactly one half
    n, ns = p // 32, src // 32
    tie = xl2.reshape(100, ns).copy()
    tie[:, : ns - n] = 0
    tie[:, ns - n - 1] = 0x80000000
    xe = np.concatenate([xe, xe2, xe2])
    xl = np.concatenate([xl, xl2, tie.reshape(-1)])
    oe, ol = O.set_precision(p, src, xe, xl, mode="rna")
    for i in range(len(xe)):
        q, carried = py_round_rna(xl.reshape(-1, ns)[i], ns, n)
        want = [(q >> (32 * k)) & 0xFFFFFFFF for k in range(n)]
        assert list(ol.reshape(-1, n)[i]) == want
        assert oe[i] == xe[i] + (2 if carried else 0)


@pytest.mark.parametrize("p", [256, 2048, 16384])
def test_cmp_sgn_vs_mpfr(p):
    rng = np.random.default_rng(p)
    n = p // 32
    ae, al = rand_values(rng, 400, p, e_range=3)
    be, bl = rand_values(rng, 400, p, e_range=3)
    # equal exponent words with mantissas differing in one word, and exact equality
    be[100:200] = ae[100:200]
    bl.reshape(400, n)[100:200] = al.reshape(400, n)[100:200]
    w = rng.integers(0, n, 100)
    bl.reshape(400, n)[100 + np.arange(100), w] ^= rng.integers(1, 2 ** 32, 100, dtype=np.uint64).astype(np.uint32)
    bl.reshape(400, n)[:, -1] |= 0x80000000
    be[200:220] = ae[200:220]
    bl.reshape(400, n)[200:220] = al.reshape(400, n)[200:220]
    got = O.cmp(p, ae, al, be, bl)
    ref, ok = O.mpfr_cmp(p, ae, al, be, bl)
    assert np.array_equal(got[ok], ref[ok])
    assert (got[200:220][ok[200:220]] == 0).all()
    s = O.sgn(ae)
    ref, ok = O.mpfr_cmp(p, ae, al)
    assert np.array_equal(s[ok], ref[ok])


def test_tables_follow_reference_text():
    if not os.path.exists(REF):
        pytest.skip("reference sources not present")
    txt = open(REF).read()
    tok = {"FPA_CMP": 2, "FPA_CMP_ERR": 3}

    def entries(name):
        body = re.search(name + r"\[\d+\]\s*=\s*\{(.*?)\};", txt, re.S).group(1)
        body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
        vals = [v.strip() for v in body.split(",") if v.strip()]
        return [tok[v] if v in tok else int(v) for v in vals]

    sg = entries("sgn_action_table")
    assert sg == [O.sgn_table(c) for c in range(9)]
    cm = entries("cmp_action_table")
    assert cm == [O.cmp_table(a, b) for a in range(9) for b in range(9)]


FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "setcmp")


@pytest.mark.parametrize("path", sorted(sum((__import__("glob").glob(os.path.join(FIX, pat))
                                              for pat in ("ref_set_[0-9]*.npz", "ref_sgn_*.npz", "ref_cmp_*.npz")), [])),
                         ids=lambda s: os.path.basename(s))
def test_oracle_matches_setcmp_fixtures(path):
    z = np.load(path)
    kind = os.path.basename(path).split("_")[1]
    p = int(z["precision"])
    if kind == "set":
        re_, rl = O.set_precision(p, int(z["source_precision"]), z["a_exps"], z["a_limbs"])
        assert O.results_equal(re_, rl, z["gold_exps"], z["gold_limbs"], p).all()
    elif kind == "sgn":
        assert np.array_equal(O.sgn(z["a_exps"]), np.sign(z["gold_exps"]))
    else:
        assert np.array_equal(O.cmp(p, z["a_exps"], z["a_limbs"], z["b_exps"], z["b_limbs"]), np.sign(z["gold_exps"]))
