"""fpa_set_ui / fpa_set_si / fpa_set_float / fpa_set_double on the CPU.

* The oracle (oracle/fpa_oracle.c orc_set_scalar_array) against the reference harness's own
  generators and MPFR gold (tests_generate_set_ui/si/float/double, tests.c:487-578, built
  from its sources into oracle/_ref/) where those sources exist, and always against the
  fixtures frozen from them (tests/golden/setcmp/ref_set_<kind>_<p>.npz).
* The oracle against exact rational arithmetic: every result must equal the source value
  exactly (Fraction), with its significand in the top two limbs and zeros below, on edge
  values (zero of both signs, INT64_MIN, UINT64_MAX, subnormal and largest floats, NaN,
  infinities) and random bit patterns.
* The library's encoder -- the function the set kernels run, also compiled for the host
  and exported as fpa_scalar_encode -- against the oracle on the same values (no GPU).
"""
import glob
import os
from fractions import Fraction

import numpy as np
import pytest

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
KINDS = ["set_ui", "set_si", "set_float", "set_double"]


def edge_values(name, rng, extra=2000):
    if name == "set_ui":
        v = [0, 1, 2, 3, 2 ** 31, 2 ** 32 - 1, 2 ** 32, 2 ** 63, 2 ** 64 - 1, 2 ** 64 - 2]
        r = rng.integers(0, 2 ** 64, extra, dtype=np.uint64)
        r[: extra // 2] >>= rng.integers(0, 64, extra // 2).astype(np.uint64)
        return np.concatenate([np.array(v, np.uint64), r])
    if name == "set_si":
        v = [0, 1, -1, 2, -2, 2 ** 31, -2 ** 31, 2 ** 63 - 1, -2 ** 63, -2 ** 63 + 1, -2 ** 32]
        r = rng.integers(-2 ** 63, 2 ** 63, extra, dtype=np.int64)
        r[: extra // 2] >>= rng.integers(0, 64, extra // 2).astype(np.int64)
        return np.concatenate([np.array(v, np.int64), r])
    if name == "set_float":
        bits = [0x00000000, 0x80000000, 0x00000001, 0x80000001, 0x007FFFFF, 0x00800000, 0x7F7FFFFF,
                0xFF7FFFFF, 0x3F800000, 0xBF800000, 0x7F800000, 0xFF800000, 0x7FC00000, 0xFFC00000,
                0x7F800001, 0x00400000]
        r = rng.integers(0, 2 ** 32, extra, dtype=np.uint64).astype(np.uint32)
        return np.concatenate([np.array(bits, np.uint32), r]).view(np.float32)
    bits = [0, 1 << 63, 1, (1 << 63) | 1, (1 << 52) - 1, 1 << 52, 0x7FEFFFFFFFFFFFFF, 0xFFEFFFFFFFFFFFFF,
            0x3FF0000000000000, 0xBFF0000000000000, 0x7FF0000000000000, 0xFFF0000000000000,
            0x7FF8000000000000, 0xFFF8000000000000, 0x7FF0000000000001, 1 << 51]
    r = rng.integers(0, 2 ** 64, extra, dtype=np.uint64)
    return np.concatenate([np.array(bits, np.uint64), r]).view(np.float64)


def exact_check(name, p, values, re_, rl):
    """Each result equals its source exactly, or is the right class code."""
    n = p // 32
    rl = rl.reshape(-1, n)
    for v, w, limbs in zip(values, re_, rl):
        if name in ("set_float", "set_double") and np.isnan(v):
            assert w == 7
            continue
        if name in ("set_float", "set_double") and np.isinf(v):
            assert w == (5 if v < 0 else 6)
            continue
        x = Fraction(int(v)) if name in ("set_ui", "set_si") else Fraction(float(v))
        if x == 0:
            assert w == 2
            continue
        assert w > 8
        assert not limbs[: n - 2].any()
        top = (int(limbs[n - 1]) << 32) | int(limbs[n - 2])
        assert top >> 63 == 1
        sign = -1 if w & 1 else 1
        E = int(w >> 1) - O.BIAS
        assert sign * Fraction(top) * Fraction(2) ** (E - 64) == x, (name, v, hex(w))


@pytest.mark.parametrize("name", KINDS)
@pytest.mark.parametrize("p", [256, 1024, 16384])
def test_oracle_exact(name, p):
    vals = edge_values(name, np.random.default_rng(p + len(name)))
    re_, rl = O.set_scalar(name, p, vals)
    exact_check(name, p, vals, re_, rl)


@pytest.mark.parametrize("name", KINDS)
def test_library_encoder_matches_oracle(name):
    import cubigfloat_amd as F
    kind = O.SCALAR_KINDS[name]
    vals = edge_values(name, np.random.default_rng(99), extra=20000)
    re_, rl = O.set_scalar(name, 256, vals)
    rl = rl.reshape(-1, 8)
    bits = vals.view(np.uint32 if name == "set_float" else np.uint64)
    for i, b in enumerate(bits):
        w, top = F.scalar_encode(kind, int(b))
        assert w == re_[i], (name, hex(int(b)), hex(w), hex(int(re_[i])))
        if w > 8:
            assert top == (int(rl[i, 7]) << 32) | int(rl[i, 6])


@pytest.mark.skipif(not O.ref_available(), reason="reference harness not built (needs the reference sources)")
@pytest.mark.parametrize("name", KINDS)
@pytest.mark.parametrize("p", [1024, 4096, 8192])
def test_oracle_matches_reference_harness(name, p):
    g = O.ref_generate(name, 2000, p)
    re_, rl = O.set_scalar(name, p, g["values"])
    assert O.results_equal(re_, rl, g["gold_exps"], g["gold_limbs"], p).all()


@pytest.mark.skipif(not O.ref_available(), reason="reference harness not built (needs the reference sources)")
@pytest.mark.parametrize("p", [1024, 4096, 8192])
def test_oracle_neg_matches_reference_harness(p):
    g = O.ref_generate("neg", 2000, p)
    re_, rl = O.op("neg", p, g["a_exps"], g["a_limbs"])
    assert O.results_equal(re_, rl, g["gold_exps"], g["gold_limbs"], p).all()


FIX = sorted(glob.glob(os.path.join(HERE, "golden", "setcmp", "ref_set_[usfd]*.npz")) +
             glob.glob(os.path.join(HERE, "golden", "setcmp", "ref_neg_*.npz")))


def test_fixtures_present():
    names = {os.path.basename(f).rsplit("_", 1)[0] for f in FIX}
    assert names == {"ref_set_ui", "ref_set_si", "ref_set_float", "ref_set_double", "ref_neg"}


@pytest.mark.parametrize("path", FIX, ids=lambda s: os.path.basename(s))
def test_oracle_matches_fixtures(path):
    z = np.load(path)
    p = int(z["precision"])
    name = os.path.basename(path)[4:].rsplit("_", 1)[0]
    if name == "neg":
        re_, rl = O.op("neg", p, z["a_exps"], z["a_limbs"])
    else:
        re_, rl = O.set_scalar(name, p, z["values"])
        exact_check(name, p, z["values"], re_, rl)
    assert O.results_equal(re_, rl, z["gold_exps"], z["gold_limbs"], p).all()
