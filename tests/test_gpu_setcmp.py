"""fpa_set / fpa_sgn / fpa_cmp on the device through the C ABI, bit-exact against the
oracle (which is pinned against the reference harness's MPFR gold and MPFR)."""
import glob
import os

import numpy as np
import pytest

import cubigfloat_amd as F
from oracle import oracle as O

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
MODES = ["rz", "ru", "rd", "rne", "rna"]


def say(*a):
    print(*a, flush=True)


def values(rng, count, p, e_range=1000):
    n = p // 32
    limbs = rng.integers(0, 2 ** 32, size=(count, n), dtype=np.uint64).astype(np.uint32)
    limbs[:, -1] |= 0x80000000
    # a quarter with all-ones tails (rounding carries out of the mantissa)
    limbs[: count // 4, :] = 0xFFFFFFFF
    e = rng.integers(-e_range, e_range + 1, count)
    sign = rng.integers(0, 2, count)
    exps = (2 * (O.BIAS + e) + sign).astype(np.uint32)
    exps[count // 4: count // 4 + 9] = np.arange(9)[[2, 3, 4, 5, 6, 7, 8, 2, 6]]
    return exps, limbs.reshape(-1)


def device_set(p, src, xe, xl, mode):
    X, R = F.FpaArray(len(xe), src), F.FpaArray(len(xe), p)
    try:
        X.upload(xe, xl)
        assert F.set_precision(R, X, F.MODES[mode]) == F.SUCCESS
        return R.download()
    finally:
        X.close()
        R.close()


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("p,src", [(256, 512), (512, 256), (1024, 4096), (4096, 1024), (4096, 4096),
                                   (3072, 8192), (256, 16384), (16384, 256), (8192, 16384)])
def test_set_vs_oracle(p, src, mode):
    rng = np.random.default_rng(p + 3 * src)
    xe, xl = values(rng, 200, src)
    say(f"[gpu] set {src} -> {p} mode={mode}")
    re_, rl = device_set(p, src, xe, xl, mode)
    oe, ol = O.set_precision(p, src, xe, xl, mode=mode)
    eq = O.results_equal(re_, rl, oe, ol, p)
    assert eq.all(), f"{int((~eq).sum())} mismatches"


def test_set_every_shape_and_in_place():
    rng = np.random.default_rng(5)
    for p, src in ((2048, 4096), (4096, 2048)):
        xe, xl = values(rng, 100, src)
        oe, ol = O.set_precision(p, src, xe, xl)
        keep = F.get_variant("neg", p)
        try:
            for v, (k, l) in enumerate(F.variants("neg", p)):
                F.set_variant("neg", p, v)
                say(f"[gpu] set {src} -> {p} shape K={k} L={l}")
                re_, rl = device_set(p, src, xe, xl, "rne")
                assert O.results_equal(re_, rl, oe, ol, p).all(), (k, l)
        finally:
            F.set_variant("neg", p, keep)
    # r and x the same array (same precision): the identity
    xe, xl = values(rng, 64, 1024)
    X = F.FpaArray(64, 1024)
    try:
        X.upload(xe, xl)
        assert F.set_precision(X, X) == F.SUCCESS
        ee, ll = X.download()
    finally:
        X.close()
    assert O.results_equal(ee, ll, xe, xl, 1024).all()


@pytest.mark.parametrize("p", [256, 1024, 4096, 16384])
def test_sgn_cmp_vs_oracle(p):
    rng = np.random.default_rng(p)
    n = p // 32
    ae, al = values(rng, 400, p, e_range=2)
    be, bl = values(rng, 400, p, e_range=2)
    # every class pair, equal words with a difference in one limb, exact equality
    ae[100:181] = np.repeat(np.array([O.BIAS * 2, O.BIAS * 2 + 1, 2, 3, 4, 5, 6, 7, 8], np.uint32), 9)
    be[100:181] = np.tile(np.array([O.BIAS * 2, O.BIAS * 2 + 1, 2, 3, 4, 5, 6, 7, 8], np.uint32), 9)
    be[200:260] = ae[200:260]
    bl.reshape(400, n)[200:260] = al.reshape(400, n)[200:260]
    w = rng.integers(0, n, 40)
    bl.reshape(400, n)[200 + np.arange(40), w] ^= 1
    bl.reshape(400, n)[:, -1] |= 0x80000000
    A, B = F.FpaArray(400, p), F.FpaArray(400, p)
    try:
        A.upload(ae, al)
        B.upload(be, bl)
        say(f"[gpu] sgn/cmp p={p}")
        s = F.sgn(A)
        c = F.cmp(A, B)
    finally:
        A.close()
        B.close()
    assert np.array_equal(s, O.sgn(ae))
    assert np.array_equal(c, O.cmp(p, ae, al, be, bl))
    assert (c[240:260] == 0).all()


def test_cmp_every_shape():
    rng = np.random.default_rng(9)
    p, n = 4096, 128
    ae, al = values(rng, 200, p, e_range=1)
    be, bl = ae.copy(), al.copy()
    bl.reshape(200, n)[np.arange(200), rng.integers(0, n, 200)] ^= 4
    bl.reshape(200, n)[:, -1] |= 0x80000000
    want = O.cmp(p, ae, al, be, bl)
    keep = F.get_variant("neg", p)
    A, B = F.FpaArray(200, p), F.FpaArray(200, p)
    try:
        A.upload(ae, al)
        B.upload(be, bl)
        for v, (k, l) in enumerate(F.variants("neg", p)):
            F.set_variant("neg", p, v)
            say(f"[gpu] cmp shape K={k} L={l}")
            assert np.array_equal(F.cmp(A, B), want), (k, l)
    finally:
        F.set_variant("neg", p, keep)
        A.close()
        B.close()


SETCMP_FIX = sorted(sum((glob.glob(os.path.join(HERE, "golden", "setcmp", pat))
                         for pat in ("ref_set_[0-9]*.npz", "ref_sgn_*.npz", "ref_cmp_*.npz")), []))


@pytest.mark.parametrize("path", SETCMP_FIX,
                         ids=lambda s: os.path.basename(s))
def test_setcmp_fixtures(path):
    z = np.load(path)
    kind = os.path.basename(path).split("_")[1]
    p = int(z["precision"])
    say(f"[gpu] fixture {os.path.basename(path)}")
    if kind == "set":
        re_, rl = device_set(p, int(z["source_precision"]), z["a_exps"], z["a_limbs"], "rne")
        assert O.results_equal(re_, rl, z["gold_exps"], z["gold_limbs"], p).all()
        return
    A = F.FpaArray(len(z["a_exps"]), p)
    B = None
    try:
        A.upload(z["a_exps"], z["a_limbs"])
        if kind == "sgn":
            got = F.sgn(A)
        else:
            B = F.FpaArray(len(z["b_exps"]), p)
            B.upload(z["b_exps"], z["b_limbs"])
            got = F.cmp(A, B)
    finally:
        A.close()
        if B is not None:
            B.close()
    assert np.array_equal(got, np.sign(z["gold_exps"]))


# ------------------------ fpa_set_ui / si / float / double, fpa_neg ------------------------
import test_set_scalar as TS  # noqa: E402

SCALAR_FIX = sorted(glob.glob(os.path.join(HERE, "golden", "setcmp", "ref_set_[usfd]*.npz")) +
                    glob.glob(os.path.join(HERE, "golden", "setcmp", "ref_neg_*.npz")))


def device_set_scalars(name, p, values):
    R = F.FpaArray(len(values), p)
    try:
        assert F.set_scalars(R, name, values) == F.SUCCESS
        return R.download()
    finally:
        R.close()


@pytest.mark.parametrize("name", TS.KINDS)
@pytest.mark.parametrize("p", F.PRECISIONS)
def test_set_scalars_vs_oracle(name, p):
    """Edge values (zeros, INT64_MIN, subnormals, NaN, infinities) and random bit patterns,
    with the result limbs pre-filled with garbage so every limb the kernel writes is seen."""
    vals = TS.edge_values(name, np.random.default_rng(p), extra=3000)
    say(f"[gpu] {name} p={p} count={len(vals)}")
    R = F.FpaArray(len(vals), p)
    try:
        rng = np.random.default_rng(1)
        R.upload(rng.integers(9, 2 ** 32, len(vals), dtype=np.uint64).astype(np.uint32),
                 rng.integers(0, 2 ** 32, len(vals) * (p // 32), dtype=np.uint64).astype(np.uint32))
        assert F.set_scalars(R, name, vals) == F.SUCCESS
        re_, rl = R.download()
    finally:
        R.close()
    oe, ol = O.set_scalar(name, p, vals)
    eq = O.results_equal(re_, rl, oe, ol, p)
    assert eq.all(), f"{int((~eq).sum())} mismatches, first at {int(np.where(~eq)[0][0])}"
    # class codes write zero limbs too
    spec = re_ <= 8
    assert not rl.reshape(-1, p // 32)[spec].any()


@pytest.mark.parametrize("name", TS.KINDS)
def test_set_scalars_every_shape(name):
    """The conversions run on the neg kernel's shape for the precision: every such shape."""
    for p in (1024, 4096, 8192):
        vals = TS.edge_values(name, np.random.default_rng(7), extra=500)
        keep = F.get_variant("neg", p)
        try:
            for v, (k, l) in enumerate(F.variants("neg", p)):
                F.set_variant("neg", p, v)
                say(f"[gpu] {name} p={p} shape {k}x{l}")
                re_, rl = device_set_scalars(name, p, vals)
                oe, ol = O.set_scalar(name, p, vals)
                assert O.results_equal(re_, rl, oe, ol, p).all(), (p, k, l)
        finally:
            F.set_variant("neg", p, keep)


def test_set_scalars_count_zero_and_one():
    for name in TS.KINDS:
        R = F.FpaArray(0, 1024)
        assert F.set_scalars(R, name, np.zeros(0)) == F.SUCCESS
        R.close()
        vals = TS.edge_values(name, np.random.default_rng(3), extra=0)[3:4]
        re_, rl = device_set_scalars(name, 1024, vals)
        oe, ol = O.set_scalar(name, 1024, vals)
        assert O.results_equal(re_, rl, oe, ol, 1024).all()


@pytest.mark.parametrize("path", SCALAR_FIX, ids=lambda s: os.path.basename(s))
def test_set_scalar_and_neg_fixtures(path):
    """The reference generator's set_ui/si/float/double and neg cases with its MPFR gold."""
    z = np.load(path)
    p = int(z["precision"])
    name = os.path.basename(path)[4:].rsplit("_", 1)[0]
    say(f"[gpu] fixture {os.path.basename(path)}")
    if name == "neg":
        re_, rl = F.run_op("neg", p, z["a_exps"], z["a_limbs"])
    else:
        re_, rl = device_set_scalars(name, p, z["values"])
    assert O.results_equal(re_, rl, z["gold_exps"], z["gold_limbs"], p).all()
