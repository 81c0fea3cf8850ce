"""CPU tests of the C ABI: the library loads, exports every declared symbol, and its host
logic (dispatch tables, argument validation, scalar helpers, workload generator) behaves."""
import ctypes
import math
import os
import random
import re

import numpy as np
import pytest

import cubigfloat_amd as F

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not h.endswith(".h"):
            continue
        src = open(os.path.join(ROOT, "include", h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"^\s*[A-Za-z_][\w\s\*]*?\b(fpa_\w+)\s*\(", src, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_every_declared_symbol_is_exported():
    names = declared_functions()
    assert "fpa_mul" in names and "fpa_kernel_launch" in names
    missing = [n for n in names if not hasattr(F.lib, n)]
    assert not missing, missing


def test_reference_api_symbols_present():
    # fpa.h:43-63 of the reference
    for n in ["fpa_init", "fpa_clear", "fpa_copy_array_to_device", "fpa_copy_array_to_host", "fpa_set_ui",
              "fpa_set_si", "fpa_set_float", "fpa_set_double", "fpa_set", "fpa_sgn", "fpa_cmp", "fpa_neg",
              "fpa_add", "fpa_sub", "fpa_mul", "fpa_div", "fpa_sqrt"]:
        assert hasattr(F.lib, n), n


def test_struct_layout_matches_reference():
    # struct fpa_s {uint32 count; uint32 precision; uint32 *gpu_exps; uint32 *gpu_limbs;}
    assert ctypes.sizeof(F.FpaStruct) == 24
    assert F.FpaStruct.gpu_exps.offset == 8 and F.FpaStruct.gpu_limbs.offset == 16
    assert ctypes.sizeof(F.FpaArguments) == 72


def test_precision_validation_without_device():
    s = F.FpaStruct()
    for bad in (0, 32, 1000, 1536, 9216, 32768):
        assert F.lib.fpa_init(ctypes.byref(s), 4, bad) == F.INVALID_PRECISIONS
    for p in F.PRECISIONS:
        assert F.lib.fpa_precision_supported(p) == 1


def test_batch_validation_without_device():
    u8 = ctypes.POINTER(ctypes.c_uint8)
    h = ctypes.c_void_p()
    def create(ops, precs, modes, branches=4):
        o = np.array(ops, np.uint8)
        p = np.array(precs, np.uint32)
        m = np.array(modes, np.uint8)
        return F.lib.fpa_batch_create_branches(ctypes.byref(h), len(o), o.ctypes.data_as(u8),
                                               p.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                               m.ctypes.data_as(u8), branches)
    assert create([F.OPS["mul"]], [1536], [3]) == F.INVALID_PRECISIONS
    assert create([F.OPS["neg"]], [256], [3]) == F.MATH_ERROR        # neg is not a batch op
    assert create([F.OPS["add"]], [256], [5]) == F.MATH_ERROR        # no sixth rounding mode
    assert create([F.OPS["add"]], [256], [3], branches=0) == F.INVALID_COUNTS
    assert create([F.OPS["add"]], [256], [3], branches=9) == F.INVALID_COUNTS
    for p in F.PRECISIONS:
        assert F.lib.fpa_batch_precision_supported(p) == 1


def test_variant_tables():
    for op in F.KERNELS:
        for p in F.PRECISIONS:
            vs = F.variants(op, p)
            assert vs, (op, p)
            for k, l in vs:
                assert k * l == p // 32 and 64 % l == 0 and l <= 32
            assert len(set(vs)) == len(vs)
            assert 0 <= F.get_variant(op, p) < len(vs)
    keep = F.get_variant("mul", 4096)
    assert F.set_variant("mul", 4096, 99) == -1
    assert F.set_variant("mul", 4096, 1) == 0 and F.get_variant("mul", 4096) == 1
    F.set_variant("mul", 4096, keep)


def test_tuning_file(tmp_path):
    f = tmp_path / "tune.txt"
    keep = F.get_variant("mul", 4096)
    k, l = F.variants("mul", 4096)[2]
    f.write_text(f"# op precision K L\n3 4096 {k} {l}\n3 4096 3 3\n")
    assert F.lib.fpa_kernel_load_tuning(str(f).encode()) == 1
    assert F.get_variant("mul", 4096) == 2
    F.set_variant("mul", 4096, keep)
    assert F.lib.fpa_kernel_load_tuning(b"/nonexistent/file") == -1


def test_load_time_tuning_matches_committed_table():
    """The library applies tuning.txt beside it when loaded; every line of the committed
    table names a compiled shape, so every line is applied."""
    path = os.path.join(ROOT, "cubigfloat_amd", "tuning.txt")
    if not os.path.exists(path):
        pytest.skip("no tuning table committed")
    lines = [ln.split() for ln in open(path) if ln.strip() and not ln.startswith("#")]
    if os.environ.get("FPA_TUNING_FILE") is None:
        assert F.lib.fpa_kernel_tuning_applied() == len(lines)
    for f in lines:
        assert len(f) in (4, 5)
        op, prec, k, l = (int(x) for x in f[:4])
        name = {v: n for n, v in F.KERNELS.items()}[op]
        assert (k, l) in F.variants(name, prec)


def test_small_launch_variants():
    keep = F.get_variant("div", 4096), F.get_variant_for_count("div", 4096, 100)
    nv = len(F.variants("div", 4096))
    try:
        F.set_variant("div", 4096, 0)
        assert F.get_variant_for_count("div", 4096, 100) == 0
        assert F.set_small_variant("div", 4096, nv - 1, 5000) == 0
        assert F.get_variant_for_count("div", 4096, 5000) == nv - 1
        assert F.get_variant_for_count("div", 4096, 5001) == 0
        assert F.get_variant("div", 4096) == 0
        assert F.set_small_variant("div", 4096, nv, 5000) == -1      # no such variant
        # tiers: a launch takes the smallest bound that holds it; a repeated bound replaces
        assert F.set_small_variant("div", 4096, 1, 2000) == 0
        assert F.set_small_variant("div", 4096, 2 % nv, 9000) == 0
        assert [F.get_variant_for_count("div", 4096, c) for c in (1, 2000, 2001, 5000, 5001, 9000, 9001)] == \
            [1, 1, nv - 1, nv - 1, 2 % nv, 2 % nv, 0]
        assert F.set_small_variant("div", 4096, 0, 2000) == 0
        assert F.get_variant_for_count("div", 4096, 1500) == 0
        assert F.set_small_variant("div", 4096, 0, 100) == 0        # fourth tier
        assert F.set_small_variant("div", 4096, 0, 200) == -1       # a fifth does not fit
        assert F.set_small_variant("div", 4096, 0, 0) == 0          # bound 0 removes every tier
        assert F.get_variant_for_count("div", 4096, 100) == 0 and F.get_variant_for_count("div", 4096, 5000) == 0
        # an explicit override applies to every launch size
        assert F.set_small_variant("div", 4096, nv - 1, 5000) == 0
        assert F.set_variant("div", 4096, 1) == 0
        assert F.get_variant_for_count("div", 4096, 10) == 1
    finally:
        # restore the committed table (its tiers included)
        F.set_variant("div", 4096, keep[0])
        F.lib.fpa_kernel_load_tuning(os.path.join(ROOT, "cubigfloat_amd", "tuning.txt").encode())


def test_scalar_helpers_vs_python():
    rnd = random.Random(1)
    alph = [0, 1, 0x7FFFFFFF, 0x80000000, 0xFFFFFFFF, 0xFFFFFFFE]
    pick = lambda: rnd.choice(alph) if rnd.random() < 0.5 else rnd.getrandbits(32)
    for _ in range(20000):
        d1, d0 = pick() | 0x80000000, pick()
        D = (d1 << 32) | d0
        v = F.lib.fpa_scalar_recip_3by2(d1, d0)
        assert v == ((1 << 96) - 1) // D - (1 << 32)
        u2 = rnd.randrange(0, d1 + 1)
        u1, u0 = pick(), pick()
        if (u2 << 32 | u1) >= D:
            continue
        U = (u2 << 64) | (u1 << 32) | u0
        assert F.lib.fpa_scalar_div3by2(u2, u1, u0, d1, d0, v) == U // D
        x = rnd.getrandbits(64)
        assert F.lib.fpa_scalar_isqrt64(x) == math.isqrt(x)
    for m in range(5):
        for s in (0, 1):
            for lsb in (0, 1):
                for rw in (0, 1, 0x7FFFFFFF, 0x80000000, 0x80000001, 0xFFFFFFFF):
                    for st in (0, 1):
                        from oracle import oracle as O
                        exp = O.round_decision(["rz", "ru", "rd", "rne", "rna"][m], s, lsb, rw >> 31,
                                               int((rw & 0x7FFFFFFF) != 0 or st))
                        assert F.lib.fpa_scalar_round_inc(m, s, lsb, rw, st) == exp


def test_workload_generator_properties():
    p = 4096
    n = p // 32
    ae, al, be, bl = F.generate(2, 2, "mul", p, 20000)
    vals = ae > 8
    assert (al.reshape(-1, n)[vals][:, -1] >= 0x80000000).all()
    e = (ae[vals].astype(np.int64) >> 1) - F.BIAS
    assert e.min() >= -1024 and e.max() <= 1025
    zero_rate = (ae == 2).mean()
    assert 0.005 < zero_rate < 0.015
    # regenerating a sub-range reproduces the same elements
    ae2, al2, _, _ = F.generate(2, 2, "mul", p, 100, first=500)
    assert (ae2 == ae[500:600]).all() and (al2 == al[500 * n:600 * n]).all()
    # near-cancellation pairs: |b| within one binade of |a| with opposite sign for add
    ae, al, be, bl = F.generate(2, 4, "add", p, 20000)
    both = (ae > 8) & (be > 8)
    close = both & (np.abs((ae.astype(np.int64) >> 1) - (be.astype(np.int64) >> 1)) <= 1) & ((ae ^ be) & 1 == 1)
    assert 0.03 < close.mean() < 0.08


def test_autotune_table_round_trip(tmp_path):
    """A table written by the autotuner is read back by the library's loader."""
    from cubigfloat_amd import autotune as T
    keep = {(op, p): F.get_variant(op, p) for op in ("mul", "div", "dot") for p in (1024, 4096)}
    table = {}
    for (op, p) in keep:
        vs = F.variants(op, p)
        pick = vs[-1]
        table[(op, p)] = (pick, [(k, l, 1.0 + i) for i, (k, l) in enumerate(vs)], 1234)
    f = tmp_path / "t.txt"
    dv = F.variants("div", 4096)
    small = {("div", 4096): [(dv[0], [(k, l, 1.0) for k, l in dv], 700, 777),
                             (dv[1], [(k, l, 1.0) for k, l in dv], 2800, 3000)]}
    assert T.tier_bounds([2048, 4096, 8192, 16384]) == [2896, 5792, 11585, 23170]
    T.write_table(str(f), table, small)
    try:
        assert F.lib.fpa_kernel_load_tuning(str(f).encode()) == len(table) + 2
        for (op, p) in keep:
            assert F.get_variant(op, p) == len(F.variants(op, p)) - 1
        assert F.get_variant_for_count("div", 4096, 777) == 0
        assert F.get_variant_for_count("div", 4096, 778) == 1
        assert F.get_variant_for_count("div", 4096, 3000) == 1
        assert F.get_variant_for_count("div", 4096, 3001) == len(dv) - 1
    finally:
        for (op, p), v in keep.items():
            F.set_variant(op, p, v)
    assert T.tune_count("add", 256) == 1 << 22
    assert T.tune_count("div", 16384) == (256 << 20) // (4 * 513) // 4
