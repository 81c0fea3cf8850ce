"""ORACLE -- test infrastructure only.

ctypes bindings for the oracle libraries built by ``oracle/Makefile``:

* ``liboracle.so`` -- the exact-integer restatement (fpa_oracle.c), the MPFR gold
  and the MPFR-backed ``fpa.h`` CPU backend (fpa_mpfr.c);
* ``_ref/libreftests.so`` -- the reference's own MPFR test harness (tests.c) compiled
  from its sources, used to draw the reference's generator cases with their gold.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg import
this module.  Arrays are numpy uint32 in the natural layout (value i's limbs at
``limbs[i*n:(i+1)*n]``, little endian), as the reference's host API uses
(fpa_management.h:17-47).
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "_build", "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libreftests.so")

OPS = {"neg": 0, "add": 1, "sub": 2, "mul": 3, "div": 4, "sqrt": 5}
MODES = {"rz": 0, "ru": 1, "rd": 2, "rne": 3, "rna": 4}
BIAS = 0x3FFFFFF3

_u32p = ctypes.POINTER(ctypes.c_uint32)
_lib = None
_ref = None


def build():
    """Compile the oracle (and the reference harness where its sources exist)."""
    subprocess.run(["make", "-C", HERE, "-s"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            build()
        _lib = ctypes.CDLL(ORACLE_SO)
        _lib.orc_op_array.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int,
                                      _u32p, _u32p, _u32p, _u32p, _u32p, _u32p]
        _lib.mpfr_gold_array.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int,
                                         _u32p, _u32p, ctypes.POINTER(ctypes.c_uint8),
                                         _u32p, _u32p, _u32p, _u32p, ctypes.c_int]
        _lib.orc_action.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32]
        _lib.orc_round_decision.argtypes = [ctypes.c_int] * 5
        _lib.fpa_cpu_mpfr_version.restype = ctypes.c_char_p
        _lib.fpa_cpu_struct_size.restype = ctypes.c_size_t
        _lib.fpa_cpu_op.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
        _lib.fpa_cpu_init.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32]
        _lib.fpa_cpu_clear.argtypes = [ctypes.c_void_p]
        _lib.fpa_cpu_copy_in.argtypes = [ctypes.c_void_p, _u32p, _u32p]
        _lib.fpa_cpu_copy_out.argtypes = [ctypes.c_void_p, _u32p, _u32p, ctypes.c_int]
        _lib.fpa_cpu_set_threads.argtypes = [ctypes.c_int]
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(_u32p)


def _c(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.uint32)


def op(name, precision, a_exps, a_limbs, b_exps=None, b_limbs=None, mode="rne"):
    """Exact-integer oracle over arrays; returns (r_exps, r_limbs)."""
    a_exps, a_limbs, b_exps, b_limbs = _c(a_exps), _c(a_limbs), _c(b_exps), _c(b_limbs)
    count = len(a_exps)
    n = precision // 32
    r_exps = np.zeros(count, np.uint32)
    r_limbs = np.zeros(count * n, np.uint32)
    lib().orc_op_array(OPS[name], count, precision, MODES[mode], _p(r_exps), _p(r_limbs),
                       _p(a_exps), _p(a_limbs), _p(b_exps), _p(b_limbs))
    return r_exps, r_limbs


def mpfr_gold(name, precision, a_exps, a_limbs, b_exps=None, b_limbs=None, mode="rne", threads=0):
    """MPFR gold; returns (r_exps, r_limbs, ok) where ok marks elements MPFR could evaluate."""
    a_exps, a_limbs, b_exps, b_limbs = _c(a_exps), _c(a_limbs), _c(b_exps), _c(b_limbs)
    count = len(a_exps)
    n = precision // 32
    r_exps = np.zeros(count, np.uint32)
    r_limbs = np.zeros(count * n, np.uint32)
    ok = np.zeros(count, np.uint8)
    lib().mpfr_gold_array(OPS[name], count, precision, MODES[mode], _p(r_exps), _p(r_limbs),
                          ok.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                          _p(a_exps), _p(a_limbs), _p(b_exps), _p(b_limbs), threads)
    return r_exps, r_limbs, ok.astype(bool)


def action(op_name, a_exp, b_exp=0):
    return lib().orc_action(OPS[op_name], a_exp, b_exp)


def round_decision(mode, sign, lsb, round_bit, sticky):
    return lib().orc_round_decision(MODES[mode], sign, lsb, round_bit, sticky)


def results_equal(ra, la, rb, lb, precision):
    """Per-element equality: exponent words equal, limbs equal when the word is a value."""
    n = precision // 32
    ra = np.asarray(ra, np.uint32)
    rb = np.asarray(rb, np.uint32)
    la = np.asarray(la, np.uint32).reshape(-1, n)
    lb = np.asarray(lb, np.uint32).reshape(-1, n)
    eq = ra == rb
    val = ra > 8
    eq &= ~val | np.all(la == lb, axis=1)
    return eq


class CpuArray:
    """MPFR-backed fpa.h CPU backend array (restates the -DCPU struct, fpa.h:28-34)."""

    def __init__(self, count, precision):
        self.buf = ctypes.create_string_buffer(lib().fpa_cpu_struct_size())
        self.count, self.precision = count, precision
        lib().fpa_cpu_init(self.buf, count, precision)

    def load(self, exps, limbs):
        lib().fpa_cpu_copy_in(self.buf, _p(_c(exps)), _p(_c(limbs)))

    def store(self, op_name):
        n = self.precision // 32
        e = np.zeros(self.count, np.uint32)
        l = np.zeros(self.count * n, np.uint32)
        lib().fpa_cpu_copy_out(self.buf, _p(e), _p(l), OPS[op_name])
        return e, l

    def close(self):
        if self.buf is not None:
            lib().fpa_cpu_clear(self.buf)
            self.buf = None


def cpu_op(name, r, a, b=None, mode="rne"):
    return lib().fpa_cpu_op(OPS[name], r.buf, a.buf, None if b is None else b.buf, MODES[mode])


def cpu_set_threads(t):
    lib().fpa_cpu_set_threads(int(t))


# ------------------------- reference harness (tests.c) -------------------------
class _Tests(ctypes.Structure):
    # tests.h:10-29
    _fields_ = [("test_type", ctypes.c_int), ("count", ctypes.c_uint32), ("precision", ctypes.c_uint32),
                ("set_source_precision", ctypes.c_uint32), ("rounding", ctypes.c_int),
                ("current", ctypes.c_uint32),
                ("gold_exps", _u32p), ("gold_limbs", _u32p), ("r_exps", _u32p), ("r_limbs", _u32p),
                ("a_exps", _u32p), ("a_limbs", _u32p), ("b_exps", _u32p), ("b_limbs", _u32p),
                ("values", ctypes.c_void_p)]


def ref_available():
    return os.path.exists(REF_SO)


def ref_lib():
    global _ref
    if _ref is None:
        _ref = ctypes.CDLL(REF_SO)
        for name in ("add", "sub", "mul", "div", "sqrt", "neg", "sgn", "cmp",
                     "set_ui", "set_si", "set_float", "set_double"):
            f = getattr(_ref, "tests_generate_" + name)
            f.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
            f.restype = ctypes.POINTER(_Tests)
        _ref.tests_generate_set.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]
        _ref.tests_generate_set.restype = ctypes.POINTER(_Tests)
        _ref.tests_free.argtypes = [ctypes.POINTER(_Tests)]
    return _ref


def ref_generate(name, count, precision, source_precision=None):
    """Draw `count` cases from the reference's own generator (tests.c:580-925).

    Returns dict with a_exps, a_limbs, (b_exps, b_limbs,) gold_exps, (gold_limbs).  For
    "set" the a operands have source_precision; for "sgn"/"cmp" gold_exps holds the
    int32 results (mpfr_sgn / mpfr_cmp) and there are no gold limbs.
    The generator's GMP random state persists inside the loaded library, so successive
    calls continue the reference's deterministic sequence.
    """
    if name == "set":
        t = ref_lib().tests_generate_set(count, precision, source_precision)
    else:
        t = getattr(ref_lib(), "tests_generate_" + name)(count, precision)
    if not t or t.contents.count != count or t.contents.precision != precision or t.contents.current != count:
        raise RuntimeError("reference harness returned an unusable tests_t")
    n = precision // 32
    na = (source_precision if name == "set" else precision) // 32
    s = t.contents

    def arr(ptr, m):
        return np.ctypeslib.as_array(ptr, shape=(m,)).copy()

    if name in SCALAR_KINDS:
        # tests_add_mpfr_ui64/float/double (tests.c:225-283): values and gold only
        ctype = {"set_ui": ctypes.c_uint64, "set_si": ctypes.c_int64, "set_float": ctypes.c_float,
                 "set_double": ctypes.c_double}[name]
        vals = np.ctypeslib.as_array(ctypes.cast(s.values, ctypes.POINTER(ctype)), shape=(count,)).copy()
        out = {"values": vals, "gold_exps": arr(s.gold_exps, count), "gold_limbs": arr(s.gold_limbs, count * n)}
        ref_lib().tests_free(t)
        out["gold_limbs"].reshape(count, -1)[out["gold_exps"] == 2] = 0
        return out

    out = {"a_exps": arr(s.a_exps, count), "a_limbs": arr(s.a_limbs, count * na),
           "gold_exps": arr(s.gold_exps, count)}
    if name in ("sgn", "cmp"):
        out["gold_exps"] = out["gold_exps"].view(np.int32)
    else:
        out["gold_limbs"] = arr(s.gold_limbs, count * n)
    if bool(s.b_exps):
        out["b_exps"] = arr(s.b_exps, count)
        out["b_limbs"] = arr(s.b_limbs, count * n)
    ref_lib().tests_free(t)
    # tests_set_mpfr leaves the limbs of zeros unset (tests.c:142-145): clear them
    for k in ("a", "b", "gold"):
        if k + "_limbs" in out:
            z = out[k + "_exps"] == 2
            out[k + "_limbs"].reshape(len(z), -1)[z] = 0
    return out


def dot(precision, lengths, x_exps, x_limbs, y_exps, y_limbs, mode="rne"):
    """Oracle dot chains (acc <- RN(acc + RN(x*y)) in index order); returns (r_exps, r_limbs)."""
    lengths = np.ascontiguousarray(lengths, dtype=np.uint32)
    count = len(lengths)
    n = precision // 32
    r_exps = np.zeros(count, np.uint32)
    r_limbs = np.zeros(count * n, np.uint32)
    f = lib().orc_dot_array
    f.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, _u32p, _u32p, _u32p, _u32p, _u32p, _u32p, _u32p]
    f(count, precision, MODES[mode], _p(lengths), _p(_c(x_exps)), _p(_c(x_limbs)), _p(_c(y_exps)),
      _p(_c(y_limbs)), _p(r_exps), _p(r_limbs))
    return r_exps, r_limbs


def mpfr_dot(precision, lengths, x_exps, x_limbs, y_exps, y_limbs, mode="rne", threads=0):
    """MPFR gold for dot chains (the same mul-then-add sequence); returns (r_exps, r_limbs)."""
    lengths = np.ascontiguousarray(lengths, dtype=np.uint32)
    count = len(lengths)
    offs = np.zeros(count + 1, np.uint64)
    offs[1:] = np.cumsum(lengths, dtype=np.uint64)
    n = precision // 32
    r_exps = np.zeros(count, np.uint32)
    r_limbs = np.zeros(count * n, np.uint32)
    f = lib().mpfr_dot_array
    f.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, _u32p, ctypes.POINTER(ctypes.c_uint64),
                  _u32p, _u32p, _u32p, _u32p, _u32p, _u32p, ctypes.c_int]
    f(count, precision, MODES[mode], _p(lengths), offs.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
      _p(_c(x_exps)), _p(_c(x_limbs)), _p(_c(y_exps)), _p(_c(y_limbs)), _p(r_exps), _p(r_limbs), threads)
    return r_exps, r_limbs


# ----------------------------- set / sgn / cmp -----------------------------
SCALAR_KINDS = {"set_ui": 0, "set_si": 1, "set_float": 2, "set_double": 3}
SCALAR_DTYPES = {"set_ui": np.uint64, "set_si": np.int64, "set_float": np.float32, "set_double": np.float64}


def set_scalar(name, precision, values):
    """fpa_set_ui/si/float/double restated (exact); returns (r_exps, r_limbs)."""
    values = np.ascontiguousarray(values, dtype=SCALAR_DTYPES[name])
    count = len(values)
    re = np.zeros(count, np.uint32)
    rl = np.zeros(count * (precision // 32), np.uint32)
    f = lib().orc_set_scalar_array
    f.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p, _u32p, _u32p]
    f(SCALAR_KINDS[name], count, precision, values.ctypes.data, _p(re), _p(rl))
    return re, rl


def sgn(x_exps):
    x_exps = _c(x_exps)
    r = np.zeros(len(x_exps), np.int32)
    f = lib().orc_sgn_array
    f.argtypes = [ctypes.c_uint32, _u32p, ctypes.POINTER(ctypes.c_int32)]
    f(len(x_exps), _p(x_exps), r.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
    return r


def cmp(precision, a_exps, a_limbs, b_exps, b_limbs):
    a_exps, a_limbs, b_exps, b_limbs = _c(a_exps), _c(a_limbs), _c(b_exps), _c(b_limbs)
    r = np.zeros(len(a_exps), np.int32)
    f = lib().orc_cmp_array
    f.argtypes = [ctypes.c_uint32, ctypes.c_uint32, _u32p, _u32p, _u32p, _u32p, ctypes.POINTER(ctypes.c_int32)]
    f(len(a_exps), precision, _p(a_exps), _p(a_limbs), _p(b_exps), _p(b_limbs),
      r.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
    return r


def set_precision(r_precision, x_precision, x_exps, x_limbs, mode="rne"):
    x_exps, x_limbs = _c(x_exps), _c(x_limbs)
    count = len(x_exps)
    re = np.zeros(count, np.uint32)
    rl = np.zeros(count * (r_precision // 32), np.uint32)
    f = lib().orc_set_array
    f.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, _u32p, _u32p, _u32p, _u32p]
    f(count, r_precision, x_precision, MODES[mode], _p(re), _p(rl), _p(x_exps), _p(x_limbs))
    return re, rl


def mpfr_set(r_precision, x_precision, x_exps, x_limbs, mode="rne"):
    """mpfr_set gold; returns (r_exps, r_limbs, ok)."""
    x_exps, x_limbs = _c(x_exps), _c(x_limbs)
    count = len(x_exps)
    re = np.zeros(count, np.uint32)
    rl = np.zeros(count * (r_precision // 32), np.uint32)
    ok = np.zeros(count, np.uint8)
    f = lib().mpfr_set_array
    f.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, _u32p, _u32p,
                  ctypes.POINTER(ctypes.c_uint8), _u32p, _u32p]
    f(count, r_precision, x_precision, MODES[mode], _p(re), _p(rl), ok.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
      _p(x_exps), _p(x_limbs))
    return re, rl, ok.astype(bool)


def mpfr_cmp(precision, a_exps, a_limbs, b_exps=None, b_limbs=None):
    """mpfr_cmp (or mpfr_sgn when b is None) clamped to -1/0/1; returns (r, ok)."""
    a_exps, a_limbs = _c(a_exps), _c(a_limbs)
    count = len(a_exps)
    r = np.zeros(count, np.int32)
    ok = np.zeros(count, np.uint8)
    f = lib().mpfr_cmp_array
    f.argtypes = [ctypes.c_uint32, ctypes.c_uint32, _u32p, _u32p, _u32p, _u32p, ctypes.POINTER(ctypes.c_int32),
                  ctypes.POINTER(ctypes.c_uint8)]
    f(count, precision, _p(a_exps), _p(a_limbs), _p(_c(b_exps)) if b_exps is not None else None,
      _p(_c(b_limbs)) if b_limbs is not None else None, r.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
      ok.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)))
    return r, ok.astype(bool)


def sgn_table(cls):
    f = lib().orc_sgn_table
    f.argtypes = [ctypes.c_uint32]
    return f(cls)


def cmp_table(ca, cb):
    f = lib().orc_cmp_table
    f.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
    return f(ca, cb)
