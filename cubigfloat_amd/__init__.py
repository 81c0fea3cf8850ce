"""Host-side Python mirror of the fpa.h C API (ctypes over the in-tree libfpa.so).

The product is ``libfpa.so`` (C host layer + HIP kernels for gfx950); this module
only binds it.  It raises at import if the library has not been built -- there is no
fallback path.

Mirrors the reference's interface (fpa.h:43-63): ``FpaArray`` is an ``fpa_t``
(count values at one precision in HBM, natural limb layout); ``add/sub/mul/div/sqrt/
neg`` take (r, a[, b], mode) and return the ``fpa_result`` code like the C functions.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfpa.so")

ROUND_TRUNCATE, ROUND_UP, ROUND_DOWN, ROUND_TIES_TO_EVEN, ROUND_TIES_AWAY = range(5)
MODES = {"rz": 0, "ru": 1, "rd": 2, "rne": 3, "rna": 4}
SUCCESS, MATH_ERROR, INVALID_PRECISIONS, INVALID_COUNTS = range(4)
OPS = {"neg": 0, "add": 1, "sub": 2, "mul": 3, "div": 4, "sqrt": 5}
# variant-table ids: the element-wise ops plus the dot-chain kernel (FPA_KERNEL_DOT)
KERNELS = dict(OPS, dot=6)
PRECISIONS = (256, 512, 1024, 2048, 3072, 4096, 5120, 6144, 7168, 8192, 16384)
BIAS = 0x3FFFFFF3

if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() or make -C cubigfloat_amd")

_u32p = ctypes.POINTER(ctypes.c_uint32)


class FpaStruct(ctypes.Structure):
    """struct fpa_s (fpa.h): count, precision, gpu_exps, gpu_limbs."""
    _fields_ = [("count", ctypes.c_uint32), ("precision", ctypes.c_uint32),
                ("gpu_exps", ctypes.c_void_p), ("gpu_limbs", ctypes.c_void_p)]


class FpaArguments(ctypes.Structure):
    """fpa_arguments (fpa_kernels.h), same fields as the reference's fpa_internal.h:34-46."""
    _fields_ = [("count", ctypes.c_uint32), ("precision", ctypes.c_uint32),
                ("set_source_precision", ctypes.c_uint32), ("mode", ctypes.c_int),
                ("r_exps", ctypes.c_void_p), ("r_limbs", ctypes.c_void_p),
                ("a_exps", ctypes.c_void_p), ("a_limbs", ctypes.c_void_p),
                ("b_exps", ctypes.c_void_p), ("b_limbs", ctypes.c_void_p),
                ("values", ctypes.c_void_p)]


lib = ctypes.CDLL(LIB_PATH)
_P = ctypes.POINTER(FpaStruct)
for _name in ("fpa_init",):
    getattr(lib, _name).argtypes = [_P, ctypes.c_uint32, ctypes.c_uint32]
lib.fpa_clear.argtypes = [_P]
lib.fpa_copy_array_to_device.argtypes = [_P, _u32p, _u32p]
lib.fpa_copy_array_to_host.argtypes = [_P, _u32p, _u32p]
for _name in ("fpa_add", "fpa_sub", "fpa_mul", "fpa_div"):
    getattr(lib, _name).argtypes = [_P, _P, _P, ctypes.c_int]
lib.fpa_sqrt.argtypes = [_P, _P, ctypes.c_int]
lib.fpa_neg.argtypes = [_P, _P]
lib.fpa_set_ui.argtypes = [_P, ctypes.POINTER(ctypes.c_uint64)]
lib.fpa_set_si.argtypes = [_P, ctypes.POINTER(ctypes.c_int64)]
lib.fpa_set_float.argtypes = [_P, ctypes.POINTER(ctypes.c_float)]
lib.fpa_set_double.argtypes = [_P, ctypes.POINTER(ctypes.c_double)]
lib.fpa_kernel_launch.argtypes = [ctypes.c_int, ctypes.POINTER(FpaArguments), ctypes.c_int, ctypes.c_void_p]
lib.fpa_kernel_variants.argtypes = [ctypes.c_int, ctypes.c_uint32]
lib.fpa_kernel_variant_shape.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
lib.fpa_kernel_get_variant.argtypes = [ctypes.c_int, ctypes.c_uint32]
lib.fpa_kernel_set_variant.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.c_int]
lib.fpa_kernel_load_tuning.argtypes = [ctypes.c_char_p]
lib.fpa_set_stream.argtypes = [ctypes.c_void_p]
lib.fpa_get_stream.restype = ctypes.c_void_p
lib.fpa_precision_supported.argtypes = [ctypes.c_uint32]
lib.fpa_scalar_round_inc.argtypes = [ctypes.c_int] + [ctypes.c_uint32] * 4
lib.fpa_scalar_round_inc.restype = ctypes.c_uint32
lib.fpa_scalar_recip_3by2.argtypes = [ctypes.c_uint32] * 2
lib.fpa_scalar_recip_3by2.restype = ctypes.c_uint32
lib.fpa_scalar_div3by2.argtypes = [ctypes.c_uint32] * 6
lib.fpa_scalar_div3by2.restype = ctypes.c_uint32
lib.fpa_scalar_isqrt64.argtypes = [ctypes.c_uint64]
lib.fpa_scalar_isqrt64.restype = ctypes.c_uint32
lib.fpa_scalar_encode.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)]
lib.fpa_scalar_encode.restype = ctypes.c_uint32
lib.fpa_workload_generate.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_int, ctypes.c_uint32,
                                      ctypes.c_uint64, ctypes.c_uint32, _u32p, _u32p, _u32p, _u32p]


def _ptr(a):
    return None if a is None else a.ctypes.data_as(_u32p)


class FpaArray:
    """An fpa_t: `count` values of `precision` bits in device memory."""

    def __init__(self, count, precision):
        self.s = FpaStruct()
        rc = lib.fpa_init(ctypes.byref(self.s), count, precision)
        if rc != SUCCESS:
            raise ValueError(f"fpa_init({count}, {precision}) returned {rc}")
        self.count, self.precision, self.n = count, precision, precision // 32

    @property
    def ref(self):
        return ctypes.byref(self.s)

    def upload(self, exps, limbs):
        exps = np.ascontiguousarray(exps, dtype=np.uint32)
        limbs = np.ascontiguousarray(limbs, dtype=np.uint32)
        assert exps.size == self.count and limbs.size == self.count * self.n
        return lib.fpa_copy_array_to_device(self.ref, _ptr(exps), _ptr(limbs))

    def download(self):
        e = np.empty(self.count, np.uint32)
        l = np.empty(self.count * self.n, np.uint32)
        rc = lib.fpa_copy_array_to_host(self.ref, _ptr(e), _ptr(l))
        if rc != SUCCESS:
            raise RuntimeError(f"fpa_copy_array_to_host returned {rc}")
        return e, l

    def close(self):
        if self.s is not None:
            lib.fpa_clear(self.ref)
            self.s = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def add(r, a, b, mode=ROUND_TIES_TO_EVEN):
    return lib.fpa_add(r.ref, a.ref, b.ref, mode)


def sub(r, a, b, mode=ROUND_TIES_TO_EVEN):
    return lib.fpa_sub(r.ref, a.ref, b.ref, mode)


def mul(r, a, b, mode=ROUND_TIES_TO_EVEN):
    return lib.fpa_mul(r.ref, a.ref, b.ref, mode)


def div(r, a, b, mode=ROUND_TIES_TO_EVEN):
    return lib.fpa_div(r.ref, a.ref, b.ref, mode)


def sqrt(r, x, mode=ROUND_TIES_TO_EVEN):
    return lib.fpa_sqrt(r.ref, x.ref, mode)


def neg(r, a):
    return lib.fpa_neg(r.ref, a.ref)


lib.fpa_set.argtypes = [_P, _P, ctypes.c_int]
lib.fpa_sgn.argtypes = [ctypes.c_void_p, _P]
lib.fpa_cmp.argtypes = [ctypes.c_void_p, _P, _P]


class DeviceInt32:
    """A device int32 array (hipMalloc) for fpa_sgn / fpa_cmp results."""

    _hip = None

    def __init__(self, count):
        if DeviceInt32._hip is None:
            h = ctypes.CDLL("libamdhip64.so")
            h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
            h.hipFree.argtypes = [ctypes.c_void_p]
            h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
            DeviceInt32._hip = h
        self.count = count
        self.ptr = ctypes.c_void_p()
        assert DeviceInt32._hip.hipMalloc(ctypes.byref(self.ptr), max(4 * count, 4)) == 0

    def download(self):
        out = np.empty(self.count, np.int32)
        if self.count:
            assert DeviceInt32._hip.hipMemcpy(out.ctypes.data, self.ptr, 4 * self.count, 2) == 0  # device to host
        return out

    def close(self):
        if self.ptr:
            DeviceInt32._hip.hipFree(self.ptr)
            self.ptr = None


SCALAR_SETTERS = {"set_ui": (lib.fpa_set_ui, np.uint64, ctypes.c_uint64),
                  "set_si": (lib.fpa_set_si, np.int64, ctypes.c_int64),
                  "set_float": (lib.fpa_set_float, np.float32, ctypes.c_float),
                  "set_double": (lib.fpa_set_double, np.float64, ctypes.c_double)}


def set_scalars(r, name, values):
    """fpa_set_ui / fpa_set_si / fpa_set_float / fpa_set_double (name) of a host array."""
    fn, dt, ct = SCALAR_SETTERS[name]
    v = np.ascontiguousarray(values, dtype=dt)
    assert v.size == r.count
    return fn(r.ref, v.ctypes.data_as(ctypes.POINTER(ct)))


lib.fpa_set_scalars_noalloc.argtypes = [_P, ctypes.c_int, ctypes.c_void_p]
NOALLOC_MAX_VALUES = 1 << 19


def set_scalars_noalloc(r, name, values):
    """fpa_set_scalars_noalloc: the fpa_set_* conversion `name` with no device allocation."""
    _, dt, _ = SCALAR_SETTERS[name]
    v = np.ascontiguousarray(values, dtype=dt)
    assert v.size == r.count
    kind = ("set_ui", "set_si", "set_float", "set_double").index(name)
    return lib.fpa_set_scalars_noalloc(r.ref, kind, v.ctypes.data)


def scalar_encode(kind, bits):
    """(exponent word, left-aligned significand) of one fpa_set_* source (host helper)."""
    top = ctypes.c_uint64(0)
    w = lib.fpa_scalar_encode(kind, bits, ctypes.byref(top))
    return w, top.value


def set_precision(r, x, mode=ROUND_TIES_TO_EVEN):
    """fpa_set: r (its own precision) = x rounded in `mode`."""
    return lib.fpa_set(r.ref, x.ref, mode)


def sgn(x):
    """fpa_sgn into a device array; returns the int32 results on the host."""
    d = DeviceInt32(x.count)
    try:
        rc = lib.fpa_sgn(d.ptr, x.ref)
        if rc != SUCCESS:
            raise RuntimeError(f"fpa_sgn returned {rc}")
        return d.download()
    finally:
        d.close()


def cmp(a, b):
    """fpa_cmp into a device array; returns the int32 results on the host."""
    d = DeviceInt32(a.count)
    try:
        rc = lib.fpa_cmp(d.ptr, a.ref, b.ref)
        if rc != SUCCESS:
            raise RuntimeError(f"fpa_cmp returned {rc}")
        return d.download()
    finally:
        d.close()


def apply(op, r, a, b=None, mode=ROUND_TIES_TO_EVEN):
    """Run op by name ('add', 'mul', ...) through the fpa.h entry point."""
    if op in ("sqrt",):
        return sqrt(r, a, mode)
    if op == "neg":
        return neg(r, a)
    return {"add": add, "sub": sub, "mul": mul, "div": div}[op](r, a, b, mode)


def variants(op, precision):
    """[(K, L), ...] compiled kernel shapes for (op, precision)."""
    out = []
    for v in range(lib.fpa_kernel_variants(KERNELS[op], precision)):
        k, l = ctypes.c_int(), ctypes.c_int()
        lib.fpa_kernel_variant_shape(KERNELS[op], precision, v, ctypes.byref(k), ctypes.byref(l))
        out.append((k.value, l.value))
    return out


def set_variant(op, precision, variant):
    return lib.fpa_kernel_set_variant(KERNELS[op], precision, variant)


lib.fpa_kernel_set_small_variant.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.c_int, ctypes.c_uint32]
lib.fpa_kernel_get_variant_for_count.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.c_uint64]


def set_small_variant(op, precision, variant, max_count):
    """Variant for launches of at most max_count values (max_count 0 removes it)."""
    return lib.fpa_kernel_set_small_variant(KERNELS[op], precision, variant, max_count)


def get_variant_for_count(op, precision, count):
    return lib.fpa_kernel_get_variant_for_count(KERNELS[op], precision, count)


def get_variant(op, precision):
    return lib.fpa_kernel_get_variant(KERNELS[op], precision)


def generate(kind, seed, op, precision, count, first=0, binary=True):
    """Synthetic operands (fpa_workload_generate): returns (a_exps, a_limbs, b_exps, b_limbs)."""
    n = precision // 32
    ae = np.empty(count, np.uint32)
    al = np.empty(count * n, np.uint32)
    be = np.empty(count, np.uint32) if binary else None
    bl = np.empty(count * n, np.uint32) if binary else None
    rc = lib.fpa_workload_generate(kind, seed, OPS[op], precision, first, count,
                                   _ptr(ae), _ptr(al), _ptr(be), _ptr(bl))
    if rc != 0:
        raise ValueError(f"unknown workload kind {kind}")
    return ae, al, be, bl


def run_op(op, precision, a_exps, a_limbs, b_exps=None, b_limbs=None, mode=ROUND_TIES_TO_EVEN):
    """Upload, run one op through fpa.h, download.  Returns (r_exps, r_limbs)."""
    count = len(a_exps)
    A = FpaArray(count, precision)
    R = FpaArray(count, precision)
    B = None
    try:
        A.upload(a_exps, a_limbs)
        if b_exps is not None:
            B = FpaArray(count, precision)
            B.upload(b_exps, b_limbs)
        rc = apply(op, R, A, B, mode)
        if rc != SUCCESS:
            raise RuntimeError(f"fpa_{op} returned {rc}")
        return R.download()
    finally:
        for x in (A, B, R):
            if x is not None:
                x.close()


# ------------------------------ batches (fpa_batch.h) ------------------------------
_vp = ctypes.c_void_p
lib.fpa_batch_create.argtypes = [ctypes.POINTER(_vp), ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint8), _u32p,
                                 ctypes.POINTER(ctypes.c_uint8)]
lib.fpa_batch_set_operands.argtypes = [_vp, _u32p, _u32p, _u32p, _u32p]
lib.fpa_batch_create_branches.argtypes = [ctypes.POINTER(_vp), ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint8),
                                          _u32p, ctypes.POINTER(ctypes.c_uint8), ctypes.c_uint32]
lib.fpa_batch_create_persistent.argtypes = [ctypes.POINTER(_vp), ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint8),
                                            _u32p, ctypes.POINTER(ctypes.c_uint8)]
for _n in ("fpa_batch_tiles", "fpa_batch_grid"):
    getattr(lib, _n).argtypes = [_vp]
    getattr(lib, _n).restype = ctypes.c_uint32
lib.fpa_batch_is_persistent.argtypes = [_vp]
lib.fpa_batch_run.argtypes = [_vp, _vp]
lib.fpa_batch_run_plain.argtypes = [_vp, _vp]
lib.fpa_batch_get_results.argtypes = [_vp, _vp, _u32p, _u32p]
lib.fpa_batch_destroy.argtypes = [_vp]
for _n in ("fpa_batch_segments", "fpa_batch_branches"):
    getattr(lib, _n).argtypes = [_vp]
    getattr(lib, _n).restype = ctypes.c_uint32
lib.fpa_batch_blocks.argtypes = [_vp]
lib.fpa_batch_blocks.restype = ctypes.c_uint64
lib.fpa_batch_device_bytes.argtypes = [_vp]
lib.fpa_batch_device_bytes.restype = ctypes.c_size_t
lib.fpa_batch_total_limbs.argtypes = [_vp]
lib.fpa_batch_total_limbs.restype = ctypes.c_uint64
lib.fpa_batch_precision_supported.argtypes = [ctypes.c_uint32]
lib.fpa_dot_chains.argtypes = [_P, _P, _P, _u32p, ctypes.c_uint32, ctypes.c_int]
lib.fpa_dot_plan_create.argtypes = [ctypes.POINTER(_vp), _u32p, ctypes.c_uint32]
lib.fpa_dot_plan_run.argtypes = [_vp, _P, _P, _P, ctypes.c_int, _vp]
lib.fpa_dot_plan_destroy.argtypes = [_vp]
lib.fpa_dot_plan_time.argtypes = [_vp, _P, _P, _P, ctypes.c_int, ctypes.c_int, _vp, ctypes.POINTER(ctypes.c_float)]

BATCH_PRECISIONS = (256, 512, 1024, 2048, 4096, 8192, 16384)
PERSIST_PRECISIONS = (256, 512, 1024, 2048, 4096, 8192)
PERSIST_MAX_SEGMENTS = 40


class Batch:
    """A planned mixed-precision batch (fpa_batch_t): one captured fork/join graph, or with
    persistent=True one captured launch of the persistent tile kernel."""

    def __init__(self, ops, precisions, modes=None, branches=None, persistent=False):
        ops = np.ascontiguousarray([OPS[o] if isinstance(o, str) else o for o in ops], dtype=np.uint8)
        precisions = np.ascontiguousarray(precisions, dtype=np.uint32)
        modes = np.ascontiguousarray(np.full(len(ops), ROUND_TIES_TO_EVEN) if modes is None else modes, dtype=np.uint8)
        self.h = _vp()
        u8 = ctypes.POINTER(ctypes.c_uint8)
        if persistent:
            rc = lib.fpa_batch_create_persistent(ctypes.byref(self.h), len(ops), ops.ctypes.data_as(u8),
                                                 _ptr(precisions), modes.ctypes.data_as(u8))
        elif branches is None:
            rc = lib.fpa_batch_create(ctypes.byref(self.h), len(ops), ops.ctypes.data_as(u8), _ptr(precisions),
                                      modes.ctypes.data_as(u8))
        else:
            rc = lib.fpa_batch_create_branches(ctypes.byref(self.h), len(ops), ops.ctypes.data_as(u8),
                                               _ptr(precisions), modes.ctypes.data_as(u8), branches)
        if rc != SUCCESS:
            raise ValueError(f"fpa_batch_create returned {rc}")
        self.ops, self.precisions, self.nops = ops, precisions, len(ops)
        self.total_limbs = lib.fpa_batch_total_limbs(self.h)

    def set_operands(self, a_exps, a_limbs, b_exps=None, b_limbs=None):
        c = lambda x: None if x is None else np.ascontiguousarray(x, dtype=np.uint32)
        a_exps, a_limbs, b_exps, b_limbs = c(a_exps), c(a_limbs), c(b_exps), c(b_limbs)
        rc = lib.fpa_batch_set_operands(self.h, _ptr(a_exps), _ptr(a_limbs), _ptr(b_exps), _ptr(b_limbs))
        if rc != SUCCESS:
            raise ValueError(f"fpa_batch_set_operands returned {rc}")

    def run(self, stream=None):
        return lib.fpa_batch_run(self.h, stream)

    def run_plain(self, stream=None):
        return lib.fpa_batch_run_plain(self.h, stream)

    def results(self, stream=None):
        e = np.empty(self.nops, np.uint32)
        l = np.empty(self.total_limbs, np.uint32)
        lib.fpa_batch_get_results(self.h, stream, _ptr(e), _ptr(l))
        return e, l

    def info(self):
        d = {"segments": lib.fpa_batch_segments(self.h), "branches": lib.fpa_batch_branches(self.h),
             "blocks": lib.fpa_batch_blocks(self.h), "device_bytes": lib.fpa_batch_device_bytes(self.h)}
        if lib.fpa_batch_is_persistent(self.h):
            d.update(persistent=True, tiles=lib.fpa_batch_tiles(self.h), grid=lib.fpa_batch_grid(self.h))
        return d

    def close(self):
        if self.h:
            lib.fpa_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DotPlan:
    """fpa_dot_plan_t: chain offsets and longest-first order resident in device memory."""

    def __init__(self, lengths):
        self.lengths = np.ascontiguousarray(lengths, dtype=np.uint32)
        self.h = _vp()
        assert lib.fpa_dot_plan_create(ctypes.byref(self.h), _ptr(self.lengths), len(self.lengths)) == SUCCESS

    def run(self, r, x, y, mode=ROUND_TIES_TO_EVEN, stream=None):
        return lib.fpa_dot_plan_run(self.h, r.ref, x.ref, y.ref, mode, stream)

    def close(self):
        if self.h:
            lib.fpa_dot_plan_destroy(self.h)
            self.h = None


def dot_chains(precision, lengths, x_exps, x_limbs, y_exps, y_limbs, mode=ROUND_TIES_TO_EVEN):
    """Upload, run fpa_dot_chains, download. Returns (r_exps, r_limbs)."""
    lengths = np.ascontiguousarray(lengths, dtype=np.uint32)
    total = int(lengths.sum())
    X, Y, R = FpaArray(total, precision), FpaArray(total, precision), FpaArray(len(lengths), precision)
    try:
        if total:
            X.upload(x_exps, x_limbs)
            Y.upload(y_exps, y_limbs)
        rc = lib.fpa_dot_chains(R.ref, X.ref, Y.ref, _ptr(lengths), len(lengths), mode)
        if rc != SUCCESS:
            raise RuntimeError(f"fpa_dot_chains returned {rc}")
        return R.download()
    finally:
        for a in (X, Y, R):
            a.close()


# ------------------------------ server (fpa_serve.h) ------------------------------
class Request(ctypes.Structure):
    """fpa_request (fpa_serve.h)."""
    _fields_ = [("nops", ctypes.c_uint32), ("precision", ctypes.c_uint32),
                ("ops", ctypes.c_void_p), ("modes", ctypes.c_void_p),
                ("a_exps", ctypes.c_void_p), ("a_limbs", ctypes.c_void_p),
                ("b_exps", ctypes.c_void_p), ("b_limbs", ctypes.c_void_p),
                ("r_exps", ctypes.c_void_p), ("r_limbs", ctypes.c_void_p),
                ("arrival", ctypes.c_double), ("t_admit", ctypes.c_double), ("t_done", ctypes.c_double),
                ("status", ctypes.c_int32), ("slot", ctypes.c_uint32)]


lib.fpa_server_create.argtypes = [ctypes.POINTER(_vp), ctypes.c_size_t, ctypes.c_uint32, ctypes.c_size_t]
lib.fpa_server_destroy.argtypes = [_vp]
lib.fpa_server_submit.argtypes = [_vp, ctypes.POINTER(Request)]
lib.fpa_server_wait.argtypes = [_vp, ctypes.POINTER(Request)]
lib.fpa_server_run_trace.argtypes = [_vp, ctypes.POINTER(Request), ctypes.c_uint32]
lib.fpa_serve_plain_trace.argtypes = [ctypes.POINTER(Request), ctypes.c_uint32]
lib.fpa_request_bytes.argtypes = [ctypes.POINTER(Request)]
lib.fpa_request_bytes.restype = ctypes.c_size_t
lib.fpa_server_peak_bytes.argtypes = [_vp]
lib.fpa_server_peak_bytes.restype = ctypes.c_size_t
lib.fpa_server_budget_waits.argtypes = [_vp]
lib.fpa_server_budget_waits.restype = ctypes.c_uint64


class Trace:
    """A list of requests with host buffers kept alive for the C server.

    Each request: dict(precision, ops (uint8), modes (uint8), a_exps, a_limbs, b_exps,
    b_limbs, arrival)."""

    def __init__(self, reqs):
        self.reqs = reqs
        self.arr = (Request * len(reqs))()
        self.keep = []
        for i, q in enumerate(reqs):
            n = q["precision"] // 32
            nops = len(q["ops"])
            bufs = {k: np.ascontiguousarray(q[k], dtype=np.uint8 if k in ("ops", "modes") else np.uint32)
                    for k in ("ops", "modes", "a_exps", "a_limbs", "b_exps", "b_limbs")}
            bufs["r_exps"] = np.zeros(nops, np.uint32)
            bufs["r_limbs"] = np.zeros(nops * n, np.uint32)
            self.keep.append(bufs)
            r = self.arr[i]
            r.nops, r.precision, r.arrival = nops, q["precision"], float(q["arrival"])
            for k, v in bufs.items():
                setattr(r, k, v.ctypes.data)

    def __len__(self):
        return len(self.reqs)

    def results(self, i):
        return self.keep[i]["r_exps"], self.keep[i]["r_limbs"]

    def times(self):
        a = np.array([r.arrival for r in self.arr])
        s = np.array([r.t_admit for r in self.arr])
        d = np.array([r.t_done for r in self.arr])
        st = np.array([r.status for r in self.arr])
        return a, s, d, st

    def reset(self):
        for r in self.arr:
            r.status, r.t_admit, r.t_done = -1, 0.0, 0.0
        for b in self.keep:
            b["r_exps"][:] = 0
            b["r_limbs"][:] = 0


class Server:
    """fpa_server_t: resident pool of budget_bytes, `slots` streams."""

    def __init__(self, budget_bytes, slots=8, max_request_bytes=0):
        self.h = _vp()
        rc = lib.fpa_server_create(ctypes.byref(self.h), budget_bytes, slots, max_request_bytes)
        if rc != SUCCESS:
            raise ValueError(f"fpa_server_create returned {rc}")

    def run_trace(self, trace):
        return lib.fpa_server_run_trace(self.h, trace.arr, len(trace))

    def peak_bytes(self):
        return lib.fpa_server_peak_bytes(self.h)

    def budget_waits(self):
        return lib.fpa_server_budget_waits(self.h)

    def close(self):
        if self.h:
            lib.fpa_server_destroy(self.h)
            self.h = None


def serve_plain(trace):
    return lib.fpa_serve_plain_trace(trace.arr, len(trace))


def request_bytes(trace, i):
    return lib.fpa_request_bytes(ctypes.byref(trace.arr[i]))
