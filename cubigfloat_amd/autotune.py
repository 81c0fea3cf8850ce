"""Autotuner of the kernel shape per (op, precision).

Each op at each precision has several compiled shapes (K limbs per lane, L lanes per
value, K*L = precision/32).  For every (op, precision) this times every shape on the
same synthetic operands (fpa_workload_generate kind 2: e ~ U[-1024, 1024], 1% zeros, 5%
near-cancellation pairs), checks that every shape returns bit-identical results, and
writes the fastest shape as one "op precision K L" line (op 6 = the dot-chain kernel,
timed on 16,384 chains of U[8, 512] elements, the size of the 2048-bit dot workload, with
shorter chains above 2048 bits; its tiers count chains, 512 and 2,048 of them).  It then times every shape again on small
launches (2,048 to 16,384 values by default, the sizes of the segments of a 2^18-op mixed
batch or of a served request): such a launch cannot fill the chip and the shape with the
shortest per-value latency wins, a different one at different sizes.  Where a size's
winner differs from the large-launch shape, "op precision K L max_count" lines (a fifth
field) give a shape per size tier: a launch of at most max_count values takes the
smallest tier that holds it.  A tier timed at c values reaches to the geometric middle
between c and the next size timed (c*sqrt(2) for the largest).  libfpa.so applies the file
when it is loaded (tuning.txt beside the library, or $FPA_TUNING_FILE), so the fpa.h
calls, the batch graphs, the server and the dot chains all use the tuned shapes.

The reference picks one fixed shape per size in its precision switch
(fpa_dispatch.h:294-321); this replaces that static choice.

    python -m cubigfloat_amd.autotune --out cubigfloat_amd/tuning.txt
"""
import argparse
import ctypes
import sys
import time

import numpy as np

from . import (DotPlan, FpaArguments, FpaArray, KERNELS, OPS, PRECISIONS, generate, get_variant, lib,
               set_small_variant, set_variant, variants)

lib.fpa_kernel_time.argtypes = [ctypes.c_int, ctypes.POINTER(FpaArguments), ctypes.c_int, ctypes.c_int,
                                ctypes.c_void_p, ctypes.POINTER(ctypes.c_float)]
lib.fpa_kernel_time.restype = ctypes.c_int
lib.fpa_kernel_tuning_applied.restype = ctypes.c_int

TUNED_OPS = ("add", "sub", "mul", "div", "sqrt", "neg", "dot")


def tune_count(op, precision, operand_bytes=256 << 20):
    """Values per timed launch: one operand array of about operand_bytes, a quarter of
    that for the long-running div/sqrt, at least 4096 and at most 2^22 values."""
    per_value = 4 * (precision // 32 + 1)
    c = operand_bytes // per_value
    if op in ("div", "sqrt"):
        c //= 4
    return int(min(max(c, 4096), 1 << 22))


def time_variant(op, precision, args, variant, target_ms=20.0):
    ms = ctypes.c_float()
    rc = lib.fpa_kernel_time(OPS[op], ctypes.byref(args), variant, 1, None, ctypes.byref(ms))
    if rc != 0:
        raise RuntimeError(f"fpa_kernel_time({op}, {precision}, variant {variant}) returned {rc}")
    reps = int(min(50, max(3, target_ms / max(ms.value, 1e-3))))
    rc = lib.fpa_kernel_time(OPS[op], ctypes.byref(args), variant, reps, None, ctypes.byref(ms))
    if rc != 0:
        raise RuntimeError(f"fpa_kernel_time({op}, {precision}, variant {variant}) returned {rc}")
    return float(ms.value)


DOT_CHAINS = 16384
DOT_SMALL_CHAINS = (512, 2048)


def tune_dot(precision, count=None, seed=1, log=None):
    """Dot chains (fpa_dot_plan_run): `count` chains (default DOT_CHAINS) of U[8, m]
    elements, m = 512 up to 2048 bits and 512*2048/precision above (at least 16), so an
    operand array stays near 1 GB.  A dot kernel keeps every chain of a wave in registers
    for the wave's whole life, so how many chains there are (how many waves, and whether
    they are all resident at once) decides the shape as much as the precision does: the
    chain count is what the plan's shape lookup (and its tiers) go by."""
    chains = count or DOT_CHAINS
    m = 512 if precision <= 2048 else max(16, 512 * 2048 // precision)
    rng = np.random.default_rng(seed + precision)
    lengths = rng.integers(8, m + 1, chains).astype(np.uint32)
    total = int(lengths.sum())
    xe, xl, ye, yl = generate(1, seed + precision, "mul", precision, total)
    X, Y, R = FpaArray(total, precision), FpaArray(total, precision), FpaArray(len(lengths), precision)
    plan = DotPlan(lengths)
    keep = get_variant("dot", precision)
    try:
        X.upload(xe, xl)
        Y.upload(ye, yl)
        shapes = variants("dot", precision)
        timings, first = [], None
        for v, (k, l) in enumerate(shapes):
            set_variant("dot", precision, v)
            ms = ctypes.c_float()
            rc = lib.fpa_dot_plan_time(plan.h, R.ref, X.ref, Y.ref, 3, 3, None, ctypes.byref(ms))
            if rc != 0:
                raise RuntimeError(f"fpa_dot_plan_time({precision}, variant {v}) returned {rc}")
            e, lm = R.download()
            if first is None:
                first = (e, lm)
            elif not (np.array_equal(e, first[0]) and np.array_equal(lm, first[1])):
                raise RuntimeError(f"dot {precision}: shape K={k} L={l} differs from K={shapes[0][0]} "
                                   f"L={shapes[0][1]}")
            timings.append((k, l, float(ms.value)))
        best = min(timings, key=lambda t: t[2])
        if log:
            parts = ", ".join(f"{k}x{l} {ms:.4g} ms{'*' if (k, l) == best[:2] else ''}" for k, l, ms in timings)
            log(f"dot {precision} ({total} elements, {len(lengths)} chains): {parts}")
        return best[:2], timings, chains
    finally:
        set_variant("dot", precision, keep)
        for x in (X, Y, R):
            x.close()
        plan.close()


def tune_one(op, precision, count=None, seed=1, log=None):
    """Times every shape of (op, precision).  Returns (best (K, L), [(K, L, ms)], count)."""
    if op == "dot":
        return tune_dot(precision, count, seed, log)
    count = count or tune_count(op, precision)
    unary = op in ("sqrt", "neg")
    ae, al, be, bl = generate(2, seed + precision, "sqrt" if op == "neg" else op, precision, count,
                              binary=not unary)
    A, R = FpaArray(count, precision), FpaArray(count, precision)
    B = None
    try:
        A.upload(ae, al)
        if not unary:
            B = FpaArray(count, precision)
            B.upload(be, bl)
        args = FpaArguments()
        args.count, args.precision, args.mode = count, precision, 3
        args.r_exps, args.r_limbs = R.s.gpu_exps, R.s.gpu_limbs
        args.a_exps, args.a_limbs = A.s.gpu_exps, A.s.gpu_limbs
        src = B if B is not None else A
        args.b_exps, args.b_limbs = src.s.gpu_exps, src.s.gpu_limbs
        shapes = variants(op, precision)
        timings, first = [], None
        for v, (k, l) in enumerate(shapes):
            ms = time_variant(op, precision, args, v)
            e, lm = R.download()
            if first is None:
                first = (e, lm)
            elif not (np.array_equal(e, first[0]) and np.array_equal(lm, first[1])):
                raise RuntimeError(f"{op} {precision}: shape K={k} L={l} differs from K={shapes[0][0]} "
                                   f"L={shapes[0][1]}")
            timings.append((k, l, ms))
        best = min(timings, key=lambda t: t[2])
        if log:
            parts = ", ".join(f"{k}x{l} {ms:.4g} ms{'*' if (k, l) == best[:2] else ''}" for k, l, ms in timings)
            log(f"{op} {precision} ({count} values): {parts}")
        return best[:2], timings, count
    finally:
        for x in (A, B, R):
            if x is not None:
                x.close()


SMALL_COUNTS = (2048, 4096, 8192, 16384)


def tier_bounds(counts):
    """Largest launch each timed size serves: the geometric middle to the next size."""
    c = sorted(counts)
    return [int((c[i] * c[i + 1]) ** 0.5) for i in range(len(c) - 1)] + [int(c[-1] * 2 ** 0.5)]


def tune(ops=TUNED_OPS, precisions=PRECISIONS, log=None, small_counts=SMALL_COUNTS):
    """Returns ({(op, precision): ((K, L), timings, count)}, {(op, precision): [((K, L),
    timings, count, bound) per small-launch tier written]}); also applies each choice.  Tiers are
    written up to the largest size whose winner differs from the large-launch shape (a
    launch takes the smallest tier that holds it, so the tiers below that one are needed
    even where their winner is the large shape)."""
    out, small = {}, {}
    for op in ops:
        for p in precisions:
            best, timings, count = tune_one(op, p, log=log)
            out[(op, p)] = (best, timings, count)
            set_variant(op, p, variants(op, p).index(best))
            if not small_counts:
                continue
            sizes = DOT_SMALL_CHAINS if op == "dot" else small_counts
            bounds = tier_bounds(sizes)
            tiers = [tune_one(op, p, count=c, log=log) + (b,) for c, b in zip(sorted(sizes), bounds)]
            differ = [i for i, t in enumerate(tiers) if t[0] != best]
            if differ:
                small[(op, p)] = tiers[:differ[-1] + 1]
                for sbest, _, _, bound in small[(op, p)]:
                    set_small_variant(op, p, variants(op, p).index(sbest), bound)
    return out, small


def write_table(path, table, small=None):
    small = small or {}
    lines = ["# kernel shapes chosen by cubigfloat_amd.autotune, " + time.strftime("%Y-%m-%d"),
             "# op precision K L [max_count]   (per-shape mean launch time; * = chosen)"]
    for (op, p), (best, timings, count) in sorted(table.items(), key=lambda kv: (KERNELS[kv[0][0]], kv[0][1])):
        parts = ", ".join(f"{k}x{l} {ms:.4g} ms" + ("*" if (k, l) == best else "") for k, l, ms in timings)
        lines.append(f"# {op} {p}, {count} {'chains' if op == 'dot' else 'values'}: {parts}")
        lines.append(f"{KERNELS[op]} {p} {best[0]} {best[1]}")
        for sbest, stimings, scount, bound in small.get((op, p), []):
            parts = ", ".join(f"{k}x{l} {ms:.4g} ms" + ("*" if (k, l) == sbest else "") for k, l, ms in stimings)
            unit = "chains" if op == "dot" else "values"
            lines.append(f"# {op} {p}, small launches of {scount} {unit} (tier up to {bound}): {parts}")
            lines.append(f"{KERNELS[op]} {p} {sbest[0]} {sbest[1]} {bound}")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--out", required=True, help="tuning table to write")
    ap.add_argument("--ops", default=",".join(TUNED_OPS))
    ap.add_argument("--precisions", default=",".join(str(p) for p in PRECISIONS))
    ap.add_argument("--small-counts", default=",".join(str(c) for c in SMALL_COUNTS),
                    help="launch sizes timed for the small-launch tiers ('' to skip)")
    args = ap.parse_args()
    ops = [o for o in args.ops.split(",") if o]
    precs = [int(p) for p in args.precisions.split(",") if p]
    for op in ops:
        if op not in KERNELS:
            sys.exit(f"unknown op {op}")
    log = lambda s: print(s, flush=True)
    table, small = tune(ops, precs, log=log, small_counts=[int(c) for c in args.small_counts.split(",") if c])
    write_table(args.out, table, small)
    log(f"wrote {len(table)} entries and {sum(len(v) for v in small.values())} small-launch tiers to {args.out}")


if __name__ == "__main__":
    main()
