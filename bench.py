#!/usr/bin/env python3
"""Benchmark: 4096-bit correctly rounded multiplication throughput on one MI355X.

Workload (default): 2^20 operand pairs at 4096-bit precision (128 limbs), exponents
uniform in [-1024, 1024], 1% exact zeros, 5% near-cancellation pairs; one step = one
fpa_mul launch over all 2^20 pairs, operands resident in HBM.  The same run times the
other plain per-op launches of that workload (add/sub/div/sqrt, 1024-bit add/sub/mul),
checks a sample of every op bit-exact against MPFR, and times MPFR on the host cores
as the CPU baseline.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--count N] [--no-cpu]
Prints one JSON line.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# MI355X: 256 CUs, 4 SIMD16 per CU, 2.4 GHz peak engine clock, 8 TB/s HBM3E.
CUS, CLOCK_HZ, HBM_GBS = 256, 2.4e9, 8000.0
# 32x32->64 multiplies (v_mad_u64_u32) per second at quarter rate: 256 CU * 64 lanes/4 clk * 2.4 GHz
MAD_PEAK_T = CUS * 16 * CLOCK_HZ / 1e12


class Hip:
    """Minimal HIP runtime bindings for events on the null stream (the stream fpa.h uses)."""

    def __init__(self):
        self.h = ctypes.CDLL("libamdhip64.so")
        self.h.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        self.h.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        self.h.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        self.h.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p, ctypes.c_void_p]
        self.h.hipEventDestroy.argtypes = [ctypes.c_void_p]

    def event(self):
        e = ctypes.c_void_p()
        assert self.h.hipEventCreate(ctypes.byref(e)) == 0
        return e

    def record(self, e):
        assert self.h.hipEventRecord(e, None) == 0

    def elapsed_ms(self, a, b):
        assert self.h.hipEventSynchronize(b) == 0
        ms = ctypes.c_float()
        assert self.h.hipEventElapsedTime(ctypes.byref(ms), a, b) == 0
        return ms.value

    def sync(self):
        assert self.h.hipDeviceSynchronize() == 0


def algorithmic(op, p):
    """Per-op algorithmic figure: (products, bytes).  products = 32x32->64 multiplies of
    the basecase algorithm (mul n^2, div n^2, sqrt n^2/2); bytes = exponent words + limbs
    read and written."""
    n = p // 32
    prods = {"mul": n * n, "div": n * n, "sqrt": n * n // 2}.get(op, 0)
    nops = 1 if op == "sqrt" else 2
    byts = (nops + 1) * (4 * n + 4)
    return prods, byts


def time_op(F, hip, op, p, count, steps, warmup, kind, seed, check_sample):
    ae, al, be, bl = F.generate(kind, seed, op, p, count, binary=op != "sqrt")
    A, R = F.FpaArray(count, p), F.FpaArray(count, p)
    B = F.FpaArray(count, p) if op != "sqrt" else None
    A.upload(ae, al)
    if B is not None:
        B.upload(be, bl)
    for _ in range(warmup):
        assert F.apply(op, R, A, B) == F.SUCCESS
    hip.sync()
    evs = [hip.event() for _ in range(steps + 1)]
    t0 = time.perf_counter()
    hip.record(evs[0])
    for i in range(steps):
        assert F.apply(op, R, A, B) == F.SUCCESS
        hip.record(evs[i + 1])
    hip.sync()
    wall = time.perf_counter() - t0
    per = [hip.elapsed_ms(evs[i], evs[i + 1]) for i in range(steps)]
    re, rl = R.download()
    for x in (A, B, R):
        if x is not None:
            x.close()
    n = p // 32
    sample = slice(0, check_sample)
    res = {"op": op, "precision": p, "count": count, "ms_per_launch": float(np.mean(per)),
           "ms_min": float(np.min(per)), "wall_s": wall}
    res["ops_per_s"] = count / (res["ms_per_launch"] / 1e3)
    prods, byts = algorithmic(op, p)
    res["tproducts_per_s"] = prods * res["ops_per_s"] / 1e12
    res["gb_per_s"] = byts * res["ops_per_s"] / 1e9
    sample_in = (ae[sample], al[sample.start * n:sample.stop * n],
                 None if be is None else be[sample], None if bl is None else bl[sample.start * n:sample.stop * n])
    return res, sample_in, (re[sample], rl[sample.start * n:sample.stop * n])


def cpu_baseline(op, p, sample_in, gpu_out, threads):
    """MPFR-backed fpa.h CPU backend (oracle/fpa_mpfr.c) on the host cores; also checks the
    GPU's results for the same elements bit-exact against it."""
    from oracle import oracle as O
    ae, al, be, bl = sample_in
    cnt = len(ae)
    O.cpu_set_threads(threads)
    A, R = O.CpuArray(cnt, p), O.CpuArray(cnt, p)
    A.load(ae, al)
    B = None
    if be is not None:
        B = O.CpuArray(cnt, p)
        B.load(be, bl)
    O.cpu_op(op, R, A, B)  # warm
    reps, t = 0, 0.0
    t0 = time.perf_counter()
    while t < 1.0 or reps < 2:
        O.cpu_op(op, R, A, B)
        reps += 1
        t = time.perf_counter() - t0
    ce, cl = R.store(op)
    ok = O.results_equal(gpu_out[0], gpu_out[1], ce, cl, p)
    for x in (A, B, R):
        if x is not None:
            x.close()
    return cnt * reps / t, reps * cnt, bool(ok.all()), int((~ok).sum())


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--count", type=int, default=1 << 20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and MPFR checks")
    ap.add_argument("--only", default="", help="comma list of extra ops to time (default all)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (got --gpus {args.gpus}); "
              "the path shards by element ranges with no exchange and is not run on several GPUs.",
              file=sys.stderr)
        sys.exit(1)

    import cubigfloat_amd as F
    hip = Hip()
    F.lib.fpa_ubench_mad.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(ctypes.c_float)]
    blocks, iters = 256 * 4 * 8, 4096
    ms = ctypes.c_float()
    assert F.lib.fpa_ubench_mad(blocks, iters, ctypes.byref(ms)) == 0
    mad_measured_t = blocks * 256 * iters * 8 / (ms.value / 1e3) / 1e12
    print(f"[bench] v_mad_u64_u32 probe: {mad_measured_t:.2f} T/s ({ms.value:.2f} ms)", file=sys.stderr, flush=True)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    p, count = 4096, args.count
    print(f"[bench] mul {p}-bit x{count}: warmup {args.warmup}, steps {args.steps}", file=sys.stderr, flush=True)
    head, sample_in, gpu_out = time_op(F, hip, "mul", p, count, args.steps, args.warmup, 2, 2, 4096)

    extras = []
    others = [("add", 4096, 2), ("sub", 4096, 2), ("div", 4096, 2), ("sqrt", 4096, 2),
              ("add", 1024, 1), ("sub", 1024, 1), ("mul", 1024, 1)]
    wanted = set(args.only.split(",")) if args.only else None
    checks = {}
    for op, pp, kind in others:
        if wanted is not None and f"{op}{pp}" not in wanted:
            continue
        cnt = count if kind == 2 else 65536
        print(f"[bench] {op} {pp}-bit x{cnt}", file=sys.stderr, flush=True)
        r, si, go = time_op(F, hip, op, pp, cnt, max(3, args.steps // 4), 1, kind, kind, 2048)
        if not args.no_cpu:
            rate, _, ok, bad = cpu_baseline(op, pp, si, go, threads)
            r["cpu_mpfr_ops_per_s"] = rate
            r["mpfr_sample_bit_exact"] = ok
            checks[f"{op}{pp}"] = ok
        extras.append(r)

    cpu = None
    if not args.no_cpu:
        print("[bench] cpu baseline (MPFR, host cores)", file=sys.stderr, flush=True)
        rate_all, done, ok, bad = cpu_baseline("mul", p, sample_in, gpu_out, threads)
        rate_1, _, _, _ = cpu_baseline("mul", p, (sample_in[0][:512], sample_in[1][:512 * 128],
                                                  sample_in[2][:512], sample_in[3][:512 * 128]),
                                       (gpu_out[0][:512], gpu_out[1][:512 * 128]), 1)
        checks["mul4096"] = ok
        cpu = {"value": rate_all, "unit": "mul/s", "cores": threads, "kind": "port",
               "sample": f"MPFR 4.0.2 mpfr_mul RNDN via the MPFR-backed fpa.h CPU backend on the first 4096 "
                         f"elements of the same workload, repeated >=1 s; {done} muls",
               "single_core_mul_per_s": rate_1, "gpu_sample_bit_exact_vs_mpfr": ok}

    ops = head["ops_per_s"]
    res = {
        "metric": "4096-bit correctly-rounded mul/s", "value": ops, "unit": "mul/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": head["ms_per_launch"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": "plain per-op launch, 2^20 x 4096-bit mul (RNE), e~U[-1024,1024], "
                               "1% zeros, 5% near-cancellation pairs", "count": count, "precision": p,
                   "kernel_shape": F.variants("mul", p)[F.get_variant("mul", p)]},
        "roofline": {"bound": "valu", "achieved": head["tproducts_per_s"], "peak": MAD_PEAK_T,
                     "unit": "T 32x32->64 products/s", "frac": head["tproducts_per_s"] / MAD_PEAK_T,
                     "traffic": None, "peak_measured_mad_probe": mad_measured_t,
                     "frac_of_measured": head["tproducts_per_s"] / mad_measured_t,
                     "note": "INT32 VALU multiply roofline (no MFMA, no HBM bound at 4096 bits); "
                             "peak = v_mad_u64_u32 at quarter rate, 256 CU x 16 lanes/clk x 2.4 GHz"},
        "cpu_baseline": cpu,
        "checks_bit_exact_vs_mpfr": checks,
        "extras": extras,
        "wall_s": head["wall_s"],
    }
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
