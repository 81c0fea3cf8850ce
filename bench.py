#!/usr/bin/env python3
"""Benchmark: 4096-bit correctly rounded multiplication throughput on one MI355X.

Workload (default): 2^20 operand pairs at 4096-bit precision (128 limbs), exponents
uniform in [-1024, 1024], 1% exact zeros, 5% near-cancellation pairs; one step = one
fpa_mul launch over all 2^20 pairs, operands resident in HBM.  The same run times the
other plain per-op launches of that workload (add/sub/div/sqrt, 1024-bit add/sub/mul),
checks every result of every timed launch bit-exact against MPFR gold computed on all
host threads (full_launch_checks_vs_mpfr, mismatches_vs_mpfr), and times MPFR on the host
cores as the CPU baseline.

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
# INT32 VALU issue ceiling: one 32x32->64 multiply-add (v_mad_u64_u32) per lane per clock,
# 256 CU * 4 SIMD * 16 lanes * 2.4 GHz.  (A quarter-rate multiplier would give a quarter of
# this; the probe below measures the device's sustained rate.)
MAD_PEAK_T = CUS * 64 * CLOCK_HZ / 1e12
MAD_QUARTER_RATE_T = MAD_PEAK_T / 4
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_4096.md")


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
    nops = 1 if op in ("sqrt", "neg") else 2
    byts = (nops + 1) * (4 * n + 4)
    return prods, byts


def time_op(F, hip, op, p, count, steps, warmup, kind, seed, check_sample):
    ae, al, be, bl = F.generate(kind, seed, op, p, count, binary=op != "sqrt")
    A, R = F.FpaArray(count, p), F.FpaArray(count, p)
    B = F.FpaArray(count, p) if op != "sqrt" else None
    t_up = time.perf_counter()
    A.upload(ae, al)
    if B is not None:
        B.upload(be, bl)
    t_up = time.perf_counter() - t_up
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
    full = ((ae, al, be, bl), (re, rl))
    res = {"op": op, "precision": p, "count": count, "ms_per_launch": float(np.mean(per)),
           "ms_min": float(np.min(per)), "wall_s": wall, "upload_s": t_up}
    res["ops_per_s"] = count / (res["ms_per_launch"] / 1e3)
    prods, byts = algorithmic(op, p)
    res["tproducts_per_s"] = prods * res["ops_per_s"] / 1e12
    res["gb_per_s"] = byts * res["ops_per_s"] / 1e9
    # the path's own roofline: INT32 multiply issue for mul/div/sqrt, HBM for add/sub
    if prods:
        res["frac_of_int32_mad_peak"] = res["tproducts_per_s"] / MAD_PEAK_T
    else:
        res["frac_of_hbm_peak"] = res["gb_per_s"] / HBM_GBS
    sample_in = (ae[sample], al[sample.start * n:sample.stop * n],
                 None if be is None else be[sample], None if bl is None else bl[sample.start * n:sample.stop * n])
    return res, sample_in, (re[sample], rl[sample.start * n:sample.stop * n]), full


def full_check(op, p, full, threads):
    """Every result of the timed launch against MPFR gold (all host threads): returns
    {"checked": elements MPFR could hold, "mismatches": n, "seconds": t}."""
    from oracle import oracle as O
    (ae, al, be, bl), (re, rl) = full
    t0 = time.perf_counter()
    ge, gl, ok = O.mpfr_gold(op, p, ae, al, be, bl, threads=threads)
    eq = O.results_equal(re, rl, ge, gl, p)
    return {"checked": int(ok.sum()), "of": int(len(ok)), "mismatches": int((~eq & ok).sum()),
            "seconds": time.perf_counter() - t0}


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


def measured_traffic():
    """HBM bytes per launch from the committed rocprofv3 counter summary (FETCH_SIZE and
    WRITE_SIZE passes, scaled by the same counters on a copy kernel of known size):
    lines 'traffic_bytes_per_launch <op><precision> <bytes>'.  {} if absent."""
    out = {}
    if os.path.exists(TRAFFIC_FILE):
        for line in open(TRAFFIC_FILE):
            f = line.split()
            if len(f) == 3 and f[0] == "traffic_bytes_per_launch":
                out[f[1]] = float(f[2])
    return out


def kernels_only(F, hip, args):
    """Short run for hardware-counter passes: neg (a pure copy of known bytes), mul and add
    (or the ops named by --only) at 4096 bits over the same 2^20-value workload, one
    warmup and `steps` launches each."""
    res = []
    ops = [o for o in args.only.split(",") if o] or ["neg", "mul", "add"]
    for op in ops:
        print(f"[bench] kernels-only {op} 4096 x{args.count}", file=sys.stderr, flush=True)
        r, _, _, _ = time_op(F, hip, op, 4096, args.count, args.steps, 1, 2, 2, 16)
        r["algorithmic_bytes_per_launch"] = algorithmic(op, 4096)[1] * args.count
        res.append(r)
    print(json.dumps({"kernels_only": res}), flush=True)


def pct(x, q):
    return float(np.percentile(np.asarray(x), q)) if len(x) else None


def bench_config3(F, hip, nops, steps, cpu, threads):
    """2^18 mixed ops (precision log-uniform over 256..8192, 40% mul / 15% add / 15% sub /
    20% div / 10% sqrt, config-2 operands) in one graph launch (fork/join of per-segment
    kernels) vs the same kernels as plain launches on one stream, MPFR on host cores."""
    from cubigfloat_amd import workloads as W
    ops, precs, modes = W.mixed_ops(nops, 3)
    ae, al, be, bl = W.mixed_operands(ops, precs, 3)
    B = F.Batch(ops, precs, modes)
    B.set_operands(ae, al, be, bl)
    res = {"workload": f"{nops} mixed ops, precision log-uniform over 256..8192, op mix 40/15/15/20/10",
           "plan": B.info()}
    for name, fn in (("batch_graph", B.run), ("plain_per_op", B.run_plain)):
        fn()
        hip.sync()
        e0, e1 = hip.event(), hip.event()
        hip.record(e0)
        for _ in range(steps):
            fn()
        hip.record(e1)
        ms = hip.elapsed_ms(e0, e1) / steps
        res[name] = {"ms": ms, "ops_per_s": nops / (ms / 1e3)}
        if name == "batch_graph":
            be_, bl_ = B.results()
        else:
            pe_, pl_ = B.results()
    off = W.limb_offsets(precs)
    same = bool((be_ == pe_).all() and (bl_ == pl_).all())
    res["batch_equals_plain"] = same
    # the same batch as one persistent kernel launch (tiles from a work queue), captured too
    P = F.Batch(ops, precs, modes, persistent=True)
    P.set_operands(ae, al, be, bl)
    P.run()
    hip.sync()
    e0, e1 = hip.event(), hip.event()
    hip.record(e0)
    for _ in range(steps):
        P.run()
    hip.record(e1)
    ms = hip.elapsed_ms(e0, e1) / steps
    xe_, xl_ = P.results()
    res["persistent_kernel"] = {"ms": ms, "ops_per_s": nops / (ms / 1e3), "plan": P.info(),
                                "equals_graph": bool((be_ == xe_).all() and (bl_ == xl_).all())}
    P.close()
    if cpu:
        from oracle import oracle as O
        t_cpu, sample_ok, bad = 0.0, True, 0
        O.cpu_set_threads(threads)
        inv = {v: k for k, v in F.OPS.items()}
        for op, p, m, idx, e1, l1, pos in W.split_by_op(ops, precs, modes, ae, al):
            e2, l2 = be[idx], bl[pos]
            cnt = len(idx)
            A, R = O.CpuArray(cnt, p), O.CpuArray(cnt, p)
            A.load(e1, l1)
            Bc = None
            if inv[op] != "sqrt":
                Bc = O.CpuArray(cnt, p)
                Bc.load(e2, l2)
            t0 = time.perf_counter()
            O.cpu_op(inv[op], R, A, Bc, ["rz", "ru", "rd", "rne", "rna"][m])
            t_cpu += time.perf_counter() - t0
            ce, cl = R.store(inv[op])
            eq = O.results_equal(be_[idx], bl_[pos], ce, cl, p)
            sample_ok &= bool(eq.all())
            bad += int((~eq).sum())
            for x in (A, Bc, R):
                if x is not None:
                    x.close()
        res["cpu_mpfr"] = {"ops_per_s": nops / t_cpu, "cores": threads}
        res["bit_exact_vs_mpfr"] = sample_ok
        res["mismatches_vs_mpfr"] = bad
    B.close()
    return res


def bench_config4(F, hip, nchains, steps, cpu, threads):
    """16,384 dot products at 2048 bits, lengths U[8,512], per-element RN mul then RN add
    accumulated in registers, one launch."""
    from cubigfloat_amd import workloads as W
    lengths, xe, xl, ye, yl = W.dot_workload(nchains, 2048, 4)
    total = int(lengths.sum())
    X, Y, R = F.FpaArray(total, 2048), F.FpaArray(total, 2048), F.FpaArray(nchains, 2048)
    X.upload(xe, xl)
    Y.upload(ye, yl)
    plan = F.DotPlan(lengths)
    plan.run(R, X, Y)
    hip.sync()
    e0, e1 = hip.event(), hip.event()
    hip.record(e0)
    for _ in range(steps):
        plan.run(R, X, Y)
    hip.record(e1)
    ms = hip.elapsed_ms(e0, e1) / steps
    re, rl = R.download()
    res = {"workload": f"{nchains} dot products at 2048 bits, lengths U[8,512], {total} elements",
           "ms": ms, "dots_per_s": nchains / (ms / 1e3), "mul_add_pairs_per_s": total / (ms / 1e3)}
    if cpu:
        from oracle import oracle as O
        k = nchains
        off = np.concatenate([[0], np.cumsum(lengths[:k], dtype=np.int64)])
        n = 64
        t0 = time.perf_counter()
        me, ml = O.mpfr_dot(2048, lengths[:k], xe[:off[-1]], xl[:off[-1] * n], ye[:off[-1]], yl[:off[-1] * n],
                            threads=threads)
        dt = time.perf_counter() - t0
        res["cpu_mpfr"] = {"dots_per_s": k / dt, "cores": threads, "sample": f"all {k} chains"}
        eq = O.results_equal(re[:k], rl[:k * n], me, ml, 2048)
        res["bit_exact_vs_mpfr"] = bool(eq.all())
        res["mismatches_vs_mpfr"] = int((~eq).sum())
    for a in (X, Y, R):
        a.close()
    plan.close()
    return res


def bench_config5(F, nreq, budget):
    """Served workload: Poisson arrivals, requests of 1..4096 ops (log-uniform) at one
    precision each (256..16384), op mix of the mixed batch; continuous-batching server
    (resident pool = budget, 8 streams) vs plain per-request launches."""
    from cubigfloat_amd import workloads as W
    reqs = W.served_requests(nreq, 5, rate=0.0)
    nops = sum(len(q["ops"]) for q in reqs)

    def run(kind, rate):
        rng = np.random.default_rng(17)
        t = np.cumsum(rng.exponential(1.0 / rate, len(reqs))) if rate > 0 else np.zeros(len(reqs))
        tr = F.Trace(reqs)
        for i, q in enumerate(tr.arr):
            q.arrival = float(t[i])
        if kind == "server":
            srv = F.Server(budget, slots=8, max_request_bytes=64 << 20)
            rc = srv.run_trace(tr)
            extra = {"peak_pool_bytes": srv.peak_bytes(), "budget_waits": srv.budget_waits()}
            srv.close()
        else:
            rc = F.serve_plain(tr)
            extra = {}
        a, s, d, st = tr.times()
        assert rc == F.SUCCESS and (st == 0).all()
        lat = (d - a) * 1e3
        span = d.max() - a.min()
        out = {"rate_req_per_s": rate, "p50_ms": pct(lat, 50), "p99_ms": pct(lat, 99),
               "mean_ms": float(lat.mean()), "ops_per_s": nops / span, "req_per_s": len(reqs) / span}
        out.update(extra)
        return out, tr

    cap_plain, tp = run("plain", 0.0)
    cap_server, ts = run("server", 0.0)
    same = all((tp.results(i)[0] == ts.results(i)[0]).all() and (tp.results(i)[1] == ts.results(i)[1]).all()
               for i in range(len(reqs)))
    res = {"workload": f"{nreq} requests, {nops} ops, sizes log-uniform 1..4096, precisions 256..16384",
           "budget_bytes": budget, "slots": 8, "capacity_plain": cap_plain, "capacity_server": cap_server,
           "server_results_equal_plain": same, "loads": []}
    for frac in (0.5, 0.8):
        lam = frac * cap_plain["req_per_s"]
        ps, _ = run("plain", lam)
        ss, _ = run("server", lam)
        res["loads"].append({"fraction_of_plain_capacity": frac, "lambda_req_per_s": lam, "plain": ps, "server": ss})
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--count", type=int, default=1 << 20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and MPFR checks")
    ap.add_argument("--only", default="", help="comma list of extra ops to time (default all)")
    ap.add_argument("--configs", default="3,4,5", help="batched/chained/served configurations to run ('' for none)")
    ap.add_argument("--kernels-only", action="store_true",
                    help="only time neg/mul/add at 4096 bits (for rocprofv3 counter passes)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (got --gpus {args.gpus}); "
              "the path shards by element ranges with no exchange and is not run on several GPUs.",
              file=sys.stderr)
        sys.exit(1)

    import cubigfloat_amd as F
    hip = Hip()
    if args.kernels_only:
        kernels_only(F, hip, args)
        return
    F.lib.fpa_ubench_mad.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(ctypes.c_float)]
    blocks, iters = 256 * 4 * 8, 4096
    ms = ctypes.c_float()
    assert F.lib.fpa_ubench_mad(blocks, iters, ctypes.byref(ms)) == 0
    mad_measured_t = blocks * 256 * iters * 8 / (ms.value / 1e3) / 1e12
    print(f"[bench] v_mad_u64_u32 probe: {mad_measured_t:.2f} T/s ({ms.value:.2f} ms)", file=sys.stderr, flush=True)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    p, count = 4096, args.count
    print(f"[bench] mul {p}-bit x{count}: warmup {args.warmup}, steps {args.steps}", file=sys.stderr, flush=True)
    head, sample_in, gpu_out, full = time_op(F, hip, "mul", p, count, args.steps, args.warmup, 2, 2, 4096)
    full_checks = {}
    if not args.no_cpu:
        full_checks["mul4096"] = full_check("mul", p, full, threads)
    del full

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
        r, si, go, full = time_op(F, hip, op, pp, cnt, max(3, args.steps // 4), 1, kind, kind, 2048)
        if not args.no_cpu:
            full_checks[f"{op}{pp}"] = full_check(op, pp, full, threads)
        del full
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

    configs = {}
    wanted_cfg = set(c for c in args.configs.split(",") if c)
    if "3" in wanted_cfg:
        print("[bench] config 3: mixed batch, one graph launch", file=sys.stderr, flush=True)
        configs["config3"] = bench_config3(F, hip, 1 << 18, 20, not args.no_cpu, threads)
    if "4" in wanted_cfg:
        print("[bench] config 4: dot chains", file=sys.stderr, flush=True)
        configs["config4"] = bench_config4(F, hip, 16384, 3, not args.no_cpu, threads)
    if "5" in wanted_cfg:
        print("[bench] config 5: served workload", file=sys.stderr, flush=True)
        configs["config5"] = bench_config5(F, 400, 4 << 30)

    ops = head["ops_per_s"]
    traffic = measured_traffic().get("mul4096")
    shape = F.variants("mul", p)[F.get_variant("mul", p)]
    roofline = {"bound": "valu", "achieved": head["tproducts_per_s"], "peak": MAD_PEAK_T,
                "unit": "T 32x32->64 products/s", "frac": head["tproducts_per_s"] / MAD_PEAK_T,
                "peak_measured_mad_probe": mad_measured_t,
                "frac_of_measured": head["tproducts_per_s"] / mad_measured_t,
                "peak_if_quarter_rate": MAD_QUARTER_RATE_T,
                "note": "INT32 VALU multiply roofline (the path is exact integer VALU work; at 4096 bits "
                        "mul is far from the HBM bound); achieved = n^2 = 16384 word products per mul x "
                        "mul/s; peak = one v_mad_u64_u32 per lane per clock, 256 CU x 64 lanes x 2.4 GHz; "
                        "traffic = HBM bytes per launch from rocprofv3 FETCH_SIZE+WRITE_SIZE (profiles/)"}
    roofline["traffic"] = traffic
    roofline["traffic_algorithmic"] = algorithmic("mul", p)[1] * count
    res = {
        "metric": "4096-bit correctly-rounded mul/s", "value": ops, "unit": "mul/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": head["ms_per_launch"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": "plain per-op launch, 2^20 x 4096-bit mul (RNE), e~U[-1024,1024], "
                               "1% zeros, 5% near-cancellation pairs", "count": count, "precision": p,
                   "kernel_shape": shape},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "checks_bit_exact_vs_mpfr": checks,
        "full_launch_checks_vs_mpfr": full_checks,
        "mismatches_vs_mpfr": sum(c["mismatches"] for c in full_checks.values()) if full_checks else None,
        "extras": extras,
        "configs": configs,
        "wall_s": head["wall_s"],
        "host_buffer_rate": {"upload_s": head["upload_s"],
                             "mul_per_s_including_operand_upload":
                                 count / (head["upload_s"] + head["ms_per_launch"] / 1e3),
                             "note": "fpa_copy_array_to_device of both operand arrays from pageable host "
                                     "memory, then one launch; not the metric"},
    }
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
