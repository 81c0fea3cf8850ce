"""A longer parity sweep, run on demand on a GPU (not collected by pytest):

    python tests/sweep_parity.py

Every op x precision x compiled shape (the mul shapes with K >= 28, which spill and run
40x slower, are left to the parity tests) on the reference generator's limb alphabet and on
config-2 random operands, in RNE and RZ, against the oracle: about 4.0 M values.  Prints
each mismatching case and the total; exits 1 if there is any."""
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, HERE)

import cubigfloat_amd as F  # noqa: E402
from oracle import oracle as O  # noqa: E402
import test_gpu_parity as T  # noqa: E402


def main():
    t0 = time.time()
    bad_total = 0
    for op in ("mul", "div", "sqrt", "add", "sub"):
        for p in F.PRECISIONS:
            cnt = max(256, min(20000, (1 << 24) // (p * p // 1024)))
            ae, al, be, bl = T.adversarial(op, p, cnt, 1000 + p)
            ae2, al2, be2, bl2 = F.generate(2, 2000 + p, op, p, cnt, binary=op != "sqrt")
            default = F.get_variant(op, p)
            for v, (k, l) in enumerate(F.variants(op, p)):
                if op == "mul" and k >= 28:
                    continue  # the spilling shapes are 40x slower; covered by the parity tests
                F.set_variant(op, p, v)
                for (xe, xl, ye, yl, kind) in ((ae, al, be, bl, "alph"), (ae2, al2, be2, bl2, "rand")):
                    for mode in ("rne", "rz"):
                        yb = None if op == "sqrt" else ye
                        ylb = None if op == "sqrt" else yl
                        re, rl = F.run_op(op, p, xe, xl, yb, ylb, mode=F.MODES[mode])
                        oe, ol = O.op(op, p, xe, xl, yb, ylb, mode=mode)
                        eq = O.results_equal(re, rl, oe, ol, p)
                        nb = int((~eq).sum())
                        bad_total += nb
                        if nb:
                            print(f"MISMATCH {op} p={p} K={k} L={l} {kind} {mode}: {nb} of {cnt}", flush=True)
            F.set_variant(op, p, default)
            print(f"{op} p={p} cnt={cnt} done t={time.time()-t0:.0f}s", flush=True)
    print("total mismatches", bad_total)
    return 1 if bad_total else 0


if __name__ == "__main__":
    sys.exit(main())
