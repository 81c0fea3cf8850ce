"""Writes the committed golden fixtures (run here, where the reference sources exist).

* ref_<op>_<p>.npz  -- cases drawn by the reference's own generator
  (tests.c:727-925, compiled from its sources into oracle/_ref/) with the reference
  harness's MPFR gold (round to nearest even).
* mpfr_<op>_<p>.npz -- config-2-style operands (zeros, near-cancellation pairs) from the
  package generator at precisions outside the reference's range, with MPFR gold.
* setcmp/ref_set_<p>_<src>.npz, setcmp/ref_sgn_<p>.npz, setcmp/ref_cmp_<p>.npz -- the
  reference generator's set / sgn / cmp cases (tests.c:580-699) with its MPFR gold
  (mpfr_set in round to nearest even, mpfr_sgn, mpfr_cmp).
* setcmp/ref_set_{ui,si,float,double}_<p>.npz, setcmp/ref_neg_<p>.npz -- the reference
  generator's conversion and negation cases (tests.c:487-578, 700-725: source values, or
  operands, with mpfr_set_ui/si/flt/d or mpfr_neg gold).
Usage: python tests/golden/make_golden.py            (element-wise ops)
       python tests/golden/make_golden.py --setcmp   (set / sgn / cmp only)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402

OPS = ["add", "sub", "mul", "div", "sqrt"]


def main():
    O.build()
    for op in OPS:
        for p, cnt in ((1024, 64), (2048, 48)):
            g = O.ref_generate(op, cnt, p)
            np.savez_compressed(os.path.join(HERE, f"ref_{op}_{p}.npz"), precision=np.uint32(p), **g)
    import cubigfloat_amd as F
    for op in OPS:
        for p, cnt in ((256, 64), (4096, 24), (16384, 6)):
            ae, al, be, bl = F.generate(2, 1000 + p, op, p, cnt, binary=op != "sqrt")
            ge, gl, ok = O.mpfr_gold(op, p, ae, al, be, bl)
            assert ok.all()
            d = dict(a_exps=ae, a_limbs=al, gold_exps=ge, gold_limbs=gl, precision=np.uint32(p))
            if be is not None:
                d.update(b_exps=be, b_limbs=bl)
            np.savez_compressed(os.path.join(HERE, f"mpfr_{op}_{p}.npz"), **d)
    print("fixtures written to", HERE)


def main_setcmp():
    O.build()
    out = os.path.join(HERE, "setcmp")
    os.makedirs(out, exist_ok=True)
    for p, src in ((1024, 3072), (3072, 1024), (2048, 2048), (1024, 8192), (8192, 1024)):
        g = O.ref_generate("set", 48, p, src)
        np.savez_compressed(os.path.join(out, f"ref_set_{p}_{src}.npz"), precision=np.uint32(p),
                            source_precision=np.uint32(src), **g)
    for name in ("sgn", "cmp"):
        for p in (1024, 4096):
            g = O.ref_generate(name, 64, p)
            np.savez_compressed(os.path.join(out, f"ref_{name}_{p}.npz"), precision=np.uint32(p), **g)
    for name in ("set_ui", "set_si", "set_float", "set_double", "neg"):
        for p in (1024, 4096):
            g = O.ref_generate(name, 64, p)
            np.savez_compressed(os.path.join(out, f"ref_{name}_{p}.npz"), precision=np.uint32(p), **g)
    print("fixtures written to", out)


if __name__ == "__main__":
    if "--setcmp" in sys.argv:
        main_setcmp()
    else:
        main()
