"""Parity at the configurations' own sizes and generators, every element against MPFR.

* Config 1: the kind-1 generator (e ~ U[-64, 64], random signs, no zeros), 65,536 pairs at
  1024 bits, add / sub / mul in round to nearest even -- all elements.
* Config 2 at full size: 2^20 pairs at 4096 bits (the kind-2 generator: 1% zeros, 5%
  near-cancellation pairs).  count * precision = 2^32 here, the size where the reference's
  32-bit size products overflow (fpa_management.h:7, 18); every element of one mul and one
  sub launch is checked.
MPFR gold runs on all host threads (oracle/fpa_mpfr.c)."""
import sys

import numpy as np
import pytest

import cubigfloat_amd as F
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def full_vs_mpfr(op, p, count, kind, seed):
    ae, al, be, bl = F.generate(kind, seed, op, p, count, binary=op != "sqrt")
    say(f"[gpu] {op} p={p} count={count} kind={kind}")
    re_, rl = F.run_op(op, p, ae, al, be, bl)
    ge, gl, ok = O.mpfr_gold(op, p, ae, al, be, bl)
    assert ok.all(), f"MPFR could not hold {int((~ok).sum())} elements"
    eq = O.results_equal(re_, rl, ge, gl, p)
    assert eq.all(), f"{op} p={p}: {int((~eq).sum())} of {count} differ from MPFR, first {int(np.where(~eq)[0][0])}"
    return ae, be, re_


@pytest.mark.parametrize("op", ["add", "sub", "mul"])
def test_config1_generator_all_elements(op):
    ae, be, re_ = full_vs_mpfr(op, 1024, 65536, 1, 1)
    # kind 1 has no zeros and exponents within [-64, 64]
    assert (ae > 8).all() and (be > 8).all()
    e = (ae >> 1).astype(np.int64) - O.BIAS
    assert e.min() >= -64 and e.max() <= 64


@pytest.mark.parametrize("op", ["mul", "sub"])
def test_config2_full_size_count_times_precision_2_pow_32(op):
    count, p = 1 << 20, 4096
    assert count * p == 1 << 32
    ae, be, re_ = full_vs_mpfr(op, p, count, 2, 2)
    # the workload's zeros and (for sub) its exact and near cancellations are present
    assert ((ae == 2) | (be == 2)).sum() > count // 100
    if op == "sub":
        assert (re_ == 2).sum() > 0 or ((re_ >> 1).astype(np.int64) < (ae >> 1).astype(np.int64) - 64).sum() > 1000
