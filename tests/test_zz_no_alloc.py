"""fpa_set_scalars_noalloc through the C ABI: the allocation-free form of fpa_set_ui / si /
float / double gives results identical to those calls, on consecutive calls that reuse its
fixed staging buffer (the second call smaller, so stale staging words lie past its end),
and rejects counts above its limit."""
import numpy as np
import pytest

import cubigfloat_amd as F
from oracle import oracle as O

import test_set_scalar as TS

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def via(fn, name, p, vals, fill_seed):
    R = F.FpaArray(len(vals), p)
    try:
        rng = np.random.default_rng(fill_seed)
        R.upload(rng.integers(9, 2 ** 32, len(vals), dtype=np.uint64).astype(np.uint32),
                 rng.integers(0, 2 ** 32, len(vals) * (p // 32), dtype=np.uint64).astype(np.uint32))
        assert fn(R, name, vals) == F.SUCCESS
        return R.download()
    finally:
        R.close()


@pytest.mark.parametrize("name", TS.KINDS)
def test_noalloc_equals_set_scalars_twice(name):
    for p, count, seed in ((4096, 300000, 1), (1024, 5000, 2)):
        vals = TS.edge_values(name, np.random.default_rng(seed), extra=count - 20)[:count]
        say(f"[gpu] noalloc {name} p={p} count={len(vals)}")
        e1, l1 = via(F.set_scalars, name, p, vals, 10 + seed)
        e2, l2 = via(F.set_scalars_noalloc, name, p, vals, 20 + seed)
        assert np.array_equal(e1, e2)
        assert np.array_equal(l1, l2)
        oe, ol = O.set_scalar(name, p, vals)
        assert O.results_equal(e2, l2, oe, ol, p).all()


def test_noalloc_at_limit_and_above():
    n = F.NOALLOC_MAX_VALUES
    vals = TS.edge_values("set_double", np.random.default_rng(4), extra=n)[:n]
    say(f"[gpu] noalloc set_double p=256 count={n}")
    e1, l1 = via(F.set_scalars, "set_double", 256, vals, 5)
    e2, l2 = via(F.set_scalars_noalloc, "set_double", 256, vals, 6)
    assert np.array_equal(e1, e2) and np.array_equal(l1, l2)
    R = F.FpaArray(n + 1, 256)
    try:
        assert F.set_scalars_noalloc(R, "set_ui", np.zeros(n + 1, np.uint64)) == F.INVALID_COUNTS
    finally:
        R.close()
    R = F.FpaArray(0, 256)
    try:
        assert F.set_scalars_noalloc(R, "set_ui", np.zeros(0, np.uint64)) == F.SUCCESS
    finally:
        R.close()
