"""Autotuner on the device: every shape of an (op, precision) is timed and returns the
same results; the chosen shape is one of the compiled ones."""
import pytest

import cubigfloat_amd as F
from cubigfloat_amd import autotune as T

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("op,p", [("mul", 1024), ("div", 2048), ("sqrt", 4096), ("add", 8192), ("dot", 2048)])
def test_tune_one_small(op, p):
    print(f"[gpu] autotune {op} {p}", flush=True)
    keep = F.get_variant(op, p)
    best, timings, count = T.tune_one(op, p, count=4096)
    assert count == 4096
    assert best in F.variants(op, p)
    assert [t[:2] for t in timings] == F.variants(op, p)
    assert all(t[2] > 0 for t in timings)
    F.set_variant(op, p, keep)
