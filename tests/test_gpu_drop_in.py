"""The reference's own test harness (tests.c) driving the library from C through fpa.h.

tests/c/drop_in.c mirrors the reference's fpa_test.c test_<op> flow: generate with the
reference's generator + MPFR gold, fpa_init, copy, op, copy back, tests_compare_results
(which exits non-zero on the first mismatch)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c", "_build", "drop_in")


@pytest.mark.skipif(not os.path.exists(EXE), reason="drop-in driver not built (needs the reference's tests.h at build time)")
@pytest.mark.parametrize("op,count,precision", [
    ("add", 20000, 1024), ("sub", 20000, 2048), ("mul", 20000, 2048), ("mul", 5000, 5120),
    ("div", 20000, 1024), ("div", 5000, 2048), ("sqrt", 20000, 1024), ("sqrt", 5000, 5120),
    ("add", 4000, 8192), ("sub", 4000, 3072), ("sgn", 20000, 1024), ("cmp", 20000, 2048), ("cmp", 4000, 8192),
    ("set 1024", 10000, 3072), ("set 4096", 10000, 1024), ("set 2048", 10000, 2048)])
def test_reference_harness_through_c_api(op, count, precision):
    print(f"[gpu] drop_in {op} {count} {precision}", flush=True)
    args = [EXE, op, str(count), str(precision)]
    if op.startswith("set"):
        args = [EXE, "set", str(count), str(precision), op.split()[1]]
    r = subprocess.run(args, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "tests passed" in r.stdout
