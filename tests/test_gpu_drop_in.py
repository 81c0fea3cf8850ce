"""The reference's own test harness (tests.c) driving the library from C through fpa.h.

tests/c/drop_in.c mirrors the reference's fpa_test.c test_<op> flow: generate with the
reference's generator + MPFR gold, fpa_init, copy (or fpa_set_* from host values), op,
copy back, tests_compare_results (which exits non-zero on the first mismatch).

The driver is built twice (tests/c/Makefile): against this library's include/fpa.h, and
against the reference's own, unmodified fpa.h with this repository's include/ absent from
the include path -- a caller compiled for the reference links and runs unchanged."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
BUILD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c", "_build")
EXES = {"own_header": os.path.join(BUILD, "drop_in"), "reference_header": os.path.join(BUILD, "drop_in_refhdr")}

CASES = [
    ("add", 20000, 1024), ("sub", 20000, 2048), ("mul", 20000, 2048), ("mul", 5000, 5120),
    ("div", 20000, 1024), ("div", 5000, 2048), ("sqrt", 20000, 1024), ("sqrt", 5000, 5120),
    ("add", 4000, 8192), ("sub", 4000, 3072), ("neg", 20000, 1024), ("neg", 4000, 8192),
    ("sgn", 20000, 1024), ("cmp", 20000, 2048), ("cmp", 4000, 8192),
    ("set 1024", 10000, 3072), ("set 4096", 10000, 1024), ("set 2048", 10000, 2048),
    ("set_ui", 20000, 1024), ("set_si", 20000, 2048), ("set_float", 20000, 1024), ("set_double", 20000, 4096),
    ("set_ui", 3000, 8192), ("set_si", 3000, 5120), ("set_float", 3000, 8192), ("set_double", 3000, 8192),
]


@pytest.mark.parametrize("header", sorted(EXES))
@pytest.mark.parametrize("op,count,precision", CASES)
def test_reference_harness_through_c_api(header, op, count, precision):
    exe = EXES[header]
    if not os.path.exists(exe):
        pytest.skip("drop-in driver not built (needs the reference's tests.h at build time)")
    print(f"[gpu] drop_in ({header}) {op} {count} {precision}", flush=True)
    args = [exe, op, str(count), str(precision)]
    if op.startswith("set "):
        args = [exe, "set", str(count), str(precision), op.split()[1]]
    r = subprocess.run(args, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "tests passed" in r.stdout
