"""Continuous-batching server (fpa_serve.h): every request's results are bit-identical
when served alone and when served in an overlapping mix, admission respects the
device-memory budget, and results match the oracle."""
import numpy as np
import pytest

import cubigfloat_amd as F
from cubigfloat_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu
INV = {v: k for k, v in F.OPS.items()}
MODES = ["rz", "ru", "rd", "rne", "rna"]


def say(*a):
    print(*a, flush=True)


def test_same_alone_and_in_mix():
    reqs = W.served_requests(80, 5, rate=20000.0, max_ops=512,
                             precisions=(256, 1024, 4096, 16384))
    mixed = F.Trace(reqs)
    srv = F.Server(budget_bytes=24 << 20, slots=6)
    say(f"[gpu] server trace of {len(reqs)} requests")
    assert srv.run_trace(mixed) == F.SUCCESS
    a, s, d, st = mixed.times()
    assert (st == 0).all()
    # requests overlapped on the device: some started before an earlier one finished
    overlapped = sum(1 for i in range(1, len(reqs)) if s[i] < d[:i].max())
    say(f"[gpu] overlapped requests: {overlapped}, budget waits {srv.budget_waits()}, peak {srv.peak_bytes()}")
    assert overlapped > 0
    srv.close()
    alone = F.Trace(reqs)
    for q in alone.arr:
        q.arrival = 0.0
    assert F.serve_plain(alone) == F.SUCCESS
    for i in range(len(reqs)):
        me, ml = mixed.results(i)
        ae, al = alone.results(i)
        p = reqs[i]["precision"]
        assert O.results_equal(me, ml, ae, al, p).all(), f"request {i} differs alone vs in the mix"


def test_server_results_vs_oracle():
    reqs = W.served_requests(12, 9, rate=5000.0, max_ops=200, precisions=(512, 2048))
    tr = F.Trace(reqs)
    srv = F.Server(budget_bytes=64 << 20, slots=4)
    assert srv.run_trace(tr) == F.SUCCESS
    srv.close()
    for i, q in enumerate(reqs):
        re, rl = tr.results(i)
        p, n = q["precision"], q["precision"] // 32
        for op in np.unique(q["ops"]):
            for m in np.unique(q["modes"][q["ops"] == op]):
                idx = np.nonzero((q["ops"] == op) & (q["modes"] == m))[0]
                pos = (idx[:, None] * n + np.arange(n)[None, :]).reshape(-1)
                unary = INV[int(op)] == "sqrt"
                oe, ol = O.op(INV[int(op)], p, q["a_exps"][idx], q["a_limbs"][pos],
                              None if unary else q["b_exps"][idx], None if unary else q["b_limbs"][pos],
                              mode=MODES[int(m)])
                assert O.results_equal(re[idx], rl[pos], oe, ol, p).all(), (i, INV[int(op)])


def test_budget_admission_and_oversize_rejection():
    reqs = W.served_requests(30, 3, rate=50000.0, max_ops=1024, precisions=(4096,))
    tr = F.Trace(reqs)
    sizes = [F.request_bytes(tr, i) for i in range(len(reqs))]
    budget = max(sizes) + 4096      # room for about one large request at a time
    srv = F.Server(budget_bytes=budget, slots=8)
    assert srv.run_trace(tr) == F.SUCCESS
    assert srv.peak_bytes() <= budget
    assert srv.budget_waits() > 0
    srv.close()
    # a request larger than the whole budget is rejected, the others still complete
    big = W.served_requests(2, 4, rate=1e6, max_ops=4096, precisions=(16384,))
    big[0]["ops"] = big[0]["ops"][:1]
    for k in ("a_exps", "b_exps", "modes"):
        big[0][k] = big[0][k][:1]
    n = 16384 // 32
    big[0]["a_limbs"] = big[0]["a_limbs"][:n]
    big[0]["b_limbs"] = big[0]["b_limbs"][:n]
    tr2 = F.Trace(big)
    small = F.request_bytes(tr2, 0)
    srv = F.Server(budget_bytes=small + 256, slots=2)
    srv.run_trace(tr2)
    _, _, _, st = tr2.times()
    assert st[0] == F.SUCCESS
    assert st[1] == F.INVALID_COUNTS or F.request_bytes(tr2, 1) <= small + 256
    srv.close()
