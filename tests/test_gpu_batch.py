"""Batched mixed-precision launches (fpa_batch.h) and dot chains: parity with the oracle
and with the plain per-op launches; graph replays are each checked."""
import numpy as np
import pytest

import cubigfloat_amd as F
from cubigfloat_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu
INV = {v: k for k, v in F.OPS.items()}
MODE_NAMES = ["rz", "ru", "rd", "rne", "rna"]


def say(*a):
    print(*a, flush=True)


def oracle_packed(ops, precs, modes, ae, al, be, bl):
    off = W.limb_offsets(precs)
    re = np.zeros(len(ops), np.uint32)
    rl = np.zeros(int(off[-1]), np.uint32)
    for op, p, m, idx, e1, l1, pos in W.split_by_op(ops, precs, modes, ae, al):
        _, _, _, _, e2, l2, _ = next(x for x in W.split_by_op(ops, precs, modes, be, bl) if x[:3] == (op, p, m))
        unary = INV[op] == "sqrt"
        oe, ol = O.op(INV[op], p, e1, l1, None if unary else e2, None if unary else l2, mode=MODE_NAMES[m])
        re[idx] = oe
        rl[pos] = ol
    return re, rl


def compare_packed(precs, ra, la, rb, lb):
    off = W.limb_offsets(precs)
    bad = []
    for i in range(len(precs)):
        n = int(precs[i]) // 32
        if ra[i] != rb[i] or (ra[i] > 8 and not np.array_equal(la[off[i]:off[i] + n], lb[off[i]:off[i] + n])):
            bad.append(i)
    return bad


def test_batch_mixed_ops_precisions_modes():
    ops, precs, modes = W.mixed_ops(3000, 1, precisions=F.BATCH_PRECISIONS, modes=(0, 1, 2, 3, 4))
    # keep 16384-bit ops few: the oracle is slow there
    keep = (precs < 16384) | (np.arange(len(precs)) % 16 == 0)
    ops, precs, modes = ops[keep], precs[keep], modes[keep]
    ae, al, be, bl = W.mixed_operands(ops, precs, 2)
    say(f"[gpu] batch nops={len(ops)}")
    B = F.Batch(ops, precs, modes)
    say(f"[gpu] batch plan {B.info()}")
    B.set_operands(ae, al, be, bl)
    oe, ol = oracle_packed(ops, precs, modes, ae, al, be, bl)
    for rep in range(3):
        assert B.run() == F.SUCCESS
        re, rl = B.results()
        bad = compare_packed(precs, re, rl, oe, ol)
        assert not bad, f"replay {rep}: {len(bad)} mismatches, first op {bad[0]}: op={INV[int(ops[bad[0]])]} p={precs[bad[0]]}"
    assert B.run_plain() == F.SUCCESS
    re, rl = B.results()
    assert not compare_packed(precs, re, rl, oe, ol)
    B.close()


def test_batch_config3_distribution_matches_plain():
    ops, precs, modes = W.mixed_ops(1 << 14, 3)
    ae, al, be, bl = W.mixed_operands(ops, precs, 4)
    say(f"[gpu] batch config-3 mix nops={len(ops)}")
    B = F.Batch(ops, precs, modes)
    B.set_operands(ae, al, be, bl)
    B.run_plain()
    pe, pl = B.results()
    for rep in range(2):
        B.run()
        re, rl = B.results()
        assert not compare_packed(precs, re, rl, pe, pl), f"replay {rep}"
    # a sample against the oracle
    sel = np.arange(0, len(ops), 37)
    sub_ops, sub_p, sub_m = ops[sel], precs[sel], modes[sel]
    off = W.limb_offsets(precs)
    def take(e, l):
        return e[sel], np.concatenate([l[off[i]:off[i + 1]] for i in sel])
    sae, sal = take(ae, al)
    sbe, sbl = take(be, bl)
    oe, ol = oracle_packed(sub_ops, sub_p, sub_m, sae, sal, sbe, sbl)
    sre, srl = take(re, rl)
    assert not compare_packed(sub_p, sre, srl, oe, ol)
    B.close()


def test_batch_rejects_bad_precision():
    with pytest.raises(ValueError):
        F.Batch([F.OPS["mul"]], [1536])


@pytest.mark.parametrize("branches", [1, 3, 8])
def test_batch_branch_counts_agree(branches):
    """Every branch count gives the same results as the plain launches; many small
    segments (all five modes at every precision) so each branch gets several nodes."""
    ops, precs, modes = W.mixed_ops(4000, 7, precisions=(256, 512, 1024, 2048, 3072, 4096),
                                    modes=(0, 1, 2, 3, 4))
    ae, al, be, bl = W.mixed_operands(ops, precs, 8)
    B = F.Batch(ops, precs, modes, branches=branches)
    say(f"[gpu] batch branches={branches} plan {B.info()}")
    assert B.info()["branches"] == min(branches, B.info()["segments"])
    B.set_operands(ae, al, be, bl)
    B.run_plain()
    pe, pl = B.results()
    for rep in range(3):
        B.run()
        re, rl = B.results()
        assert not compare_packed(precs, re, rl, pe, pl), f"replay {rep}"
    B.close()


@pytest.mark.parametrize("p", [256, 2048, 4096])
def test_dot_chains_vs_oracle(p):
    rng = np.random.default_rng(p)
    lengths = rng.integers(0, 40, 150).astype(np.uint32)
    lengths[:3] = [0, 1, 64]
    total = int(lengths.sum())
    xe, xl, ye, yl = F.generate(2, 9, "mul", p, total)
    for mode in ("rne", "rz", "rd"):
        say(f"[gpu] dot p={p} chains={len(lengths)} mode={mode}")
        re, rl = F.dot_chains(p, lengths, xe, xl, ye, yl, mode=F.MODES[mode])
        oe, ol = O.dot(p, lengths, xe, xl, ye, yl, mode=mode)
        eq = O.results_equal(re, rl, oe, ol, p)
        assert eq.all(), f"{int((~eq).sum())} mismatches"


@pytest.mark.parametrize("p", [512, 2048, 8192])
def test_dot_every_shape_vs_oracle(p):
    rng = np.random.default_rng(p + 1)
    lengths = rng.integers(0, 24, 70).astype(np.uint32)
    total = int(lengths.sum())
    xe, xl, ye, yl = F.generate(2, 5, "mul", p, total)
    oe, ol = O.dot(p, lengths, xe, xl, ye, yl, mode="rne")
    keep = F.get_variant("dot", p)
    try:
        for v, (k, l) in enumerate(F.variants("dot", p)):
            F.set_variant("dot", p, v)
            say(f"[gpu] dot p={p} shape K={k} L={l}")
            re, rl = F.dot_chains(p, lengths, xe, xl, ye, yl)
            assert O.results_equal(re, rl, oe, ol, p).all(), (k, l)
    finally:
        F.set_variant("dot", p, keep)


@pytest.mark.parametrize("p", [2048, 4096, 8192])
def test_dot_every_shape_exact_products(p):
    """Products whose low half is zero (two-limb mantissas, an exact product) mixed with
    random ones, so within a wave some groups cannot decide their rounding from the
    short product and form their product in full while the others go on, on chains of
    different lengths."""
    rng = np.random.default_rng(p + 2)
    lengths = rng.integers(0, 20, 90).astype(np.uint32)
    total = int(lengths.sum())
    n = p // 32
    xe, xl, ye, yl = F.generate(2, 6, "mul", p, total)
    xl, yl = xl.reshape(-1, n).copy(), yl.reshape(-1, n).copy()
    short = rng.random(total) < 0.4
    xl[short, :n - 2] = 0
    yl[short, :n - 2] = 0
    xl, yl = xl.reshape(-1), yl.reshape(-1)
    for mode in ("rne", "ru"):
        oe, ol = O.dot(p, lengths, xe, xl, ye, yl, mode=mode)
        keep = F.get_variant("dot", p)
        try:
            for v, (k, l) in enumerate(F.variants("dot", p)):
                F.set_variant("dot", p, v)
                say(f"[gpu] dot exact products p={p} shape K={k} L={l} mode={mode}")
                re, rl = F.dot_chains(p, lengths, xe, xl, ye, yl, mode=F.MODES[mode])
                assert O.results_equal(re, rl, oe, ol, p).all(), (k, l, mode)
        finally:
            F.set_variant("dot", p, keep)


def test_dot_config4_shape_sample():
    lengths, xe, xl, ye, yl = W.dot_workload(256, 2048, 11)
    say(f"[gpu] dot config-4 sample chains=256 elements={int(lengths.sum())}")
    re, rl = F.dot_chains(2048, lengths, xe, xl, ye, yl)
    me, ml = O.mpfr_dot(2048, lengths, xe, xl, ye, yl)
    assert O.results_equal(re, rl, me, ml, 2048).all()


# ------------------------- persistent batch (one kernel launch) -------------------------
def test_persistent_batch_vs_oracle_many_tiles_per_wave():
    """The persistent kernel on a batch with far more tiles than waves (so waves take many
    tiles from the queue), every replay checked, eager launches between replays."""
    ops, precs, modes = W.mixed_ops(24000, 11, precisions=F.PERSIST_PRECISIONS, modes=(3,))
    ae, al, be, bl = W.mixed_operands(ops, precs, 12)
    say(f"[gpu] persistent batch nops={len(ops)}")
    B = F.Batch(ops, precs, modes, persistent=True)
    info = B.info()
    say(f"[gpu] persistent plan {info}")
    assert info["persistent"] and info["tiles"] > 4 * info["grid"]
    B.set_operands(ae, al, be, bl)
    oe, ol = oracle_packed(ops, precs, modes, ae, al, be, bl)
    A = F.FpaArray(4096, 1024)
    try:
        for rep in range(4):
            assert B.run() == F.SUCCESS
            re, rl = B.results()
            bad = compare_packed(precs, re, rl, oe, ol)
            assert not bad, f"replay {rep}: {len(bad)} mismatches, first op {bad[0]}: op={INV[int(ops[bad[0]])]} p={precs[bad[0]]}"
            assert F.mul(A, A, A) == F.SUCCESS      # eager work between replays
        assert B.run_plain() == F.SUCCESS
        re, rl = B.results()
        assert not compare_packed(precs, re, rl, oe, ol)
    finally:
        A.close()
        B.close()


def test_persistent_batch_equals_graph_batch_all_modes():
    ops, precs, modes = W.mixed_ops(6000, 13, precisions=(256, 1024, 4096), modes=(1, 4))
    ae, al, be, bl = W.mixed_operands(ops, precs, 14)
    say(f"[gpu] persistent vs graph nops={len(ops)}")
    G = F.Batch(ops, precs, modes)
    P = F.Batch(ops, precs, modes, persistent=True)
    try:
        for b in (G, P):
            b.set_operands(ae, al, be, bl)
            assert b.run() == F.SUCCESS
        ge, gl = G.results()
        pe, pl = P.results()
        assert np.array_equal(ge, pe) and np.array_equal(gl, pl)
        oe, ol = oracle_packed(ops, precs, modes, ae, al, be, bl)
        assert not compare_packed(precs, pe, pl, oe, ol)
    finally:
        G.close()
        P.close()


def test_persistent_batch_limits():
    ops = np.array([F.OPS["mul"]] * 4, np.uint8)
    with pytest.raises(ValueError):
        F.Batch(ops, np.array([16384] * 4, np.uint32), persistent=True)      # no kind for 16384 bits
    n = F.PERSIST_MAX_SEGMENTS + 1
    ops2 = np.array([F.OPS[o] for o in ("add", "sub", "mul", "div", "sqrt")] * 9, np.uint8)[:n]
    precs2 = np.array([p for p in F.PERSIST_PRECISIONS for _ in range(5)] * 2, np.uint32)[:n]
    modes2 = np.array(([3] * 30 + [0] * 30), np.uint8)[:n]
    assert len(set(zip(ops2.tolist(), precs2.tolist(), modes2.tolist()))) == n
    with pytest.raises(ValueError):
        F.Batch(ops2, precs2, modes2, persistent=True)                       # more segments than the table holds
