"""Synthetic workloads of the batched, chained and served configurations.

Mixed batch: each op draws its precision uniformly from a class list (log-uniform over
powers of two), its op from a mix, and its operands from the config-2 operand model of
fpa_workload_generate (kind 2: e ~ U[-1024, 1024], 1% zeros, 5% near-cancellation
pairs), all from a seed.  Host arrays are packed in op order (fpa_batch.h layout).
"""
import numpy as np

from . import OPS, generate

CONFIG3_PRECISIONS = (256, 512, 1024, 2048, 4096, 8192)
CONFIG3_MIX = (("mul", 0.40), ("add", 0.15), ("sub", 0.15), ("div", 0.20), ("sqrt", 0.10))


def mixed_ops(nops, seed, precisions=CONFIG3_PRECISIONS, mix=CONFIG3_MIX, modes=None):
    """Returns (ops uint8, precisions uint32, modes uint8) for nops ops."""
    rng = np.random.default_rng(seed)
    names = [m[0] for m in mix]
    probs = np.array([m[1] for m in mix], float)
    op_idx = rng.choice(len(names), size=nops, p=probs / probs.sum())
    ops = np.array([OPS[names[i]] for i in op_idx], np.uint8)
    precs = np.array(precisions, np.uint32)[rng.integers(0, len(precisions), nops)]
    if modes is None:
        md = np.full(nops, 3, np.uint8)
    else:
        md = np.array(modes, np.uint8)[rng.integers(0, len(modes), nops)]
    return ops, precs, md


def limb_offsets(precisions):
    n = np.asarray(precisions, np.uint64) // 32
    off = np.zeros(len(n) + 1, np.uint64)
    off[1:] = np.cumsum(n)
    return off


def mixed_operands(ops, precisions, seed, kind=2):
    """Operands for a mixed op list, packed in op order: (a_exps, a_limbs, b_exps, b_limbs)."""
    inv = {v: k for k, v in OPS.items()}
    off = limb_offsets(precisions)
    total = int(off[-1])
    ae = np.zeros(len(ops), np.uint32)
    be = np.zeros(len(ops), np.uint32)
    al = np.zeros(total, np.uint32)
    bl = np.zeros(total, np.uint32)
    for op in np.unique(ops):
        for p in np.unique(precisions[ops == op]):
            idx = np.nonzero((ops == op) & (precisions == p))[0]
            n = int(p) // 32
            e1, l1, e2, l2 = generate(kind, seed * 1000003 + int(op) * 131 + int(p), inv[int(op)], int(p), len(idx),
                                      binary=int(op) != OPS["sqrt"])
            pos = (off[idx][:, None] + np.arange(n, dtype=np.uint64)[None, :]).reshape(-1)
            ae[idx] = e1
            al[pos] = l1
            if e2 is not None:
                be[idx] = e2
                bl[pos] = l2
    return ae, al, be, bl


def split_by_op(ops, precisions, modes, exps, limbs):
    """Yield (op, precision, mode, idx, exps[idx], limbs of idx) groups of a packed array."""
    off = limb_offsets(precisions)
    for op in np.unique(ops):
        for p in np.unique(precisions[ops == op]):
            for m in np.unique(modes[(ops == op) & (precisions == p)]):
                idx = np.nonzero((ops == op) & (precisions == p) & (modes == m))[0]
                n = int(p) // 32
                pos = (off[idx][:, None] + np.arange(n, dtype=np.uint64)[None, :]).reshape(-1)
                yield int(op), int(p), int(m), idx, exps[idx], limbs[pos], pos


def dot_workload(nchains, precision, seed, lmin=8, lmax=512):
    """Chain lengths ~ U[lmin, lmax]; elements: random mantissas, e ~ U[-32, 32], random signs
    (fpa_workload_generate kind 1 draws e in [-64, 64]; exponents are re-drawn here)."""
    rng = np.random.default_rng(seed)
    lengths = rng.integers(lmin, lmax + 1, nchains).astype(np.uint32)
    total = int(lengths.sum())
    xe, xl, ye, yl = generate(1, seed, "mul", precision, total)
    from . import BIAS
    for e in (xe, ye):
        sign = e & 1
        ex = rng.integers(-32, 33, total)
        e[:] = (2 * (BIAS + ex) + sign).astype(np.uint32)
    return lengths, xe, xl, ye, yl


SERVE_PRECISIONS = (256, 512, 1024, 2048, 4096, 8192, 16384)


def served_requests(nreq, seed, rate, max_ops=4096, precisions=SERVE_PRECISIONS, mix=CONFIG3_MIX):
    """Poisson arrivals at `rate` requests/s; request size log-uniform in [1, max_ops];
    one precision per request, uniform over the (power-of-two) class list; op mix as the
    mixed batch.  Returns a list of request dicts for cubigfloat_amd.Trace."""
    rng = np.random.default_rng(seed)
    t = 0.0
    reqs = []
    for i in range(nreq):
        t += rng.exponential(1.0 / rate) if rate > 0 else 0.0
        nops = int(np.floor(np.exp(rng.uniform(0, np.log(max_ops + 1)))))
        nops = min(max(nops, 1), max_ops)
        p = int(precisions[rng.integers(0, len(precisions))])
        ops, _, modes = mixed_ops(nops, seed * 7919 + i, precisions=(p,), mix=mix)
        precs = np.full(nops, p, np.uint32)
        ae, al, be, bl = mixed_operands(ops, precs, seed * 104729 + i)
        reqs.append({"precision": p, "ops": ops, "modes": modes, "a_exps": ae, "a_limbs": al,
                     "b_exps": be, "b_limbs": bl, "arrival": t})
    return reqs
