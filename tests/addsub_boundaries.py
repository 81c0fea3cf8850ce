"""add and sub operands whose exact sums sit on the rounding boundaries (the smaller
operand's bits below the larger one's last place are exactly half, just below or above
half, or zero; exponent gaps 0..p+2; same and opposite signs, so carries out and
cancellations move the rounding position), and the correctly rounded sums computed with
Python integers by the definition of each rounding mode.

A value is A * 2^(E - BIAS - p) with A a p-bit integer whose top bit is set; its exponent
word is 2E + sign.  Pairs whose exact sum is zero are left out (their sign is a
convention, covered by the special-class tests)."""
import numpy as np

from oracle import oracle as O

MODES = ("rne", "rna", "rz", "ru", "rd")


def operands(p, cnt, seed):
    rng = np.random.default_rng(seed)
    n = p // 32
    top = 1 << (p - 1)
    gaps = [0, 1, 2, 3, 31, 32, 33, p - 2, p - 1, p, p + 1, p + 2]
    A, B, EA, EB, SA, SB = [], [], [], [], [], []
    while len(A) < cnt:
        g = int(gaps[rng.integers(0, len(gaps))])
        a = top | int.from_bytes(rng.bytes(n * 4), "little") & (top - 1)
        if rng.integers(0, 3) == 0:
            a = top | (a & ((1 << int(rng.integers(1, p))) - 1))       # sparse low part
        # b's low min(g, p) bits (those below a's last place) form the boundary pattern
        low = min(g, p)
        rest = int.from_bytes(rng.bytes(n * 4), "little") >> low << low
        if low == 0:
            tail = 0
        else:
            pat = int(rng.integers(0, 4))
            tail = [1 << (low - 1), (1 << (low - 1)) - 1, (1 << (low - 1)) | 1, 0][pat]
        b = (rest | tail | top) & ((1 << p) - 1)
        if g == 0 and rng.integers(0, 2):
            b = a ^ (1 << int(rng.integers(0, p - 1)))                  # near-cancellation
        ea = O.BIAS + int(rng.integers(-40, 40))
        sa, sb = int(rng.integers(0, 2)), int(rng.integers(0, 2))
        eb = ea - g
        if rng.integers(0, 2):
            a, b, ea, eb, sa, sb = b, a, eb, ea, sb, sa
        x = (-1) ** sa * (a << (ea - min(ea, eb))) + (-1) ** sb * (b << (eb - min(ea, eb)))
        y = (-1) ** sa * (a << (ea - min(ea, eb))) - (-1) ** sb * (b << (eb - min(ea, eb)))
        if x == 0 or y == 0:
            continue
        A.append(a); B.append(b); EA.append(ea); EB.append(eb); SA.append(sa); SB.append(sb)
    limbs = lambda xs: np.array([[(x >> (32 * i)) & 0xFFFFFFFF for i in range(n)] for x in xs],
                                np.uint32).reshape(-1)
    ae = (2 * np.array(EA, np.int64) + np.array(SA)).astype(np.uint32)
    be = (2 * np.array(EB, np.int64) + np.array(SB)).astype(np.uint32)
    return ae, limbs(A), be, limbs(B)


def _int(row):
    return sum(int(w) << (32 * i) for i, w in enumerate(row))


def reference(op, p, ae, al, be, bl, mode):
    n = p // 32
    re = np.zeros(len(ae), np.uint32)
    rl = np.zeros((len(ae), n), np.uint32)
    for i in range(len(ae)):
        a, b = _int(al[i * n:(i + 1) * n]), _int(bl[i * n:(i + 1) * n])
        ea, eb = int(ae[i]) >> 1, int(be[i]) >> 1
        sa, sb = int(ae[i]) & 1, (int(be[i]) & 1) ^ (1 if op == "sub" else 0)
        e0 = min(ea, eb)
        s = (-1) ** sa * (a << (ea - e0)) + (-1) ** sb * (b << (eb - e0))   # sum * 2^(e0-BIAS-p)
        sign = 1 if s < 0 else 0
        m = abs(s)
        k = p - m.bit_length()                                    # m * 2^k has p bits
        if k >= 0:
            q, r, den = m << k, 0, 1
        else:
            q, r, den = m >> -k, m & ((1 << -k) - 1), 1 << -k
        up = {"rz": False, "ru": sign == 0 and r != 0, "rd": sign == 1 and r != 0,
              "rne": 2 * r > den or (2 * r == den and q & 1 == 1), "rna": 2 * r >= den}[mode]
        if up:
            q += 1
            if q >> p:
                q >>= 1
                k -= 1
        re[i] = 2 * (e0 - k) + sign                               # q * 2^(e0 - k - BIAS - p)
        rl[i] = [(q >> (32 * j)) & 0xFFFFFFFF for j in range(n)]
    return re, rl.reshape(-1)
