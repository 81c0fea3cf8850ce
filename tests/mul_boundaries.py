"""Mul operands whose exact products sit on the rounding boundaries, and the correctly
rounded product computed with Python integers (the definition of each rounding mode), for
the oracle's CPU test and the GPU parity test."""
import numpy as np

from oracle import oracle as O

MODES = ("rne", "rna", "rz", "ru", "rd")


def operands(p, per_class, seed):
    """Mantissa pairs whose exact product sits on each rounding boundary: bits below the
    kept p are zero, exactly half (with the kept part even and odd), just below and just
    above half, each with the product's top bit clear (p kept from bit p-1) and set (from
    bit p).  Operands are 2^(p-1) plus or 2^p minus a few powers of two, so the product's
    low part is sparse and every class turns up quickly; classes are decided on the
    exact product with Python integers."""
    rng = np.random.default_rng(seed)
    n = p // 32
    top = 1 << (p - 1)
    want = {(c, f): per_class for c in ("exact", "tie_even", "tie_odd", "below", "above") for f in (0, 1)}
    pairs = []
    for _ in range(400 * per_class):
        if not any(want.values()):
            break
        bits = lambda k: sum(1 << int(b) for b in rng.integers(0, p - 1, size=k))
        a = top + bits(int(rng.integers(1, 4)))
        if rng.integers(0, 2):
            b = top + bits(int(rng.integers(1, 4)))
        else:
            b = (1 << p) - bits(int(rng.integers(1, 3)))
            if b < top:
                continue
        if a >= 1 << p:
            continue
        aim = int(rng.integers(0, 4))
        if aim == 0:
            # a tie with the top bit clear: (2^(p-1) + 2^i)(2^(p-1) + 2^(p-2-i)); the kept
            # part is odd for i = 0 or p-2
            i = int(rng.choice([0, p - 2])) if rng.integers(0, 2) else int(rng.integers(0, p - 1))
            a, b = top + (1 << i), top + (1 << (p - 2 - i))
        elif aim == 1:
            # a tie with the top bit set: (2^p - 2^i)(2^(p-1) + 2^(p-1-i)), 1 <= i <= (p-1)/2;
            # the kept part is even for i = 1
            i = 1 if rng.integers(0, 2) else int(rng.integers(1, (p - 1) // 2 + 1))
            a, b = (1 << p) - (1 << i), top + (1 << (p - 1 - i))
        elif aim == 2:
            # top bit set, low part 2^(p-1) - x*y just below half: b = 2^p - y with y odd
            i = int(rng.integers(2, p - 1))
            j = int(rng.integers(1, i))
            a, b = top + (1 << i), (1 << p) - 1 - (1 << j)
        prod = a * b
        full = prod >> (2 * p - 1)
        cut = p if full else p - 1
        low, half = prod & ((1 << cut) - 1), 1 << (cut - 1)
        odd = (prod >> cut) & 1
        cls = ("exact" if low == 0 else "tie_odd" if low == half and odd else "tie_even" if low == half
               else "below" if low < half else "above")
        if want[(cls, full)]:
            want[(cls, full)] -= 1
            pairs.append((a, b))
    missing = [k for k, v in want.items() if v]
    assert not missing, f"p={p}: no operands found for {missing}"
    cnt = len(pairs)
    limbs = lambda x: [(x >> (32 * i)) & 0xFFFFFFFF for i in range(n)]
    al = np.array([limbs(a) for a, _ in pairs], np.uint32).reshape(-1)
    bl = np.array([limbs(b) for _, b in pairs], np.uint32).reshape(-1)
    ea = O.BIAS + rng.integers(-40, 40, size=cnt)
    eb = O.BIAS + rng.integers(-40, 40, size=cnt)
    sa, sb = rng.integers(0, 2, size=cnt), rng.integers(0, 2, size=cnt)
    return (2 * ea + sa).astype(np.uint32), al, (2 * eb + sb).astype(np.uint32), bl


def _int(row):
    return sum(int(w) << (32 * i) for i, w in enumerate(row))


def reference(p, ae, al, be, bl, mode):
    """The correctly rounded products: sign = xor of the signs, exponent = sum less the bias
    (less one when the product's top bit is clear), p bits kept from the top set bit and
    rounded by the mode's rule on the exact remainder; (exponent words, limbs)."""
    n = p // 32
    cnt = len(ae)
    A, B = al.reshape(cnt, n), bl.reshape(cnt, n)
    re = np.zeros(cnt, np.uint32)
    rl = np.zeros((cnt, n), np.uint32)
    for k in range(cnt):
        s = (int(ae[k]) ^ int(be[k])) & 1
        prod = _int(A[k]) * _int(B[k])
        full = prod >> (2 * p - 1)
        cut = p if full else p - 1
        keep, low, half = prod >> cut, prod & ((1 << cut) - 1), 1 << (cut - 1)
        up = {"rz": False, "ru": s == 0 and low != 0, "rd": s == 1 and low != 0,
              "rne": low > half or (low == half and keep & 1 == 1), "rna": low >= half}[mode]
        e = (int(ae[k]) >> 1) + (int(be[k]) >> 1) - O.BIAS - (0 if full else 1)
        if up:
            keep += 1
            if keep >> p:
                keep >>= 1
                e += 1
        re[k] = 2 * e + s
        rl[k] = [(keep >> (32 * i)) & 0xFFFFFFFF for i in range(n)]
    return re, rl.reshape(-1)
