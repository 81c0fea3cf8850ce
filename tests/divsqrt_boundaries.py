"""div and sqrt operands whose exact results are representable (remainder zero) or sit one
unit in the last place of an operand away from such a result, and the correctly rounded
results computed with Python integers by the definition of each rounding mode, for the
oracle's CPU test and the GPU parity test.

A value is A * 2^(E - BIAS - p) with A a p-bit integer whose top bit is set; its exponent
word is 2E + sign."""
import math

import numpy as np

from oracle import oracle as O

MODES = ("rne", "rna", "rz", "ru", "rd")


def _norm(x, p):
    """x shifted left so its top set bit is bit p-1, and the shift"""
    s = p - x.bit_length()
    return x << s, s


def _limbs(xs, p):
    n = p // 32
    return np.array([[(x >> (32 * i)) & 0xFFFFFFFF for i in range(n)] for x in xs], np.uint32).reshape(-1)


def div_operands(p, cnt, seed):
    """a = u*v and b = v for short odd u, v (so a/b = u exactly), and a moved by one unit
    in its last place either way (a result just off a representable value)."""
    rng = np.random.default_rng(seed)
    half = max(2, p // 2 - 1)
    A, B = [], []
    while len(A) < cnt:
        u = int(rng.integers(1, 1 << 30)) | 1
        v = int(rng.integers(1, 1 << 30)) | 1
        for _ in range(int(rng.integers(0, half // 30))):
            u = (u << 30) | int(rng.integers(0, 1 << 30))
            v = (v << 30) | int(rng.integers(0, 1 << 30))
        u, v = u >> max(0, u.bit_length() - half), v >> max(0, v.bit_length() - half)
        a, _ = _norm(u * v, p)
        b, _ = _norm(v, p)
        for d in (0, 1, -1):
            x = a + d
            if x.bit_length() == p:
                A.append(x)
                B.append(b)
    A, B = A[:cnt], B[:cnt]
    ea = O.BIAS + rng.integers(-40, 40, size=cnt)
    eb = O.BIAS + rng.integers(-40, 40, size=cnt)
    sa, sb = rng.integers(0, 2, size=cnt), rng.integers(0, 2, size=cnt)
    return ((2 * ea + sa).astype(np.uint32), _limbs(A, p), (2 * eb + sb).astype(np.uint32), _limbs(B, p))


def sqrt_operands(p, cnt, seed):
    """x = r^2 for short r with an exponent that keeps the scaled value a square (so the
    root is exact), and x moved by one unit in its last place either way."""
    rng = np.random.default_rng(seed)
    half = max(2, p // 2 - 1)
    X, E = [], []
    while len(X) < cnt:
        r = int(rng.integers(1, 1 << 30))
        for _ in range(int(rng.integers(0, half // 30))):
            r = (r << 30) | int(rng.integers(0, 1 << 30))
        r >>= max(0, r.bit_length() - half)
        r = max(r, 1)
        x, s = _norm(r * r, p)
        # x * 2^(e - BIAS - p) = r^2 * 2^(s + e - BIAS - p): a square when s + e - p is even
        e = O.BIAS + int(rng.integers(-40, 40))
        if (s + e - O.BIAS - p) % 2:
            e += 1
        for d in (0, 1, -1):
            y = x + d
            if y.bit_length() == p:
                X.append(y)
                E.append(e)
    X, E = X[:cnt], E[:cnt]
    return (2 * np.array(E, np.int64)).astype(np.uint32), _limbs(X, p)


def _round(num, den, p, sign, mode):
    """num/den > 0 rounded to p bits: (M, k) with the result M * 2^-k"""
    k = p - (num.bit_length() - den.bit_length())
    while True:
        q, r = divmod(num << k, den) if k >= 0 else divmod(num, den << -k)
        if q >> p:
            k -= 1
        elif not (q >> (p - 1)):
            k += 1
        else:
            break
    up = {"rz": False, "ru": sign == 0 and r != 0, "rd": sign == 1 and r != 0,
          "rne": 2 * r > (den << max(0, -k)) or (2 * r == (den << max(0, -k)) and q & 1 == 1),
          "rna": 2 * r >= (den << max(0, -k))}[mode]
    if up:
        q += 1
        if q >> p:
            q >>= 1
            k -= 1
    return q, k


def _int(row):
    return sum(int(w) << (32 * i) for i, w in enumerate(row))


def _pack(results, p):
    n = p // 32
    re = np.array([2 * e + s for e, s, _ in results], np.uint32)
    rl = np.array([[(m >> (32 * i)) & 0xFFFFFFFF for i in range(n)] for _, _, m in results], np.uint32)
    return re, rl.reshape(-1)


def div_reference(p, ae, al, be, bl, mode):
    n = p // 32
    out = []
    for i in range(len(ae)):
        a, b = _int(al[i * n:(i + 1) * n]), _int(bl[i * n:(i + 1) * n])
        s = (int(ae[i]) ^ int(be[i])) & 1
        d = (int(ae[i]) >> 1) - (int(be[i]) >> 1)          # a/b = (A/B) * 2^d
        num, den = (a << d, b) if d >= 0 else (a, b << -d)
        m, k = _round(num, den, p, s, mode)
        out.append((O.BIAS + p - k, s, m))                 # m * 2^-k = m * 2^(E - BIAS - p)
    return _pack(out, p)


def sqrt_reference(p, ae, al, mode):
    n = p // 32
    out = []
    for i in range(len(ae)):
        a = _int(al[i * n:(i + 1) * n])
        t = (int(ae[i]) >> 1) - O.BIAS - p                 # value = a * 2^t
        if t % 2:
            a, t = a << 1, t - 1
        # sqrt(a) * 2^(t/2): m = isqrt(a * 4^k) with p bits
        k = (2 * p - a.bit_length() + 1) // 2
        while True:
            x = a << (2 * k) if k >= 0 else a >> (-2 * k)
            m = math.isqrt(x)
            if m >> p:
                k -= 1
            elif not (m >> (p - 1)):
                k += 1
            else:
                break
        rem = x - m * m                                    # k >= 0 here, so x is exact
        # no ties: (m + 1/2)^2 is never an integer; above half iff rem > m
        up = {"rz": False, "ru": rem != 0, "rd": False, "rne": rem > m, "rna": rem > m}[mode]
        if up:
            m += 1
            if m >> p:
                m >>= 1
                k -= 1
        out.append((O.BIAS + p + t // 2 - k, 0, m))
    return _pack(out, p)
