"""CPU check of the short (upper-half) product that mul uses (fpa_ops.h core_mul_short).

For every compiled mul shape with a short product, the lane schedules of
cubigfloat_amd/csrc/fpa_mul_sched.h are replayed on Python integers exactly as the kernel
sums them: each lane's two segments of chunk pairs (window of 2*KD columns, chunks below
F0 dropped, chunk F0 from column T on) plus the guard corners.  Pinned here:
  * the schedule covers every chunk pair with alpha + beta >= F0 exactly once, and no pair
    of the guard anti-diagonal F0-1 twice;
  * Q <= A*B < Q + 2^(27T+36) for random, all-ones and alphabet operands (the error bound
    the kernel's decision rests on);
  * whenever the 64 bits of Q below the round word are neither all zeros nor all ones,
    mantissa, round word and sticky = 1 taken from Q give the same rounded result as the
    exact product, in all five rounding modes.
The constants (T, G, corner) are recomputed here from their definition, not read from the
kernel, and compared with what the header's tables imply.
"""
import os
import random
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "cubigfloat_amd", "csrc")
RADIX = 27


def shapes():
    txt = open(os.path.join(CSRC, "fpa_shapes.h")).read()
    body = txt[txt.index("#define FPA_SHAPES(X)"):]
    body = body[:body.index("\n\n")]
    maxk = int(re.search(r"#define FPA_MUL_MAX_K (\d+)", txt).group(1))
    return [(int(n), int(k), int(l)) for n, k, l in re.findall(r"X\((\d+), (\d+), (\d+)\)", body) if int(k) <= maxk]


def schedules():
    """{(L, F0): (ok, R1, R2, lane words)}"""
    txt = open(os.path.join(CSRC, "fpa_mul_sched.h")).read()
    out = {}
    for m in re.finditer(r"k_mul_sched_(\d+)_(\d+)\[\d+\] = \{([^}]*)\}", txt):
        L, F0 = int(m.group(1)), int(m.group(2))
        words = [int(w, 16) for w in re.findall(r"0x([0-9a-f]+)u", m.group(3))]
        r = re.search(r"struct MulSched<%d, %d> \{[^}]*ok = (true|false);[^}]*R1 = (\d+), R2 = (\d+)" % (L, F0), txt)
        out[(L, F0)] = (r.group(1) == "true", int(r.group(2)), int(r.group(3)), words)
    return out


SCHED = schedules()


def params(K, L):
    KD = (32 * K + RADIX - 1) // RADIX
    ND = KD * L
    p = 32 * K * L
    tmax = (p - 133) // RADIX
    F0 = min(L - 1, tmax // KD) if L == 32 else L - 1
    G = min(tmax - KD * (F0 - 1), 2 * KD - 1)
    T = KD * (F0 - 1) + G
    ok = (L, F0) in SCHED and SCHED[(L, F0)][0] and F0 >= 1 and G >= KD and ND < 512 and 2 * L <= 64
    return dict(KD=KD, ND=ND, p=p, F0=F0, G=G, T=T, ok=ok)


SHORT = [(n, K, L) for (n, K, L) in shapes() if params(K, L)["ok"]]


def digits(x, ND):
    return [(x >> (RADIX * i)) & ((1 << RADIX) - 1) for i in range(ND)]


def short_q(A, B, K, L):
    """the kernel's sum: columns as integers, then Q = sum col * 2^(27 s)"""
    P = params(K, L)
    KD, ND, G, F0 = P["KD"], P["ND"], P["G"], P["F0"]
    _, R1, R2, words = SCHED[(L, F0)]
    a, b = digits(A, ND), digits(B, ND)
    cols = {}

    C0 = P["T"] - KD * F0

    def emit(c, w):
        if c >= F0:
            for i in range(KD):
                if c > F0 or i >= C0:     # chunk F0 keeps its columns from T on
                    cols[c * KD + i] = cols.get(c * KD + i, 0) + w[i]

    for sc in words:
        for (a0, b0, rs) in ((sc & 31, (sc >> 5) & 31, R1), ((sc >> 10) & 31, (sc >> 15) & 31, R2)):
            W = [0] * (2 * KD)
            for u in range(rs):
                al = a0 + u
                for i in range(KD):
                    for j in range(KD):
                        W[i + j] += a[al * KD + i] * b[b0 * KD + j]
                c = al + b0
                emit(c, W[:KD])
                if u == rs - 1:
                    emit(c + 1, W[KD:])
                else:
                    W = W[KD:] + [0] * KD
        if (sc >> 30) & 1:
            ca, cb = (sc >> 20) & 31, (sc >> 25) & 31
            assert ca + cb == F0 - 1
            for x in range(KD):
                for y in range(KD):
                    if x + y >= G:
                        s = KD * (F0 - 1) + x + y
                        cols[s] = cols.get(s, 0) + a[ca * KD + x] * b[cb * KD + y]
    for s, v in cols.items():
        assert v < 1 << 64, "column sum overflows 64 bits"
    return sum(v << (RADIX * s) for s, v in cols.items())


def round_ref(P, p, mode, sign):
    """mantissa (p bits), exponent adjust, from the exact product P < 2^(2p); fpa_round modes
    0 RNE, 1 RZ, 2 RU (+inf), 3 RD (-inf), 4 away from zero"""
    full = P >> (2 * p - 1)
    sh = p if full else p - 1
    mant = P >> sh
    rem = P - (mant << sh)
    half = 1 << (sh - 1)
    if mode == 0:
        inc = rem > half or (rem == half and (mant & 1))
    elif mode == 1:
        inc = False
    elif mode == 2:
        inc = rem != 0 and sign == 0
    elif mode == 3:
        inc = rem != 0 and sign == 1
    else:
        inc = rem != 0
    mant += int(inc)
    return mant, full


def round_from_q(Q, p, mode, sign):
    """the kernel's decided case: mantissa + round word from Q, sticky 1"""
    full = Q >> (2 * p - 1)
    sh = p if full else p - 1
    mant = Q >> sh
    rw = (Q >> (sh - 32)) & 0xFFFFFFFF
    if mode == 0:
        inc = rw > 0x80000000 or rw == 0x80000000     # sticky 1: at or above half is above
    elif mode == 1:
        inc = False
    elif mode == 2:
        inc = sign == 0
    elif mode == 3:
        inc = sign == 1
    else:
        inc = True
    return mant + int(inc), full


def decided(Q, p):
    w = (Q >> (p - 97)) & ((1 << 64) - 1)
    return w != 0 and w != (1 << 64) - 1


def operands(p, rng):
    top = 1 << (p - 1)
    yield rng.getrandbits(p) | top, rng.getrandbits(p) | top
    yield (1 << p) - 1, (1 << p) - 1
    alpha = [0, 1, 0x7FFFFFFF, 0x80000000, 0xFFFFFFFF]
    for _ in range(2):
        A = sum(rng.choice(alpha) << (32 * i) for i in range(p // 32)) | top
        B = sum(rng.choice(alpha) << (32 * i) for i in range(p // 32)) | top
        yield A, B
    yield top, top                                   # exact: Q's window is all zeros
    yield top | 1, (1 << p) - 1


def test_schedule_covers_each_pair_once():
    for (L, F0), (ok, R1, R2, words) in SCHED.items():
        assert len(words) == L
        seen = {}
        for sc in words:
            for (a0, b0, rs) in ((sc & 31, (sc >> 5) & 31, R1), ((sc >> 10) & 31, (sc >> 15) & 31, R2)):
                for u in range(rs):
                    al = a0 + u
                    assert 0 <= al < L and 0 <= b0 < L
                    if al + b0 >= F0 - 1:
                        assert (al, b0) not in seen
                        seen[(al, b0)] = "whole"
            if (sc >> 30) & 1:
                key = ((sc >> 20) & 31, (sc >> 25) & 31)
                assert sum(key) == F0 - 1 and key not in seen
                seen[key] = "corner"
        for al in range(L):
            for be in range(L):
                if al + be >= F0 - 1:
                    assert (al, be) in seen, (L, F0, al, be)
        # a usable schedule has fewer rounds than the L of the full product scan
        assert ok == (R1 + R2 < L)


def test_short_shapes_exist():
    # the 4096-bit mul shape of the headline workload is among them
    assert (128, 16, 8) in SHORT


@pytest.mark.parametrize("n,K,L", SHORT)
def test_short_product_bound_and_rounding(n, K, L):
    P = params(K, L)
    p, T = P["p"], P["T"]
    assert T >= P["KD"] * P["F0"] and RADIX * T + 36 <= p - 97
    rng = random.Random(1000 * K + L)
    for A, B in operands(p, rng):
        Q = short_q(A, B, K, L)
        E = A * B - Q
        assert 0 <= E < 1 << (RADIX * T + 36), (n, K, L)
        if decided(Q, p):
            for mode in range(5):
                for sign in (0, 1):
                    assert round_from_q(Q, p, mode, sign) == round_ref(A * B, p, mode, sign), (n, K, L, mode, sign)
    # the exact power-of-two product is never decided from Q (its window is all zeros)
    top = 1 << (p - 1)
    assert not decided(short_q(top, top, K, L), p)
