"""CPU model of the short root's claim (core_sqrt_lazy<K, L, true>, fpa_ops.h).

The lazy root keeps U = M - S^2 (scaled) as signed column sums, one per radix-2^27 slot, and
T = 2S negated in the same slots.  A step moves every slot up one (the top one folds into
the new top slot), takes the digit from the top two slots, subtracts q * T with the new
digit's own slot set to q (so q^2 goes too) and then sets that slot of T to 2q.  The short
root applies a step's products only on the frame of top slots it keeps: the slots below the
frame keep moving up (they carry radicand digits, which enter the frame at its bottom), and
once T reaches below the frame the products there are dropped.  Its claim: no later digit and
not the end's estimate reads what is dropped, so the digits and the top two slots of U at the
end are those of the whole frame.  This model runs both on Python integers with the digit
estimates formed in correctly rounded doubles as the kernel forms them (step 0: integer square
root of the top slot; step 1: the quadratic in FP64; later steps: the lazy estimate with
2^27 / (T's top two digits)) and checks the claim on the kernel's frames (sqrt_short_cols and
the quarter frames of long roots), without carry levels (Python integers need none).
"""
import math
import random
import struct
from fractions import Fraction

import pytest

RADIX = 27
BETA = 1 << RADIX


def frames(K, L):
    """(first step, bottom slot) of the short root's frames for K limbs on L lanes"""
    KD = (32 * K + RADIX - 1) // RADIX
    F = KD * L
    KB = (F + 2 + 2 * L - 1) // (2 * L)
    FB = KB * L
    B0 = F - FB
    assert KB >= 2 and KB < KD and B0 >= 1 and 2 * B0 <= F - 2
    kq = KB % 2 == 0 and L % 2 == 0 and KB // 2 >= 2 and (F >= 300 or (F >= 150 and K <= 24))
    if not kq:
        return F, [(0, B0)]
    FQ = (KB // 2) * L
    B3 = F - FQ
    return F, [(0, F - FQ), (FQ, B0), (B3 + 3, B3)]


def rnd(x):
    return float(x)


def lazy_digit(w, invs):
    x = float(Fraction(float(w)) * Fraction(invs) - Fraction(1, 2))
    y = x + 6755399441055744.0
    low = struct.unpack("<Q", struct.pack("<d", y))[0] & 0xFFFFFFFF
    return low - (1 << 32) if low >= 1 << 31 else low


def run(ad, fr):
    """all F steps; fr: [(first step, bottom slot)]"""
    F = len(ad)
    C = list(ad)
    nT = [0] * F
    qs = []
    q0 = 0
    invs = 0.0
    lo = 0
    for k in range(F):
        for first, bottom in fr:
            if k == first:
                lo = bottom
        so = C[F - 1]
        for i in range(F - 1, 0, -1):
            C[i] = C[i - 1]
        C[0] = 0
        C[F - 1] += so << RADIX
        if k == 0:
            q = math.isqrt(C[F - 1])
        elif k == 1:
            vs = float(Fraction(rnd(C[F - 1])) * Fraction(1, 32) + Fraction(C[F - 2] >> 32))
            a = float(q0 * BETA)
            v = vs * 2.0 ** 32
            q = math.floor(v / (a + math.sqrt(float(Fraction(a) * Fraction(a) + Fraction(v)))))
        else:
            q = lazy_digit(C[F - 1] + 32 * (C[F - 2] >> 32), invs)
        pos = F - 1 - k
        if pos >= lo:
            nT[pos] = -q
        for i in range(lo, F):
            C[i] += q * nT[i]
        nT[pos] = -2 * q
        if k == 0:
            q0 = q
        if k == 1:
            invs = 2.0 ** RADIX / rnd(2 * q0 * BETA + 2 * q)
        qs.append(q)
    return qs, C[F - 1], C[F - 2]


@pytest.mark.parametrize("K,L", [(32, 4), (16, 8), (8, 16), (16, 2), (8, 4)])
def test_short_root_keeps_the_digits(K, L):
    F, fr = frames(K, L)
    rng = random.Random(100 * K + L)
    for case in range(16):
        # the radicand's digits, top-aligned, its top digit large enough for a full-length root
        ad = [rng.randrange(BETA) for _ in range(F)]
        ad[-1] = rng.randrange(1 << (RADIX - 2), BETA)
        if case % 4 == 1:      # a radicand with a long run of zero digits
            ad = [0] * (F - 2) + ad[-2:]
        if case % 4 == 2:      # all-ones digits
            ad = [BETA - 1] * F
        full = run(ad, [(0, 0)])
        part = run(ad, fr)
        assert part == full, f"K={K} L={L} case {case}"


def test_a_dropped_slot_that_is_read_again_changes_the_digits():
    F, fr = frames(32, 4)
    rng = random.Random(5)
    changed = False
    for _ in range(16):
        ad = [rng.randrange(BETA) for _ in range(F)]
        ad[-1] = rng.randrange(1 << (RADIX - 2), BETA)
        full = run(ad, [(0, 0)])
        # a frame 8 slots above what the rule allows from step 100 on
        part = run(ad, fr + [(100, 100 - 1 + 8)])
        if part != full:
            changed = True
            break
    assert changed
