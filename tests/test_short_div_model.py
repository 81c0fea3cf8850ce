"""CPU model of the short division's claim (core_div_lazy<K, L, true>, fpa_ops.h).

The lazy division keeps the remainder as signed column sums, one per radix-2^27 slot; a
digit step moves every slot up one (the top one folds into the new top slot), takes the
digit from the top two slots and adds q times the negated divisor digits to every slot.
The short division drops, after the first phase, the slots below the frame that a later
digit or the end's estimate could still read (slot s after step j is read again only if
s >= j - (S - ND + 1)).  Its claim: the digits and the top two slots at the end are exactly
those of the whole frame, so where the end's estimate is sure the results agree, and the
whole frame runs wherever it is not.  This model runs both on Python integers, with the
digit estimate formed in correctly rounded doubles as the kernel forms it (W = c1 + 32
floor(c0 / 2^32) converted to double, fma(W, 2^27 / Dt, -1/2), plus 1.5 * 2^52, low word),
and checks the claim for the phases the kernel uses (the shape formulas of div_short_shape).
"""
import random
import struct
from fractions import Fraction

import pytest

RADIX = 27
BETA = 1 << RADIX


def div_short_shape(K, L):
    """The phases of the short division for K limbs on L lanes (div_short_shape, fpa_ops.h)."""
    KD = (32 * K + RADIX - 1) // RADIX
    ND = KD * L
    S = (32 * (K * L + 1) + RADIX - 1) // RADIX
    C1 = S - ND + 1
    t3 = ND >= 300 or (ND >= 150 and K <= 24)
    TA = (S + KD + KD // 2) // (3 * KD) if t3 else (S + KD) // (2 * KD)
    JA = TA * KD
    KB = (ND - JA + 2 + C1 + L - 1) // L
    B = ND - KB * L
    ok = TA >= 1 and JA < S and KB >= 2 and KB < KD and B >= 0 and B <= JA - 2 - C1
    JB = JA + (S - JA) // 2
    KC = (ND - JB + 2 + C1 + L - 1) // L
    BC = ND - KC * L
    T3 = t3 and ok and KC >= 2 and KC < KB and BC >= B and BC <= JB - 2 - C1
    return dict(ND=ND, S=S, C1=C1, JA=JA, B=B, JB=JB if T3 else S, BC=BC if T3 else B, ok=ok)


def lazy_digit(w, invs):
    """round(W * invs - 1/2) as the kernel forms it: W to double, one fma, the 1.5*2^52 add"""
    x = float(Fraction(float(w)) * Fraction(invs) - Fraction(1, 2))
    y = x + 6755399441055744.0
    low = struct.unpack("<Q", struct.pack("<d", y))[0] & 0xFFFFFFFF
    return low - (1 << 32) if low >= 1 << 31 else low


def run(rd, dd, S, frames):
    """The digit steps on slots [bottom(j), ND): frames = [(first step, bottom slot), ...]"""
    ND = len(dd)
    nd = [-d for d in dd]
    C = list(rd)
    dt = float(dd[-1] * BETA + dd[-2])   # the fma of the top two divisor digits
    invs = (1.0 / dt) * 2.0 ** RADIX
    qs = []
    lo = 0
    for j in range(S):
        for first, bottom in frames:
            if j == first:
                lo = bottom
        so = C[ND - 1]
        for i in range(ND - 1, lo, -1):
            C[i] = C[i - 1]
        C[lo] = 0
        C[ND - 1] += so << RADIX
        q = lazy_digit(C[ND - 1] + 32 * (C[ND - 2] >> 32), invs)
        for i in range(lo, ND):
            C[i] += q * nd[i]
        qs.append(q)
    return qs, C[ND - 1], C[ND - 2]


@pytest.mark.parametrize("K,L", [(32, 4), (16, 8), (8, 16), (24, 8), (12, 8)])
def test_short_division_keeps_the_digits(K, L):
    sh = div_short_shape(K, L)
    assert sh["ok"] and sh["ND"] <= 255   # no carry levels at these lengths
    rng = random.Random(1000 * K + L)
    ND, S = sh["ND"], sh["S"]
    short = [(0, 0), (sh["JA"], sh["B"]), (sh["JB"], sh["BC"])]
    for case in range(24):
        dd = [rng.randrange(BETA) for _ in range(ND)]
        dd[-1] |= 1 << (RADIX - 1)
        if case % 4 == 1:   # a divisor with a long run of zero digits
            dd = [0] * (ND - 3) + dd[-3:]
        # a remainder below the divisor: the divisor's digits with the top one lowered
        rd = [rng.randrange(BETA) for _ in range(ND)]
        rd[-1] = rng.randrange(dd[-1])
        if case % 4 == 2:   # remainder and divisor sharing their top digits
            rd = dd[:]
            rd[0] = max(rd[0] - 1, 0) if any(rd[:-1]) else 0
            rd[-1] = dd[-1] - 1
        full = run(rd, dd, S, [(0, 0)])
        part = run(rd, dd, S, short)
        assert part == full, f"K={K} L={L} case {case}"


def test_a_dropped_slot_that_is_read_again_changes_the_digits():
    """The cut is tight: keeping one slot fewer than the rule allows changes a digit for some
    input (the rule is not vacuous)."""
    K, L = 16, 8
    sh = div_short_shape(K, L)
    ND, S, C1 = sh["ND"], sh["S"], sh["C1"]
    rng = random.Random(7)
    changed = False
    for _ in range(24):
        dd = [rng.randrange(BETA) for _ in range(ND)]
        dd[-1] |= 1 << (RADIX - 1)
        rd = [rng.randrange(BETA) for _ in range(ND)]
        rd[-1] = rng.randrange(dd[-1])
        full = run(rd, dd, S, [(0, 0)])
        # drop every slot below j - C1 + 8 after step j = 60: slots a later digit reads
        j = 60
        part = run(rd, dd, S, [(0, 0), (j, j - C1 + 8)])
        if part != full:
            changed = True
            break
    assert changed
