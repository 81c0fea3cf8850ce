"""Writes fpa_mul_sched.h: the lane schedules of the short (upper-half) product in mul.

A value's operands are L digit chunks each (chunk = one lane's KD radix-2^27 digits).
The upper part of the product needs every chunk pair (alpha, beta) with
alpha + beta >= F0 (full pairs; F0 = L - 1 for most shapes, lower where the rounding
region reaches further down) and, on anti-diagonal alpha + beta = F0 - 1 (guard pairs),
the corner of digit products whose column reaches the rounding region.  Each lane
runs two segments of R1 and R2 rounds (the same for every lane, so all lanes flush their
windows together); a segment keeps one b chunk (beta) and walks alpha upward, so its
column window slides one chunk per round.  Row beta's full pairs are alpha in
[max(0, F0-beta), L-1]; a row may be extended downward, first by its guard pair (then
computed whole, so it needs no corner), then by pairs of lower anti-diagonals, whose
columns the kernel drops.  A guard pair no segment covers gets a corner on one lane.

The schedule packs every row into pieces of R1 and R2 rounds, exactly L of each, with as
few rounds R1 + R2 as it can (at least the full pairs over L), for groups of up to 8 lanes the fewest
lanes finishing the same chunk in the same round (their LDS atomics on one word run one
after another; a chunk of n pairs over R rounds forces ceil(n/R)), then as many guard
pairs computed whole as it can.
Run: python gen_mul_sched.py > fpa_mul_sched.h
"""
import sys
import time


def row_splits(L, F0, beta, R1, R2):
    """ways to cover row beta: (extension e, [(piece length, alpha_start), ...])"""
    out = []
    s0 = max(0, F0 - beta)
    for e in range(0, s0 + 1):
        a0 = s0 - e

        def rec(pos, acc):
            if pos == L:
                out.append((e, list(acc)))
                return
            for m in (R1, R2):
                if pos + m <= L:
                    acc.append((m, pos))
                    rec(pos + m, acc)
                    acc.pop()
        rec(a0, [])
    return out


def collisions(L, F0, R1, R2, pieces):
    """sum over the emission instructions (each round's finished chunk, and the flush's
    second chunk) of the most lanes of a group adding into one kept chunk; the LDS runs
    same-word atomics of one instruction one after another"""
    cost = 0
    for rk in (R1, R2):
        starts = [a + b for (b, a, m) in pieces if m == rk]
        for u in range(rk + 1):
            cnt = {}
            for d in starts:
                if d + u >= F0:
                    cnt[d + u] = cnt.get(d + u, 0) + 1
            cost += max(cnt.values()) if cnt else 0
    return cost


def guards_covered(L, F0, pieces):
    return {b for (b, a, m) in pieces if a + b <= F0 - 1 < a + b + m and b <= F0 - 1}


def solve(L, F0, R1, R2, budget_s=20.0):
    """Row pieces, exactly L of length R1 and L of length R2.  Small groups: exhaustive
    search for the fewest atomic collisions, then the most guard pairs whole.  Large groups:
    the most guard pairs whole (dynamic programme over the rows), collisions not ranked.
    Returns [(beta, alpha_start, length), ...] or None."""
    splits = [row_splits(L, F0, b, R1, R2) for b in range(L)]
    if L <= 8:
        best = [None, None]
        t0 = time.time()

        def rec(beta, n1, n2, pieces):
            if time.time() - t0 > budget_s:
                return
            if beta < 0:
                if n1 == L and n2 == L:
                    key = (collisions(L, F0, R1, R2, pieces), -len(guards_covered(L, F0, pieces)))
                    if best[1] is None or key < best[1]:
                        best[0], best[1] = list(pieces), key
                return
            for e, ps in splits[beta]:
                k1 = sum(1 for m, a in ps if m == R1)
                if n1 + k1 > L or n2 + len(ps) - k1 > L:
                    continue
                rec(beta - 1, n1 + k1, n2 + len(ps) - k1, pieces + [(beta, a, m) for m, a in ps])

        rec(L - 1, 0, 0, [])
        return best[0]

    # large groups: most guards, by rows with (pieces used) as state
    memo = {}

    def best_from(beta, u1, u2):
        if beta == L:
            return (0, []) if (u1 == L and u2 == L) else None
        key = (beta, u1, u2)
        if key in memo:
            return memo[key]
        top = None
        for e, ps in splits[beta]:
            k1 = sum(1 for m, a in ps if m == R1)
            if u1 + k1 > L or u2 + len(ps) - k1 > L:
                continue
            r = best_from(beta + 1, u1 + k1, u2 + len(ps) - k1)
            if r is None:
                continue
            gd = 1 if (e >= 1 and beta <= F0 - 1) else 0
            if top is None or r[0] + gd > top[0]:
                top = (r[0] + gd, [(beta, a, m) for m, a in ps] + r[1])
        memo[key] = top
        return top

    r = best_from(0, 0, 0)
    return None if r is None else r[1]


def full_pairs(L, F0):
    return sum(1 for a in range(L) for b in range(L) if a + b >= F0)


def schedule(L, F0):
    R = -(-full_pairs(L, F0) // L)
    while True:
        found = None
        for R2 in range(1, R // 2 + 1):
            R1 = R - R2
            if R1 == R2:
                continue
            pieces = solve(L, F0, R1, R2)
            if pieces is None:
                continue
            key = (collisions(L, F0, R1, R2, pieces), -len(guards_covered(L, F0, pieces)))
            if found is None or key < found[0]:
                found = (key, R1, R2, pieces)
        if found:
            return found[1], found[2], found[3]
        R += 1


def lanes(L, F0, R1, R2, pieces):
    p1 = sorted((a, b) for (b, a, m) in pieces if m == R1)
    p2 = sorted((a, b) for (b, a, m) in pieces if m == R2)
    assert len(p1) == L and len(p2) == L
    covered = guards_covered(L, F0, pieces)
    corners = [(F0 - 1 - beta, beta) for beta in range(max(0, F0 - L), F0) if beta not in covered]
    assert len(corners) <= L
    out = []
    for j in range(L):
        ca, cb = corners[j] if j < len(corners) else (0, 0)
        has = 1 if j < len(corners) else 0
        out.append((p1[j][0], p1[j][1], p2[j][0], p2[j][1], ca, cb, has))
    return out, len(corners)


def check(L, F0, R1, R2, table):
    """every full pair exactly once, no guard pair twice, rows inside [0, L)"""
    seen = {}
    for (a1, b1, a2, b2, ca, cb, has) in table:
        for (a, b, m) in ((a1, b1, R1), (a2, b2, R2)):
            for u in range(m):
                assert 0 <= a + u < L and 0 <= b < L
                if a + u + b >= F0 - 1:
                    key = (a + u, b)
                    assert key not in seen, key
                    seen[key] = 1
        if has:
            assert ca + cb == F0 - 1 and (ca, cb) not in seen
            seen[(ca, cb)] = 1
    for a in range(L):
        for b in range(L):
            if a + b >= F0 - 1:
                assert (a, b) in seen, (L, a, b)


# (L, F0) pairs the compiled shapes use: F0 = L - 1, and for L = 32 min(31, TMAX // KD)
# with TMAX = (p - 133) // 27 (MulShort in fpa_ops.h)
PAIRS = ((4, 3), (8, 7), (16, 15), (32, 31), (32, 30), (32, 29), (32, 28), (32, 27), (32, 23))


def main():
    out = sys.stdout
    out.write("/* Generated by gen_mul_sched.py: lane schedules of the short product in mul.\n"
              " * Per lane one 32-bit word: bits 0-4 alpha of segment 1's first round, 5-9 its beta,\n"
              " * 10-14 / 15-19 the same for segment 2, 20-24 / 25-29 the corner's guard pair\n"
              " * (alpha, beta), bit 30 set when the lane has a corner.  MulSched<L, F0>: groups of L\n"
              " * lanes, full chunk pairs on anti-diagonals >= F0, guard pairs on F0 - 1. */\n")
    out.write("#ifndef FPA_MUL_SCHED_H\n#define FPA_MUL_SCHED_H\n\n#include <stdint.h>\n\nnamespace fpa {\n\n")
    out.write("template <int L, int F0> struct MulSched { static constexpr bool ok = false; };\n\n")
    for L, F0 in PAIRS:
        R1, R2, pieces = schedule(L, F0)
        table, ncorner = lanes(L, F0, R1, R2, pieces)
        check(L, F0, R1, R2, table)
        words = []
        for (a1, b1, a2, b2, ca, cb, has) in table:
            words.append(a1 | b1 << 5 | a2 << 10 | b2 << 15 | ca << 20 | cb << 25 | has << 30)
        nguard = sum(1 for b in range(max(0, F0 - L), F0))
        out.write(f"/* L = {L}, F0 = {F0}: {R1} + {R2} rounds per lane ({L * (R1 + R2)} slots for "
                  f"{full_pairs(L, F0)} full pairs), {nguard - ncorner} guard pairs whole, "
                  f"{ncorner} corners; {collisions(L, F0, R1, R2, pieces)} lane-chunk collisions over "
                  f"{R1 + R2 + 2} emissions */\n")
        out.write(f"__constant__ uint32_t k_mul_sched_{L}_{F0}[{L}] = {{\n")
        for i in range(0, L, 8):
            out.write("  " + ", ".join(f"0x{w:08x}u" for w in words[i:i + 8]) + ",\n")
        out.write("};\n")
        ok = "true" if R1 + R2 < L else "false"
        out.write(f"template <> struct MulSched<{L}, {F0}> {{\n"
                  f"  static constexpr bool ok = {ok};\n"
                  f"  static constexpr int R1 = {R1}, R2 = {R2}, NCORNER = {ncorner};\n"
                  f"  __device__ static uint32_t lane(int gl) {{ return k_mul_sched_{L}_{F0}[gl]; }}\n"
                  f"}};\n\n")
        print(f"L={L} F0={F0}: R1={R1} R2={R2} corners={ncorner}", file=sys.stderr, flush=True)
    out.write("}  // namespace fpa\n\n#endif\n")


if __name__ == "__main__":
    main()
