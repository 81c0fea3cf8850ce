/*
 * Wave-group multi-precision integer primitives for CDNA4 (gfx950, wave64).
 *
 * A value's n = K*L 32-bit limbs are spread over a group of L consecutive lanes of
 * one wavefront; group lane j holds limbs [j*K, j*K+K) in registers (natural order,
 * so a lane's chunk is one contiguous K*4-byte load).  64/L values share a wave.
 * This replaces the reference's 32-lane warp layout with `size` limbs per lane
 * (int_warp_math.h:57-90): L and K are template parameters chosen per precision.
 *
 * Cross-lane traffic:
 *   - operand/neighbour exchange: ds_bpermute_b32 (any group size);
 *   - carry/borrow resolution: two 64-bit ballots per add, resolved with one scalar
 *     add of the generate/propagate masks (the reference's warp_int_fast_propagate_add,
 *     int_warp_math.h:402-469, does the same on 32-bit ballots);
 *   - shifts by arbitrary bit counts: LDS staging, each lane reads the words it needs.
 * All branch conditions in callers are uniform within a group, so every lane a group
 * reads from is active.
 */
#ifndef FPA_GROUP_H
#define FPA_GROUP_H

#include <hip/hip_runtime.h>
#include <stdint.h>
#include "fpa_scalar.h"

namespace fpa {

/* Orders LDS accesses of one wavefront: every lane's prior LDS writes are complete and
 * visible before any lane's later reads.  Groups never span waves, so no workgroup
 * barrier is needed. */
__device__ __forceinline__ void wsync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

/* f(0), f(1), ..., f(N-1) as straight-line code (no loop for the unroller to decline) */
template <int N, int I = 0, typename F>
__device__ __forceinline__ void repeat(F &&f) {
  if constexpr (I < N) {
    f(I);
    repeat<N, I + 1>(f);
  }
}

template <int K>
__device__ __forceinline__ uint32_t or_all(const uint32_t (&x)[K]) {
  uint32_t r = 0;
#pragma unroll
  for (int i = 0; i < K; i++) r |= x[i];
  return r;
}

template <int K>
__device__ __forceinline__ uint32_t and_all(const uint32_t (&x)[K]) {
  uint32_t r = 0xFFFFFFFFu;
#pragma unroll
  for (int i = 0; i < K; i++) r &= x[i];
  return r;
}

/* A zero in a VGPR that the compiler cannot fold: carry chains with a zero operand then
 * stay one v_addc_co / v_subb_co per limb instead of compare-and-select sequences. */
__device__ __forceinline__ uint32_t opaque_zero() {
  uint32_t z;
  asm("v_mov_b32 %0, 0" : "=v"(z));
  return z;
}

/* Hides the chunk's words from the compiler after a carry chain whose carry out is
 * dropped.  Without it the compiler can fold that chain's last carry into the next chain
 * on the same words (x1 = (a + c1) mod 2^32, then x1 + 0 + c becomes one a + c1 + c): the
 * sum is right but that add's carry out also counts the dropped carry, and a chunk whose
 * words were all ones then reports a carry it did not make.  (Seen on gfx950: an add-back
 * after a resolved carry in sqrt.)  Emits no instructions. */
template <int K>
__device__ __forceinline__ void opaque_chunk(uint32_t (&x)[K]) {
#pragma unroll
  for (int i = 0; i < K; i++) asm volatile("" : "+v"(x[i]));
}

/* r = a + b + cin (cin 0 or 1), returns the carry out of the chunk */
template <int K>
__device__ __forceinline__ uint32_t chunk_add(uint32_t (&r)[K], const uint32_t (&a)[K], const uint32_t (&b)[K],
                                              uint32_t cin) {
  unsigned int c = cin;
#pragma unroll
  for (int i = 0; i < K; i++) {
    unsigned int o;
    r[i] = __builtin_addc(a[i], b[i], c, &o);
    c = o;
  }
  return c;
}

/* r = a - b - bin (bin 0 or 1), returns the borrow out of the chunk */
template <int K>
__device__ __forceinline__ uint32_t chunk_sub(uint32_t (&r)[K], const uint32_t (&a)[K], const uint32_t (&b)[K],
                                              uint32_t bin) {
  unsigned int c = bin;
#pragma unroll
  for (int i = 0; i < K; i++) {
    unsigned int o;
    r[i] = __builtin_subc(a[i], b[i], c, &o);
    c = o;
  }
  return c;
}

/* x += w at word 0 (any 32-bit w), returns the carry out */
template <int K>
__device__ __forceinline__ uint32_t chunk_add_word(uint32_t (&x)[K], uint32_t w) {
  unsigned int c, o;
  x[0] = __builtin_addc(x[0], w, 0u, &o);
  c = o;
  if constexpr (K > 1) {
    const uint32_t z = opaque_zero();
#pragma unroll
    for (int i = 1; i < K; i++) {
      x[i] = __builtin_addc(x[i], z, c, &o);
      c = o;
    }
  }
  return c;
}

/* x -= w at word 0 (any 32-bit w), returns the borrow out */
template <int K>
__device__ __forceinline__ uint32_t chunk_sub_word(uint32_t (&x)[K], uint32_t w) {
  unsigned int c, o;
  x[0] = __builtin_subc(x[0], w, 0u, &o);
  c = o;
  if constexpr (K > 1) {
    const uint32_t z = opaque_zero();
#pragma unroll
    for (int i = 1; i < K; i++) {
      x[i] = __builtin_subc(x[i], z, c, &o);
      c = o;
    }
  }
  return c;
}

/* x += w at word 0, the carry taken through limbs 0..H-1 only; returns the carry out of
 * limb H-1.  After adding a full word to limb 0, a carry leaves limb 1 only if limb 1 was
 * all ones, so a two-limb head almost always settles the sum; chunk_carry_tail finishes. */
template <int K, int H>
__device__ __forceinline__ uint32_t chunk_add_word_head(uint32_t (&x)[K], uint32_t w) {
  static_assert(H >= 1 && H <= K, "head within the chunk");
  unsigned int c, o;
  x[0] = __builtin_addc(x[0], w, 0u, &o);
  c = o;
  if constexpr (H > 1) {
    const uint32_t z = opaque_zero();
#pragma unroll
    for (int i = 1; i < H; i++) {
      x[i] = __builtin_addc(x[i], z, c, &o);
      c = o;
    }
  }
  return c;
}

/* x -= w at word 0, the borrow taken through limbs 0..H-1 only; returns the borrow out of
 * limb H-1 (see chunk_add_word_head) */
template <int K, int H>
__device__ __forceinline__ uint32_t chunk_sub_word_head(uint32_t (&x)[K], uint32_t w) {
  static_assert(H >= 1 && H <= K, "head within the chunk");
  unsigned int c, o;
  x[0] = __builtin_subc(x[0], w, 0u, &o);
  c = o;
  if constexpr (H > 1) {
    const uint32_t z = opaque_zero();
#pragma unroll
    for (int i = 1; i < H; i++) {
      x[i] = __builtin_subc(x[i], z, c, &o);
      c = o;
    }
  }
  return c;
}

/* carry (or, with Sub, borrow) c into limb H, through limbs H..K-1; returns the one out */
template <int K, int H, bool Sub>
__device__ __forceinline__ uint32_t chunk_carry_tail(uint32_t (&x)[K], uint32_t c) {
  if constexpr (H < K) {
    const uint32_t z = opaque_zero();
    unsigned int o;
#pragma unroll
    for (int i = H; i < K; i++) {
      x[i] = Sub ? __builtin_subc(x[i], z, c, &o) : __builtin_addc(x[i], z, c, &o);
      c = o;
    }
  }
  return c;
}

/* x += w / x -= w at word 0 in every lane of the wave: the head, then the tail only where
 * some lane of the wave carries out of the head; returns this lane's carry/borrow out */
template <int K, bool Sub>
__device__ __forceinline__ uint32_t chunk_word_wave(uint32_t (&x)[K], uint32_t w) {
  constexpr int H = K < 2 ? K : 2;
  uint32_t c = Sub ? chunk_sub_word_head<K, H>(x, w) : chunk_add_word_head<K, H>(x, w);
  if constexpr (H < K) {
    if (__builtin_expect(__ballot(c != 0) != 0, 0)) c = chunk_carry_tail<K, H, Sub>(x, c);
  }
  return c;
}

template <int K>
__device__ __forceinline__ void chunk_select(uint32_t (&r)[K], bool c, const uint32_t (&a)[K], const uint32_t (&b)[K]) {
#pragma unroll
  for (int i = 0; i < K; i++) r[i] = c ? a[i] : b[i];
}

/* Coalesced chunk load/store: lane's K limbs are contiguous (p is 4*K-byte aligned). */
template <int K>
__device__ __forceinline__ void load_chunk(uint32_t (&x)[K], const uint32_t *p) {
  if constexpr (K % 4 == 0) {
    const uint4 *q = reinterpret_cast<const uint4 *>(p);
#pragma unroll
    for (int i = 0; i < K / 4; i++) {
      uint4 t = q[i];
      x[4 * i] = t.x; x[4 * i + 1] = t.y; x[4 * i + 2] = t.z; x[4 * i + 3] = t.w;
    }
  } else if constexpr (K % 2 == 0) {
    const uint2 *q = reinterpret_cast<const uint2 *>(p);
#pragma unroll
    for (int i = 0; i < K / 2; i++) {
      uint2 t = q[i];
      x[2 * i] = t.x; x[2 * i + 1] = t.y;
    }
  } else {
#pragma unroll
    for (int i = 0; i < K; i++) x[i] = p[i];
  }
}

template <int K>
__device__ __forceinline__ void store_chunk(uint32_t *p, const uint32_t (&x)[K]) {
  if constexpr (K % 4 == 0) {
    uint4 *q = reinterpret_cast<uint4 *>(p);
#pragma unroll
    for (int i = 0; i < K / 4; i++) q[i] = make_uint4(x[4 * i], x[4 * i + 1], x[4 * i + 2], x[4 * i + 3]);
  } else if constexpr (K % 2 == 0) {
    uint2 *q = reinterpret_cast<uint2 *>(p);
#pragma unroll
    for (int i = 0; i < K / 2; i++) q[i] = make_uint2(x[2 * i], x[2 * i + 1]);
  } else {
#pragma unroll
    for (int i = 0; i < K; i++) p[i] = x[i];
  }
}

/* LDS chunk store/load (p = group's LDS base + lane offset) */
template <int K>
__device__ __forceinline__ void lds_store(uint32_t *p, const uint32_t (&x)[K]) {
#pragma unroll
  for (int i = 0; i < K; i++) p[i] = x[i];
}

template <int K>
__device__ __forceinline__ void lds_load(uint32_t (&x)[K], const uint32_t *p) {
#pragma unroll
  for (int i = 0; i < K; i++) x[i] = p[i];
}

/* 32 bits of the nw-word number in LDS starting at bit offset pos (may be negative or
 * beyond the top; missing bits read as zero). */
__device__ __forceinline__ uint32_t lds_bits(const uint32_t *x, int nw, int pos) {
  int w = pos >> 5;          /* floor division (arithmetic shift) */
  int b = pos & 31;
  uint32_t lo = (w >= 0 && w < nw) ? x[w] : 0u;
  uint32_t hi = (w + 1 >= 0 && w + 1 < nw) ? x[w + 1] : 0u;
  return b ? ((lo >> b) | (hi << (32 - b))) : lo;
}

template <int L>
struct Group {
  static_assert(L >= 1 && L <= 64 && (64 % L) == 0, "group size must divide the wave");
  int lane;   /* 0..63 */
  int gl;     /* lane within the group */
  int gbase;  /* first lane of the group */

  __device__ __forceinline__ Group() {
    lane = (int)__lane_id();
    gl = lane & (L - 1);
    gbase = lane - gl;
  }

  static constexpr uint64_t mask() { return L == 64 ? ~0ull : ((1ull << L) - 1); }

  /* x as held by group lane src (uniform or per-lane src) */
  __device__ __forceinline__ uint32_t bcast(uint32_t x, int src) const {
    if constexpr (L == 1) return x;
    return (uint32_t)__builtin_amdgcn_ds_bpermute((gbase + src) << 2, (int)x);
  }

  /* x as held by the group's top lane, through DPP moves instead of a ds_bpermute round
   * trip: it sits on the serial path of every div and sqrt digit step.
   *   L = 2, 4:  one quad_perm ([1,1,3,3] or [3,3,3,3]).
   *   L = 16:    one row_newbcast:15 (lane 15 of each 16-lane row to the whole row).
   *   L = 8:     row_newbcast:15 to the whole row, then row_newbcast:7 written only to
   *              banks 0 and 1 (row lanes 0-7; the other lanes keep the first move's value).
   *   L = 32:    ds_bpermute (two readlanes and a select measured 4-5% slower in sqrt). */
  __device__ __forceinline__ uint32_t bcast_top(uint32_t x) const {
    if constexpr (L == 1) return x;
    else if constexpr (L == 2) return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0xF5, 0xF, 0xF, false);
    else if constexpr (L == 4) return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0xFF, 0xF, 0xF, false);
    else if constexpr (L == 8) {
      const int t = __builtin_amdgcn_mov_dpp((int)x, 0x15F, 0xF, 0xF, false);
      return (uint32_t)__builtin_amdgcn_update_dpp(t, (int)x, 0x157, 0xF, 0x3, false);
    } else if constexpr (L == 16) {
      return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x15F, 0xF, 0xF, false);
    } else return bcast(x, L - 1);
  }

  /* x as held by group lane J (a constant) through DPP moves for groups inside one row:
   * quad_perm for 2 and 4 lanes, row_newbcast:J for 16, and for 8 row_newbcast:8+J to the
   * whole row then row_newbcast:J written to banks 0 and 1 (row lanes 0-7); ds_bpermute for
   * 32 and 64 */
  template <int J>
  __device__ __forceinline__ uint32_t bcast_lane(uint32_t x) const {
    static_assert(J >= 0 && J < L, "a lane of the group");
    if constexpr (L == 1) return x;
    else if constexpr (L == 2) return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, J | J << 2 | (2 + J) << 4 | (2 + J) << 6, 0xF, 0xF, false);
    else if constexpr (L == 4) return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, J * 0x55, 0xF, 0xF, false);
    else if constexpr (L == 8) {
      const int t = __builtin_amdgcn_mov_dpp((int)x, 0x158 + J, 0xF, 0xF, false);
      return (uint32_t)__builtin_amdgcn_update_dpp(t, (int)x, 0x150 + J, 0xF, 0x3, false);
    } else if constexpr (L == 16) {
      return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x150 + J, 0xF, 0xF, false);
    } else return bcast(x, J);
  }

  /* from_prev through a DPP move for the digit steps' serial path, instead of a ds_bpermute
   * round trip: a row shift (row_shr:1) for groups inside one 16-lane row, a wave shift
   * (wave_shr:1) for groups of 32.  Lane 0 of a group takes 'fill' (the shift gave it the
   * previous group's last lane, or zero at the row's or wave's start). */
  __device__ __forceinline__ uint32_t from_prev_row(uint32_t x, uint32_t fill = 0) const {
    if constexpr (L >= 2 && L <= 16) {
      const uint32_t y = (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x111, 0xF, 0xF, true);
      return gl == 0 ? fill : y;
    } else if constexpr (L == 32) {
      /* wave_shr:1 moves every lane's value one lane up across row boundaries */
      const uint32_t y = (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x138, 0xF, 0xF, true);
      return gl == 0 ? fill : y;
    } else {
      return from_prev(x, fill);
    }
  }

  /* from_prev_row with a zero fill: a 16-lane group is a whole DPP row, whose row_shr:1
   * already gives its lane 0 a zero (bound_ctrl), so no select is needed there */
  __device__ __forceinline__ uint32_t from_prev_row0(uint32_t x) const {
    if constexpr (L == 16) return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x111, 0xF, 0xF, true);
    else return from_prev_row(x, 0u);
  }

  /* from_prev_row0 with the zero in group lane 0 from a lane mask nz0 (all ones except in
   * group lane 0) that the caller keeps opaque: the AND takes the DPP move as its first
   * source (one v_and_b32 with a DPP operand instead of a move and a select) */
  __device__ __forceinline__ uint32_t from_prev_row_and(uint32_t x, uint32_t nz0) const {
    if constexpr (L == 16) return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x111, 0xF, 0xF, true);
    else if constexpr (L >= 2 && L <= 8) return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x111, 0xF, 0xF, true) & nz0;
    else if constexpr (L == 32) return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x138, 0xF, 0xF, true) & nz0;
    else return from_prev(x, 0u);
  }

  /* x of group lane gl+1, 'fill' in the group's top lane: DPP row_shl:1 for groups inside
   * one 16-lane row (a group's top lane reads the next group's lane 0, or zero at the row's
   * end, and takes 'fill' instead), ds_bpermute for larger groups */
  __device__ __forceinline__ uint32_t from_next_row(uint32_t x, uint32_t fill = 0) const {
    if constexpr (L >= 2 && L <= 16) {
      const uint32_t y = (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x101, 0xF, 0xF, true);
      return gl == L - 1 ? fill : y;
    } else {
      return from_next(x, fill);
    }
  }

  /* from_next_row with a zero fill: a 16-lane group's top lane is the row's lane 15, which
   * row_shl:1 already gives a zero (bound_ctrl), so no select is needed there */
  __device__ __forceinline__ uint32_t from_next_row0(uint32_t x) const {
    if constexpr (L == 16) return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x101, 0xF, 0xF, true);
    else return from_next_row(x, 0u);
  }

  /* x of group lane gl-1; 'fill' in group lane 0 */
  __device__ __forceinline__ uint32_t from_prev(uint32_t x, uint32_t fill = 0) const {
    if constexpr (L == 1) return fill;
    uint32_t y = (uint32_t)__builtin_amdgcn_ds_bpermute((lane - 1) << 2, (int)x);
    return gl == 0 ? fill : y;
  }

  /* x of group lane gl+1; 'fill' in the group's top lane */
  __device__ __forceinline__ uint32_t from_next(uint32_t x, uint32_t fill = 0) const {
    if constexpr (L == 1) return fill;
    uint32_t y = (uint32_t)__builtin_amdgcn_ds_bpermute((lane + 1) << 2, (int)x);
    return gl == L - 1 ? fill : y;
  }

  /* ballot restricted to the group, bit j = group lane j */
  __device__ __forceinline__ uint64_t ballot(bool p) const {
    uint64_t b = __ballot(p);
    if constexpr (L == 64) return b;
    return (b >> gbase) & mask();
  }

  __device__ __forceinline__ bool any(bool p) const { return ballot(p) != 0; }

  /*
   * Carry-lookahead over the group: lane j generated a carry (g) or passes one through
   * (p: its chunk is all ones after the local add).  Returns the carry into this lane;
   * cout = carry out of the top lane.  The carry into lane j is bit j of
   * (G|P) + G + cin, xor (G|P) xor G.
   */
  __device__ __forceinline__ uint32_t resolve(bool g, bool p, uint32_t cin, uint32_t &cout) const {
    uint64_t G = ballot(g), P = ballot(p);
    uint64_t X = G | P;
    if constexpr (L < 64) {
      uint64_t S = X + G + cin;
      cout = (uint32_t)((S >> L) & 1u);
      return (uint32_t)(((S ^ X ^ G) >> gl) & 1u);
    } else {
      uint64_t S1 = X + G;
      uint32_t c1 = S1 < X;
      uint64_t S = S1 + cin;
      uint32_t c2 = S < S1;
      cout = c1 | c2;
      return (uint32_t)(((S ^ X ^ G) >> gl) & 1u);
    }
  }

  /* r = a + b (+ cin into the lowest word of the value); returns the carry out */
  template <int K>
  __device__ __forceinline__ uint32_t add(uint32_t (&r)[K], const uint32_t (&a)[K], const uint32_t (&b)[K],
                                          uint32_t cin = 0) const {
    uint32_t c = chunk_add(r, a, b, 0u);
    uint32_t cout;
    uint32_t ci = resolve(c != 0, and_all(r) == 0xFFFFFFFFu, cin, cout);
    chunk_word_wave<K, false>(r, ci);
    opaque_chunk(r);
    return cout;
  }

  /* r = a - b (- bin at the lowest word); returns the borrow out */
  template <int K>
  __device__ __forceinline__ uint32_t sub(uint32_t (&r)[K], const uint32_t (&a)[K], const uint32_t (&b)[K],
                                          uint32_t bin = 0) const {
    uint32_t c = chunk_sub(r, a, b, 0u);
    uint32_t bout;
    uint32_t bi = resolve(c != 0, or_all(r) == 0u, bin, bout);
    chunk_word_wave<K, true>(r, bi);
    opaque_chunk(r);
    return bout;
  }

  /* x += w_lane * 2^(32*K*gl) for a per-lane word w (e.g. carries from the lane below);
   * returns the carry out of the value */
  template <int K>
  __device__ __forceinline__ uint32_t add_lane_words(uint32_t (&x)[K], uint32_t w) const {
    uint32_t c = chunk_word_wave<K, false>(x, w);
    /* A chunk passes a carry on only if adding w overflowed all its limbs: rare.  When no
     * lane of the wave did, no carry enters any lane and none leaves any value (the
     * lookahead below would give all zeros), so the resolution is skipped. */
    if (__builtin_expect(__ballot(c != 0) == 0, 1)) return 0u;
    uint32_t cout;
    uint32_t ci = resolve(c != 0, and_all(x) == 0xFFFFFFFFu, 0u, cout);
    if (__ballot(ci != 0) != 0) chunk_word_wave<K, false>(x, ci);
    opaque_chunk(x);
    return cout;
  }

  /* x -= w_lane * 2^(32*K*gl); returns the borrow out of the value */
  template <int K>
  __device__ __forceinline__ uint32_t sub_lane_words(uint32_t (&x)[K], uint32_t w) const {
    uint32_t c = chunk_word_wave<K, true>(x, w);
    /* a chunk passes a borrow on only if its limbs above word 0 were all zero: rare; with
     * no such lane in the wave nothing propagates (see add_lane_words) */
    if (__builtin_expect(__ballot(c != 0) == 0, 1)) return 0u;
    uint32_t bout;
    uint32_t bi = resolve(c != 0, or_all(x) == 0u, 0u, bout);
    if (__ballot(bi != 0) != 0) chunk_word_wave<K, true>(x, bi);
    opaque_chunk(x);
    return bout;
  }

  /* limb 'pos' (compile-time) of the group's value */
  template <int K, int POS>
  __device__ __forceinline__ uint32_t word(const uint32_t (&x)[K]) const {
    return bcast_lane<POS / K>(x[POS % K]);
  }

  /* number of leading zero bits of the n = K*L word value (32n if zero) */
  template <int K>
  __device__ __forceinline__ int clz(const uint32_t (&x)[K]) const {
    int lz = 0;
    bool found = false;
#pragma unroll
    for (int i = K - 1; i >= 0; i--) {
      if (!found) {
        if (x[i]) { lz += __clz(x[i]); found = true; }
        else lz += 32;
      }
    }
    uint64_t nz = ballot(found);
    if (nz == 0) return 32 * K * L;
    int hl = 63 - __clzll(nz);                 /* highest group lane holding a 1 bit */
    int lzh = (int)bcast((uint32_t)lz, hl);
    return (L - 1 - hl) * 32 * K + lzh;
  }
};

}  // namespace fpa

#endif
