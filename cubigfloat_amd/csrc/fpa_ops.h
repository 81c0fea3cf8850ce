/*
 * Element-wise correctly rounded operations on the fpa format, one value per group of
 * L lanes, K limbs per lane (n = K*L limbs, p = 32n bits).
 *
 * Per-op structure (reference -> here):
 *   add/sub  fpa_warp_add_sub.h:41-380 -> op_addsub: align by an LDS funnel shift with a
 *            round word + sticky below the mantissa, group add/sub with ballot carry
 *            resolution, renormalise (1 bit, or clz + LDS shift on cancellation), round.
 *   mul      fpa_warp_mul.h:14-139 -> op_mul: operands recoded to radix 2^27 so every
 *            digit product accumulates into a 64-bit column sum with one v_mad_u64_u32;
 *            rounds broadcast a's digit chunk r, windows of column sums shift down one
 *            lane per round (group lane 0's low window is complete), carries are
 *            normalised once at the end.
 *   div      fpa_warp_div.h:127-201 -> op_div: long division in radix 2^27 with the
 *            remainder kept as lazy signed 64-bit column sums (one multiply-add per column
 *            and digit), FP64 digit estimates, redundant digits; the second half of the
 *            steps on the top slots only (short division), with the whole frame and its
 *            exact end correction as the fallback where the end's estimate is unsure.
 *   sqrt     fpa_warp_sqrt.h:195-254 -> op_sqrt: digit-by-digit square root in radix
 *            2^27 on div's lazy column frame (core_sqrt_lazy), all steps on the top half
 *            of the frame (short root), the half-then-whole frame as the fallback.
 * Rounding uses fpa_round's decision (fpa_scalar.h) on the exact (round word, sticky)
 * pair; a carry out of the rounding increment becomes 0x80..0 one binade up.
 */
#ifndef FPA_OPS_H
#define FPA_OPS_H

#include <type_traits>

#include "fpa_group.h"
#include "fpa_mul_sched.h"
#include "../../include/fpa_kernels.h"

namespace fpa {

/* ---- action tables: data restated from the reference (see fpa_oracle.c) ---- */
__constant__ uint8_t k_add_tab[81] = {
  FPA_A_ADD,    FPA_A_SUB,    FPA_A_COPY_A,  FPA_A_COPY_A,  FPA_A_COPY_A,  FPA_C_NEG_INF, FPA_C_POS_INF, FPA_C_NAN, FPA_C_SNAN,
  FPA_A_SUB,    FPA_A_ADD,    FPA_A_COPY_A,  FPA_A_COPY_A,  FPA_A_COPY_A,  FPA_C_NEG_INF, FPA_C_POS_INF, FPA_C_NAN, FPA_C_SNAN,
  FPA_A_COPY_B, FPA_A_COPY_B, FPA_C_ZERO,    FPA_C_NEG_TINY, FPA_C_POS_TINY, FPA_C_NEG_INF, FPA_C_POS_INF, FPA_C_NAN, FPA_C_SNAN,
  FPA_A_COPY_B, FPA_A_COPY_B, FPA_C_NEG_TINY, FPA_C_NEG_TINY, FPA_C_NAN,    FPA_C_NEG_INF, FPA_C_POS_INF, FPA_C_NAN, FPA_C_SNAN,
  FPA_A_COPY_B, FPA_A_COPY_B, FPA_C_POS_TINY, FPA_C_NAN,     FPA_C_POS_TINY, FPA_C_NEG_INF, FPA_C_POS_INF, FPA_C_NAN, FPA_C_SNAN,
  FPA_C_NEG_INF, FPA_C_NEG_INF, FPA_C_NEG_INF, FPA_C_NEG_INF, FPA_C_NEG_INF, FPA_C_NEG_INF, FPA_C_NAN,     FPA_C_NAN, FPA_C_SNAN,
  FPA_C_POS_INF, FPA_C_POS_INF, FPA_C_POS_INF, FPA_C_POS_INF, FPA_C_POS_INF, FPA_C_NAN,     FPA_C_POS_INF, FPA_C_NAN, FPA_C_SNAN,
  FPA_C_NAN,  FPA_C_NAN,  FPA_C_NAN,  FPA_C_NAN,  FPA_C_NAN,  FPA_C_NAN,  FPA_C_NAN,  FPA_C_NAN,  FPA_C_SNAN,
  FPA_C_SNAN, FPA_C_SNAN, FPA_C_SNAN, FPA_C_SNAN, FPA_C_SNAN, FPA_C_SNAN, FPA_C_SNAN, FPA_C_SNAN, FPA_C_SNAN,
};
__constant__ uint8_t k_sub_tab[81] = {
  FPA_A_SUB,   FPA_A_ADD,   FPA_A_COPY_A,  FPA_A_COPY_A,  FPA_A_COPY_A,  FPA_C_POS_INF, FPA_C_NEG_INF, FPA_C_NAN, FPA_C_SNAN,
  FPA_A_ADD,   FPA_A_SUB,   FPA_A_COPY_A,  FPA_A_COPY_A,  FPA_A_COPY_A,  FPA_C_POS_INF, FPA_C_NEG_INF, FPA_C_NAN, FPA_C_SNAN,
  FPA_A_NEG_B, FPA_A_NEG_B, FPA_C_ZERO,    FPA_C_POS_TINY, FPA_C_NEG_TINY, FPA_C_POS_INF, FPA_C_NEG_INF, FPA_C_NAN, FPA_C_SNAN,
  FPA_A_NEG_B, FPA_A_NEG_B, FPA_C_NEG_TINY, FPA_C_NAN,    FPA_C_NEG_INF, FPA_C_POS_INF, FPA_C_NEG_INF, FPA_C_NAN, FPA_C_SNAN,
  FPA_A_NEG_B, FPA_A_NEG_B, FPA_C_POS_TINY, FPA_C_POS_TINY, FPA_C_NAN,   FPA_C_POS_INF, FPA_C_NEG_INF, FPA_C_NAN, FPA_C_SNAN,
  FPA_C_NEG_INF, FPA_C_NEG_INF, FPA_C_NEG_INF, FPA_C_NEG_INF, FPA_C_NEG_INF, FPA_C_NAN,  FPA_C_NEG_INF, FPA_C_NAN, FPA_C_SNAN,
  FPA_C_POS_INF, FPA_C_POS_INF, FPA_C_POS_INF, FPA_C_POS_INF, FPA_C_POS_INF, FPA_C_POS_INF, FPA_C_NAN,  FPA_C_NAN, FPA_C_SNAN,
  FPA_C_NAN,  FPA_C_NAN,  FPA_C_NAN,  FPA_C_NAN,  FPA_C_NAN,  FPA_C_NAN,  FPA_C_NAN,  FPA_C_NAN,  FPA_C_SNAN,
  FPA_C_SNAN, FPA_C_SNAN, FPA_C_SNAN, FPA_C_SNAN, FPA_C_SNAN, FPA_C_SNAN, FPA_C_SNAN, FPA_C_SNAN, FPA_C_SNAN,
};
__constant__ uint8_t k_mul_tab[81] = {
  FPA_A_MUL,      FPA_A_MUL,      FPA_C_ZERO, FPA_C_NEG_TINY, FPA_C_POS_TINY, FPA_C_NEG_INF, FPA_C_POS_INF, FPA_C_NAN, FPA_C_SNAN,
  FPA_A_MUL,      FPA_A_MUL,      FPA_C_ZERO, FPA_C_POS_TINY, FPA_C_NEG_TINY, FPA_C_POS_INF, FPA_C_NEG_INF, FPA_C_NAN, FPA_C_SNAN,
  FPA_C_ZERO,     FPA_C_ZERO,     FPA_C_ZERO, FPA_C_ZERO,     FPA_C_ZERO,     FPA_C_NAN,     FPA_C_NAN,     FPA_C_NAN, FPA_C_SNAN,
  FPA_C_NEG_TINY, FPA_C_POS_TINY, FPA_C_ZERO, FPA_C_POS_TINY, FPA_C_NEG_TINY, FPA_C_NAN,     FPA_C_NAN,     FPA_C_NAN, FPA_C_SNAN,
  FPA_C_POS_TINY, FPA_C_NEG_TINY, FPA_C_ZERO, FPA_C_NEG_TINY, FPA_C_POS_TINY, FPA_C_NAN,     FPA_C_NAN,     FPA_C_NAN, FPA_C_SNAN,
  FPA_C_NEG_INF,  FPA_C_POS_INF,  FPA_C_NAN,  FPA_C_NAN,      FPA_C_NAN,      FPA_C_POS_INF, FPA_C_NEG_INF, FPA_C_NAN, FPA_C_SNAN,
  FPA_C_POS_INF,  FPA_C_NEG_INF,  FPA_C_NAN,  FPA_C_NAN,      FPA_C_NAN,      FPA_C_NEG_INF, FPA_C_POS_INF, FPA_C_NAN, FPA_C_SNAN,
  FPA_C_NAN,  FPA_C_NAN,  FPA_C_NAN,  FPA_C_NAN,  FPA_C_NAN,  FPA_C_NAN,  FPA_C_NAN,  FPA_C_NAN,  FPA_C_SNAN,
  FPA_C_SNAN, FPA_C_SNAN, FPA_C_SNAN, FPA_C_SNAN, FPA_C_SNAN, FPA_C_SNAN, FPA_C_SNAN, FPA_C_SNAN, FPA_C_SNAN,
};
/* classes MPFR represents follow MPFR: x/0, inf/0 -> +-INF; inf/x signed by the
 * quotient; x/inf -> ZERO.  Entries with a TINY or SNAN operand are the reference's
 * (fpa_warp_div.h:1-12), among them value/TINY -> INF with the quotient's sign. */
__constant__ uint8_t k_div_tab[81] = {
  FPA_A_DIV,      FPA_A_DIV,      FPA_C_POS_INF, FPA_C_NEG_INF,  FPA_C_POS_INF,  FPA_C_ZERO,     FPA_C_ZERO,     FPA_C_NAN, FPA_C_SNAN,
  FPA_A_DIV,      FPA_A_DIV,      FPA_C_NEG_INF, FPA_C_POS_INF,  FPA_C_NEG_INF,  FPA_C_ZERO,     FPA_C_ZERO,     FPA_C_NAN, FPA_C_SNAN,
  FPA_C_ZERO,     FPA_C_ZERO,     FPA_C_NAN,     FPA_C_ZERO,     FPA_C_ZERO,     FPA_C_ZERO,     FPA_C_ZERO,     FPA_C_NAN, FPA_C_SNAN,
  FPA_C_NEG_TINY, FPA_C_POS_TINY, FPA_C_NAN,     FPA_C_NAN,      FPA_C_NAN,      FPA_C_POS_TINY, FPA_C_NEG_TINY, FPA_C_NAN, FPA_C_SNAN,
  FPA_C_POS_TINY, FPA_C_NEG_TINY, FPA_C_NAN,     FPA_C_NAN,      FPA_C_NAN,      FPA_C_NEG_TINY, FPA_C_POS_TINY, FPA_C_NAN, FPA_C_SNAN,
  FPA_C_NEG_INF,  FPA_C_POS_INF,  FPA_C_NEG_INF, FPA_C_POS_INF,  FPA_C_NEG_INF,  FPA_C_NAN,      FPA_C_NAN,      FPA_C_NAN, FPA_C_SNAN,
  FPA_C_POS_INF,  FPA_C_NEG_INF,  FPA_C_POS_INF, FPA_C_NEG_INF,  FPA_C_POS_INF,  FPA_C_NAN,      FPA_C_NAN,      FPA_C_NAN, FPA_C_SNAN,
  FPA_C_NAN,  FPA_C_NAN,  FPA_C_NAN,  FPA_C_NAN,  FPA_C_NAN,  FPA_C_NAN,  FPA_C_NAN,  FPA_C_NAN,  FPA_C_SNAN,
  FPA_C_SNAN, FPA_C_SNAN, FPA_C_SNAN, FPA_C_SNAN, FPA_C_SNAN, FPA_C_SNAN, FPA_C_SNAN, FPA_C_SNAN, FPA_C_SNAN,
};
__constant__ uint8_t k_sqrt_tab[9] = { FPA_A_SQRT, FPA_C_NAN, FPA_C_ZERO, FPA_C_NAN, FPA_C_POS_TINY, FPA_C_NAN,
                                       FPA_C_POS_INF, FPA_C_NAN, FPA_C_SNAN };
__constant__ uint8_t k_neg_tab[9] = { FPA_A_NEG, FPA_A_NEG, FPA_C_ZERO, FPA_C_POS_TINY, FPA_C_NEG_TINY, FPA_C_POS_INF,
                                      FPA_C_NEG_INF, FPA_C_NAN, FPA_C_SNAN };

/* exponent word for (biased E, sign) under a [lo, hi] window of E; class code outside */
__device__ __forceinline__ uint32_t pack_exp(int64_t E, uint32_t sign, int64_t lo, int64_t hi) {
  if (E < lo) return FPA_C_POS_TINY - sign;
  if (E > hi) return FPA_C_POS_INF - sign;
  return (uint32_t)(2 * E + sign);
}

/* Where a group's value lives: global index and LDS scratch. */
struct Slot {
  uint32_t v;       /* value index (clamped to count-1 for reads when invalid) */
  bool valid;
};

/* round the group's mantissa x by (round word, sticky); returns 1 if it carried into
 * a new binade (x is then 0x80..0) */
template <int K, int L>
__device__ __forceinline__ uint32_t round_mantissa(const Group<L> &g, uint32_t (&x)[K], int mode, uint32_t sign,
                                                   uint32_t rw, uint32_t sticky) {
  uint32_t lsb = g.template bcast_lane<0>(x[0]) & 1u;
  uint32_t inc = fpa_round_inc(mode, sign, lsb, rw, sticky);
  uint32_t cout = g.add_lane_words(x, g.gl == 0 ? inc : 0u);
  if (cout) {
#pragma unroll
    for (int i = 0; i < K; i++) x[i] = 0;
    if (g.gl == L - 1) x[K - 1] = 0x80000000u;
  }
  return cout;
}

template <int K>
__device__ __forceinline__ void store_result(const fpa_arguments &a, const Slot &s, int gl, uint32_t rexp,
                                             const uint32_t (&x)[K]) {
  if (!s.valid) return;
  uint32_t z[K];
#pragma unroll
  for (int i = 0; i < K; i++) z[i] = rexp > FPA_C_SNAN ? x[i] : 0u;
  const uint32_t n = a.precision >> 5;
  store_chunk(a.r_limbs + (size_t)s.v * n + (size_t)gl * K, z);
  if (gl == 0) a.r_exps[s.v] = rexp;
}

template <int K>
__device__ __forceinline__ void store_special(const fpa_arguments &a, const Slot &s, int gl, uint32_t code) {
  uint32_t z[K];
#pragma unroll
  for (int i = 0; i < K; i++) z[i] = 0;
  store_result(a, s, gl, code, z);
}

/* Loads this lane's chunk of value s.v of an operand array. */
template <int K, int L>
__device__ __forceinline__ void load_value(const Group<L> &g, const uint32_t *limbs, const Slot &s, uint32_t (&x)[K]) {
  load_chunk(x, limbs + (size_t)s.v * (K * L) + (size_t)g.gl * K);
}

/* ------------------------------- neg ------------------------------- */
template <int K, int L>
__device__ __forceinline__ void op_neg(const fpa_arguments &a, const Slot &s, const Group<L> &g) {
  uint32_t ea = a.a_exps[s.v];
  uint32_t act = k_neg_tab[fpa_classify(ea)];
  uint32_t x[K];
  load_value<K, L>(g, a.a_limbs, s, x);
  store_result(a, s, g.gl, act == FPA_A_NEG ? (ea ^ 1u) : act, x);
}

/* ----------------------------- add / sub ----------------------------- */
/* (re, R) = a +- b for the group's values (ea, A), (eb, B); lds: >= n+1 words */
template <int K, int L>
__device__ __forceinline__ void core_addsub(const Group<L> &g, uint32_t *lds, int mode, bool is_sub, uint32_t ea,
                            const uint32_t (&A)[K], uint32_t eb, const uint32_t (&B)[K], uint32_t &re,
                            uint32_t (&R)[K]) {
  constexpr int n = K * L;
  uint32_t cls = fpa_classify(ea) * 9 + fpa_classify(eb);
  uint32_t act = is_sub ? k_sub_tab[cls] : k_add_tab[cls];

  if (act != FPA_A_ADD && act != FPA_A_SUB) {
    if (act == FPA_A_COPY_A) { re = ea; chunk_select(R, true, A, A); }
    else if (act == FPA_A_COPY_B) { re = eb; chunk_select(R, true, B, B); }
    else if (act == FPA_A_NEG_B) { re = eb ^ 1u; chunk_select(R, true, B, B); }
    else {
      re = act;
#pragma unroll
      for (int i = 0; i < K; i++) R[i] = 0;
    }
    return;
  }
  if (is_sub) eb ^= 1u;

  const int64_t Ea = ea >> 1, Eb = eb >> 1;
  const bool swap = Ea < Eb;
  uint32_t X[K], Y[K];
  chunk_select(X, swap, B, A);
  chunk_select(Y, swap, A, B);
  const int64_t Ex = swap ? Eb : Ea;
  const uint32_t sx = swap ? (eb & 1u) : (ea & 1u);
  const int64_t d = swap ? Eb - Ea : Ea - Eb;
  uint32_t T[K];
  int64_t E;
  uint32_t sign;
  uint32_t rw = 0, sticky = 0;

  if (act == FPA_A_SUB && d == 0) {
    /* equal exponents: exact difference, no rounding */
    uint32_t borrow = g.sub(T, X, Y);
    sign = sx;
    if (borrow) {
      /* |Y| > |X|: negate (two's complement) */
#pragma unroll
      for (int i = 0; i < K; i++) T[i] = ~T[i];
      g.add_lane_words(T, g.gl == 0 ? 1u : 0u);
      sign = (swap ? ea : eb) & 1u;
    }
    int sh = g.clz(T);
    if (sh == 32 * n) {
      re = FPA_C_ZERO;
#pragma unroll
      for (int i = 0; i < K; i++) R[i] = 0;
      return;
    }
    if (sh) {
      lds_store(lds + g.gl * K, T);
      wsync();
#pragma unroll
      for (int i = 0; i < K; i++) T[i] = lds_bits(lds, n, 32 * (g.gl * K + i) - sh);
      wsync();
    }
    E = Ex - sh;
    re = pack_exp(E, sign, 5, 0x7FFFFFFFll);
    chunk_select(R, true, T, T);
    return;
  }

  /* align Y: Ys = Y >> dd, round word = bits [dd-32, dd), sticky = any bit below */
  const int dd = d > 32 * n + 32 ? 32 * n + 32 : (int)d;
  uint32_t Ys[K];
  lds_store(lds + g.gl * K, Y);
  wsync();
#pragma unroll
  for (int i = 0; i < K; i++) Ys[i] = lds_bits(lds, n, 32 * (g.gl * K + i) + dd);
  rw = lds_bits(lds, n, dd - 32);
  wsync();
  {
    const int lim = dd - 32;   /* sticky covers bits [0, lim) */
    uint32_t loc = 0;
#pragma unroll
    for (int i = 0; i < K; i++) {
      int lo_bit = 32 * (g.gl * K + i);
      if (lo_bit + 32 <= lim) loc |= Y[i];
      else if (lo_bit < lim) loc |= Y[i] & ((1u << (lim - lo_bit)) - 1u);
    }
    sticky = g.any(loc != 0) ? 1u : 0u;
  }
  sign = sx;

  if (act == FPA_A_ADD) {
    uint32_t c = g.add(T, X, Ys);
    E = Ex;
    if (c) {
      /* shift right one bit, the carry becomes the top bit */
      uint32_t low = g.template bcast_lane<0>(T[0]) & 1u;
      uint32_t nxt = g.from_next_row(T[0], c);
#pragma unroll
      for (int i = 0; i < K - 1; i++) T[i] = (T[i] >> 1) | (T[i + 1] << 31);
      T[K - 1] = (T[K - 1] >> 1) | (nxt << 31);
      sticky |= rw & 1u;
      rw = (rw >> 1) | (low << 31);
      E = Ex + 1;
    }
  } else {
    /* [X,0,0] - [Ys, rw, sticky] */
    uint32_t b2 = (rw != 0 || sticky) ? 1u : 0u;
    rw = 0u - rw - sticky;
    g.sub(T, X, Ys, b2);
    uint32_t top = g.bcast_top(T[K - 1]);
    E = Ex;
    if (!(top >> 31)) {
      /* normalise the (n+1)-word value [T, rw]; more than one bit only when d == 1 (exact) */
      lds_store(lds + 1 + g.gl * K, T);
      if (g.gl == 0) lds[0] = rw;
      wsync();
      int lz;
      {
        int tz = g.clz(T);
        lz = tz < 32 * n ? tz : 32 * n + __clz(rw);
      }
#pragma unroll
      for (int i = 0; i < K; i++) T[i] = lds_bits(lds, n + 1, 32 * (g.gl * K + i + 1) - lz);
      rw = lds_bits(lds, n + 1, -lz);
      wsync();
      E = Ex - lz;
    }
  }
  if (round_mantissa<K, L>(g, T, mode, sign, rw, sticky)) E += 1;
  re = pack_exp(E, sign, 5, 0x7FFFFFFFll);
  chunk_select(R, true, T, T);
}

template <int K, int L>
__device__ __forceinline__ void op_addsub(const fpa_arguments &a, const Slot &s, const Group<L> &g, uint32_t *lds, bool is_sub) {
  uint32_t A[K], B[K], R[K], re;
  load_value<K, L>(g, a.a_limbs, s, A);
  load_value<K, L>(g, a.b_limbs, s, B);
  core_addsub<K, L>(g, lds, (int)a.mode, is_sub, a.a_exps[s.v], A, a.b_exps[s.v], B, re, R);
  store_result(a, s, g.gl, re, R);
}

/* -------------------------------- mul -------------------------------- */
/*
 * The product is formed in radix 2^27: each lane recodes its 32K mantissa bits into
 * KD = ceil(32K/27) digits, so a digit product is < 2^54 and a 64-bit column
 * accumulator takes every product of a column (at most 608 at 16384 bits) without
 * overflow.  The inner loop is then one v_mad_u64_u32 per digit product, with carries
 * propagated once at the end.
 *
 * Rounds r = 0..L-1: every lane multiplies digit chunk r of a (broadcast from group lane
 * r) by its own chunk of b into a window of 2*KD column sums; group lane 0's low KD
 * columns are complete and go to LDS; windows then shift down one lane.  After the last
 * round each lane holds KD high columns.  Columns are normalised to digits (local carry
 * chains, then the multi-bit carries between chunks, then a 1-bit ballot resolution).
 * The sticky bit comes from the low digits in registers.  The mantissa words and round
 * word are funnel-shifted out of each lane's high digits and a few digits of its
 * predecessor lane, in registers, where the shape allows (mul_extract_nb); otherwise the
 * digits go to LDS and the words are read back from there.
 */
constexpr int MUL_RADIX = 27;

__host__ __device__ constexpr int mul_digits_per_lane(int K) { return (32 * K + MUL_RADIX - 1) / MUL_RADIX; }
/* digit chunks in LDS are padded to whole 16-byte rows */
__host__ __device__ constexpr int mul_digits_padded(int K) { return (mul_digits_per_lane(K) + 3) / 4 * 4; }
/* groups of 2..16 lanes sit inside one 16-lane DPP row: their window shift is a DPP
 * row shift; groups of 32 shift through ds_bpermute */
__host__ __device__ constexpr bool mul_row_path(int L) { return L >= 2 && L <= 16; }
/* groups of 2 or more lanes take a's digit chunk of each round from LDS */
__host__ __device__ constexpr bool mul_lds_broadcast(int L) { return L >= 2; }
/* word offset of a's digit chunks: the first 16-byte row at or above ND = L*KD */
__host__ __device__ constexpr int mul_adig_offset(int K, int L) { return (L * mul_digits_per_lane(K) + 3) / 4 * 4; }
/* LDS words per value: 2*ND words of 64-bit low columns (later the digits).  a's digit
 * chunks overlap their top half (chunk r at mul_adig_offset + r*KDP): lane 0 fills the
 * low columns from the bottom, 2*KD words per round, and after round r has written
 * 2*(r+1)*KD <= ND + (r+1)*KDP words (KDP >= KD), so chunks still to be read are never
 * overwritten. */
__host__ __device__ constexpr int mul_lds_words(int K, int L) {
  return !mul_lds_broadcast(L) ? 2 * L * mul_digits_per_lane(K)
         : 2 * L * mul_digits_per_lane(K) > mul_adig_offset(K, L) + L * mul_digits_padded(K)
             ? 2 * L * mul_digits_per_lane(K)
             : mul_adig_offset(K, L) + L * mul_digits_padded(K);
}

/* 32K - 27*KD <= 0: how far a lane's K mantissa words run short of its KD digits */
__host__ __device__ constexpr int mul_delta(int K) { return 32 * K - MUL_RADIX * mul_digits_per_lane(K); }
/* Mantissa extraction in registers: lane gl's first mantissa bit lies
 * s = 27*NB + delta*(L + gl) - (1 - full) bits above the start of the string made of the
 * predecessor lane's NB top digits and its own KD high digits.  NB is the fewest digits
 * (at most KD) for which every lane's s >= 0, lane 0 has the 32-bit round word below its
 * first bit inside the string (s >= 32) and all s fit one two-word window from a
 * 32-bit-aligned offset; 0 when no NB qualifies (the digits then go through LDS). */
__host__ __device__ constexpr int mul_extract_nb(int K, int L) {
  const int d = mul_delta(K), KD = mul_digits_per_lane(K);
  for (int nb = 1; nb <= KD; nb++) {
    const int smin = MUL_RADIX * nb + d * (2 * L - 1) - 1, smax = MUL_RADIX * nb + d * L;
    if (smin >= 0 && MUL_RADIX * nb + d * L - 1 >= 32 && smax - smin / 32 * 32 < 64) return nb;
  }
  return 0;
}

/* x of lane+1 within this lane's 16-lane row (DPP row_shl:1); the row's lane 15 gets 0 */
__device__ __forceinline__ uint32_t dpp_row_next(uint32_t x) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x101, 0xF, 0xF, true);
}

/* own + (nb of lane+1 within this lane's 16-lane row; 0 in the row's lane 15), mod 2^64.
 * The compiler moves each DPP operand with a v_mov_b32_dpp of its own and does not fold it
 * into a carry add, so both halves are written here as VOP2 adds whose first source reads
 * the neighbour through DPP: two VALU instead of four.  The compiler counts no wait states
 * for asm, so the block keeps its own: 2 before its DPP reads (a VALU write of a source
 * just before), 5 when Nop says so (the first block after an EXEC write in the same
 * basic block), 2 between the VCC write of the low add and its read as the high add's
 * carry (the compiler keeps the same 2 between its own carry pairs on this target), and
 * 2 after it when Tail says so (a DPP read of its results just after). */
#define FPA_ADD_ROW_NEXT64(pre, post)                                                            \
  asm(pre "\n\t"                                                                                  \
      "v_add_co_u32_dpp %0, vcc, %2, %3 row_shl:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"    \
      "s_nop 1\n\t"                                                                              \
      "v_addc_co_u32_dpp %1, vcc, %4, %5, vcc row_shl:1 row_mask:0xf bank_mask:0xf bound_ctrl:1" \
      post                                                                                        \
      : "=&v"(lo), "=v"(hi)                                                                       \
      : "v"((uint32_t)nb), "v"((uint32_t)own), "v"((uint32_t)(nb >> 32)), "v"((uint32_t)(own >> 32)) \
      : "vcc")
template <int Nop, bool Tail>
__device__ __forceinline__ uint64_t add_row_next64(uint64_t own, uint64_t nb) {
  static_assert(Nop == 1 || Nop == 4, "s_nop count");
  uint32_t lo, hi;
  if constexpr (Nop == 4 && Tail) FPA_ADD_ROW_NEXT64("s_nop 4", "\n\ts_nop 1");
  else if constexpr (Nop == 4) FPA_ADD_ROW_NEXT64("s_nop 4", "");
  else if constexpr (Tail) FPA_ADD_ROW_NEXT64("s_nop 1", "\n\ts_nop 1");
  else FPA_ADD_ROW_NEXT64("s_nop 1", "");
  return ((uint64_t)hi << 32) | lo;
}
#undef FPA_ADD_ROW_NEXT64

/* 32 bits at bit offset o >= 0 of a digit string (radix 2^27) held in LDS; the two words
 * past its top digit must read as zero */
__device__ __forceinline__ uint32_t digits_bits(const uint32_t *d, int o) {
  const int k = o / MUL_RADIX, off = o - k * MUL_RADIX;
  uint64_t t = d[k] >> off;
  t |= (uint64_t)d[k + 1] << (MUL_RADIX - off);
  t |= (uint64_t)d[k + 2] << (2 * MUL_RADIX - off);
  return (uint32_t)t;
}

/* This lane's radix-2^27 digits of an n = K*L word value: digit k = bits
 * [27*(gl*KD + k), +27).  The value is staged in LDS with zero words past its top, so
 * each digit is one two-word read and a funnel shift. */
template <int K, int L, int KD>
__device__ __forceinline__ void recode_digits(const Group<L> &g, uint32_t *lds, const uint32_t (&X)[K],
                                              uint32_t (&D)[KD]) {
  constexpr int n = K * L;
  constexpr int PADW = (MUL_RADIX * KD * L + 31) / 32 + 1 - n;   /* words [n, n+PADW) read as 0 */
  constexpr uint32_t M = (1u << MUL_RADIX) - 1;
  static_assert(PADW >= 1, "digits cover the value");
  lds_store(lds + g.gl * K, X);
  if (g.gl == L - 1) {
#pragma unroll
    for (int i = 0; i < PADW; i++) lds[n + i] = 0;
  }
  wsync();
  /* the lane's digits are bits [b, b + 27*KD): funnel-shift NW words starting at bit b
   * out of LDS once, then every digit sits at a fixed offset of that window.  The highest
   * word read, floor(b/32) + NW, is at most ceil(27*KD*L/32) = n + PADW - 1. */
  constexpr int NW = (MUL_RADIX * KD + 31) / 32;
  const int b = MUL_RADIX * KD * g.gl;
  const uint32_t *src = lds + (b >> 5);
  const uint32_t s = (uint32_t)b & 31u;
  uint32_t win[NW + 1];
  uint32_t lo = src[0];
#pragma unroll
  for (int i = 0; i < NW; i++) {
    const uint32_t hi = src[i + 1];
    win[i] = __builtin_amdgcn_alignbit(hi, lo, s);
    lo = hi;
  }
  win[NW] = 0;
#pragma unroll
  for (int k = 0; k < KD; k++) {
    const int j = (MUL_RADIX * k) >> 5, sh = (MUL_RADIX * k) & 31;
    D[k] = (sh + MUL_RADIX <= 32 ? win[j] >> sh : __builtin_amdgcn_alignbit(win[j + 1], win[j], (uint32_t)sh)) & M;
  }
  wsync();
}

/* recode_digits without LDS, for shapes whose lanes' digit windows drift by less than a
 * word over the group: lane gl's digits are bits [27*KD*gl, +27*KD) = the lane's own words
 * from bit E*gl on (E = 27*KD - 32K) plus the next lane's first words, which come through
 * a DPP row shift (zero above the value).  16 limbs per lane: E = 1, two words of the next
 * lane. */
template <int K, int L, int KD>
__host__ __device__ constexpr bool recode_dpp_ok() {
  return L >= 2 && L <= 16 && (MUL_RADIX * KD - 32 * K) * (L - 1) < 32 &&
         (MUL_RADIX * KD + 31) / 32 + 1 - K <= K;
}
template <int K, int L, int KD>
__device__ __forceinline__ void recode_digits_dpp(const Group<L> &g, const uint32_t (&X)[K], uint32_t (&D)[KD]) {
  constexpr int E = MUL_RADIX * KD - 32 * K;
  constexpr int NW = (MUL_RADIX * KD + 31) / 32;   /* window words */
  constexpr int NX = NW + 1 - K;                   /* words taken from the next lane */
  constexpr uint32_t M = (1u << MUL_RADIX) - 1;
  static_assert(recode_dpp_ok<K, L, KD>(), "digit window within the lane and the next one");
  uint32_t w[K + NX];
#pragma unroll
  for (int i = 0; i < K; i++) w[i] = X[i];
#pragma unroll
  for (int j = 0; j < NX; j++) w[K + j] = g.from_next_row0(X[j]);
  const uint32_t s = (uint32_t)(E * g.gl);
  uint32_t win[NW + 1];
#pragma unroll
  for (int i = 0; i < NW; i++) win[i] = __builtin_amdgcn_alignbit(w[i + 1], w[i], s);
  win[NW] = 0;
#pragma unroll
  for (int k = 0; k < KD; k++) {
    const int j = (MUL_RADIX * k) >> 5, sh = (MUL_RADIX * k) & 31;
    D[k] = (sh + MUL_RADIX <= 32 ? win[j] >> sh : __builtin_amdgcn_alignbit(win[j + 1], win[j], (uint32_t)sh)) & M;
  }
}

/* local radix-2^27 normalisation of KD column sums with carry-in c; returns the carry out.
 * S > 0: the columns below S are zero and c is zero (their digits are zero). */
template <int KD, int S = 0>
__device__ __forceinline__ uint64_t normalize_columns(const uint64_t (&col)[KD], uint32_t (&dg)[KD], uint64_t c) {
  constexpr uint32_t M = (1u << MUL_RADIX) - 1;
#pragma unroll
  for (int k = 0; k < S; k++) dg[k] = 0;
#pragma unroll
  for (int k = S; k < KD; k++) {
    uint64_t v = col[k] + c;
    dg[k] = (uint32_t)v & M;
    c = v >> MUL_RADIX;
  }
  return c;
}

/* add a multi-bit carry c at digit 0 of a normalised chunk; returns the 1-bit overflow */
template <int KD>
__device__ __forceinline__ uint32_t add_carry_digits(uint32_t (&dg)[KD], uint64_t c) {
  constexpr uint32_t M = (1u << MUL_RADIX) - 1;
#pragma unroll
  for (int k = 0; k < KD; k++) {
    uint64_t v = (uint64_t)dg[k] + c;
    dg[k] = (uint32_t)v & M;
    c = v >> MUL_RADIX;
  }
  return (uint32_t)c;
}

/* add_carry_digits for a carry below 2^54, called by every lane of the wave together.
 * Digit 0 passes on at most 2^27 and digit 1 then at most 1, which moves on only past a
 * digit of all ones, so the rest of the chunk is visited only when some lane of the wave
 * still carries after two digits. */
template <int KD>
__device__ __forceinline__ uint32_t add_carry_digits_wave(uint32_t (&dg)[KD], uint64_t c) {
  constexpr uint32_t M = (1u << MUL_RADIX) - 1;
  constexpr int H = KD < 2 ? KD : 2;
#pragma unroll
  for (int k = 0; k < H; k++) {
    uint64_t v = (uint64_t)dg[k] + c;
    dg[k] = (uint32_t)v & M;
    c = v >> MUL_RADIX;
  }
  uint32_t c32 = (uint32_t)c;
  if constexpr (H < KD) {
    if (__builtin_expect(__ballot(c32 != 0) != 0, 0)) {
#pragma unroll
      for (int k = H; k < KD; k++) {
        const uint32_t v = dg[k] + c32;
        dg[k] = v & M;
        c32 = v >> MUL_RADIX;
      }
    }
  }
  return c32;
}

template <int K, int L, bool kShort, int LOWC0 = -1>
__device__ __forceinline__ void mul_finish(const Group<L> &g, uint32_t *lds,
                                           const uint64_t (&cl)[mul_digits_per_lane(K)],
                                           const uint64_t (&ch)[mul_digits_per_lane(K)], int mode, uint32_t ea,
                                           uint32_t eb, uint32_t &re, uint32_t (&R)[K], bool &amb);

/* (re, R) = a * b; lds: mul_lds_words(K, L) words (64-bit low columns, then digits; a's
 * digit chunks in their top half while the rounds run).  group_stride = LDS words from this group's
 * base to the next group's in the same wave.
 * kLockstep: every group of the wave calls this together (the element-wise kernel);
 * special-class operands then run the digit rounds too and take their class result at
 * the end, so the row path can let a group's top lane read its neighbour group and
 * correct for it once.  Without it (dot chains: groups run chains of different lengths)
 * the top lane's incoming value is masked each round instead. */
template <int K, int L, bool kLockstep>
__device__ __forceinline__ void core_mul(const Group<L> &g, uint32_t *lds, int group_stride, int mode, uint32_t ea,
                         const uint32_t (&Aw)[K], uint32_t eb, const uint32_t (&Bw)[K], uint32_t &re,
                         uint32_t (&R)[K]) {
  constexpr int KD = mul_digits_per_lane(K);
  constexpr int KDP = mul_digits_padded(K);
  constexpr int ND = KD * L;               /* digits per operand */
  static_assert(2 * L <= 64, "low and high chunk carries share one 64-bit mask");
  /* a column sums at most ND products below 2^54; on the lockstep row path a window can
   * also hold one chunk of the neighbour group on top of its own */
  static_assert(ND <= 1024, "column sums of 27-bit digits must stay below 2^64");
  static_assert(!(kLockstep && mul_row_path(L) && L < 16) || ND <= 512,
                "own column sum plus one neighbour chunk must stay below 2^64");
  const uint32_t act = k_mul_tab[fpa_classify(ea) * 9 + fpa_classify(eb)];
  if (!kLockstep && act != FPA_A_MUL) {
    re = act;
#pragma unroll
    for (int i = 0; i < K; i++) R[i] = 0;
    return;
  }

  /* recode both operands into radix-2^27 digit chunks via LDS */
  uint32_t A[KD], B[KD];
  recode_digits<K, L, KD>(g, lds, Aw, A);
  recode_digits<K, L, KD>(g, lds, Bw, B);

  /* Product scanning: round r multiplies a's chunk r (broadcast to the group) by this
   * lane's chunk of b into a window of 2*KD 64-bit column sums (column chunks r+gl and
   * r+gl+1).  Lane 0's low half is then complete and goes to LDS; every lane's new low
   * half is its high half plus the next lane's low half. */
  uint64_t W[2 * KD];
#pragma unroll
  for (int i = 0; i < 2 * KD; i++) W[i] = 0;
  uint64_t *low = reinterpret_cast<uint64_t *>(lds);   /* ND 64-bit low column sums */
  if constexpr (mul_lds_broadcast(L)) {
    uint32_t *adig = lds + mul_adig_offset(K, L);       /* a's digit chunks, KDP apart */
#pragma unroll
    for (int k = 0; k < KD; k++) adig[g.gl * KDP + k] = A[k];
    wsync();
    for (int r = 0; r < L; r++) {
      uint32_t ar[KDP];
#pragma unroll
      for (int q = 0; q < KDP / 4; q++) {
        const uint4 t = reinterpret_cast<const uint4 *>(adig + r * KDP)[q];
        ar[4 * q] = t.x; ar[4 * q + 1] = t.y; ar[4 * q + 2] = t.z; ar[4 * q + 3] = t.w;
      }
#pragma unroll
      for (int i = 0; i < KD; i++) {
#pragma unroll
        for (int j = 0; j < KD; j++) W[i + j] += (uint64_t)ar[i] * B[j];
      }
      if (g.gl == 0) {
#pragma unroll
        for (int i = 0; i < KD; i++) low[r * KD + i] = W[i];
      }
      if constexpr (mul_row_path(L)) {
        /* Lockstep: a group's top lane reads lane 0 of the next group in its row (not
         * zero); that value is the next group's complete chunk r, and it ends up,
         * unchanged, in lane r's final window, where it is subtracted below. */
        if constexpr (kLockstep) {
          /* the blocks all write VCC, so they stay in this order: the first follows the
           * EXEC restore after lane 0's stores, the last precedes code that may read its
           * results through DPP */
#pragma unroll
          for (int i = 0; i < KD; i++) {
            if (KD == 1) W[i] = add_row_next64<4, true>(W[KD + i], W[i]);
            else if (i == 0) W[i] = add_row_next64<4, false>(W[KD + i], W[i]);
            else if (i == KD - 1) W[i] = add_row_next64<1, true>(W[KD + i], W[i]);
            else W[i] = add_row_next64<1, false>(W[KD + i], W[i]);
            W[KD + i] = 0;
          }
        } else {
          /* every lane moves its neighbour's words first and the top lane zeroes them
           * after: written as a select of the moved value, the compiler branches around
           * the DPP moves for the top lane */
          const bool top = g.gl == L - 1;
#pragma unroll
          for (int i = 0; i < KD; i++) {
            uint32_t lo = dpp_row_next((uint32_t)W[i]);
            uint32_t hi = dpp_row_next((uint32_t)(W[i] >> 32));
            lo = top ? 0u : lo;
            hi = top ? 0u : hi;
            unsigned int c, c2;
            const uint32_t nlo = __builtin_addc(lo, (uint32_t)W[KD + i], 0u, &c);
            const uint32_t nhi = __builtin_addc(hi, (uint32_t)(W[KD + i] >> 32), c, &c2);
            W[i] = ((uint64_t)nhi << 32) | nlo;
            W[KD + i] = 0;
          }
        }
      } else {
        /* the group spans DPP rows: next window = own high half + next lane's low half */
#pragma unroll
        for (int i = 0; i < KD; i++) {
          uint32_t lo = g.from_next((uint32_t)W[i]);
          uint32_t hi = g.from_next((uint32_t)(W[i] >> 32));
          W[i] = W[KD + i] + (((uint64_t)hi << 32) | lo);
          W[KD + i] = 0;
        }
      }
    }
  } else {
    /* one lane per value: the whole product is in this lane's window */
#pragma unroll
    for (int i = 0; i < KD; i++) {
#pragma unroll
      for (int j = 0; j < KD; j++) W[i + j] += (uint64_t)A[i] * B[j];
    }
#pragma unroll
    for (int i = 0; i < KD; i++) low[i] = W[i];
#pragma unroll
    for (int i = 0; i < KD; i++) { W[i] = W[KD + i]; W[KD + i] = 0; }
  }
  wsync();
  if constexpr (kLockstep && mul_row_path(L) && L < 16) {
    /* groups whose top lane is not the last lane of its DPP row took in the next group's
     * complete chunks: lane j's final window holds that group's chunk j on top */
    if (((g.gbase + L) & 15) != 0) {
      const uint64_t *nlow = reinterpret_cast<const uint64_t *>(lds + group_stride);
#pragma unroll
      for (int i = 0; i < KD; i++) W[i] -= nlow[g.gl * KD + i];
    }
  }

  uint64_t cl[KD], ch[KD];
  const uint64_t *lrow = low + g.gl * KD;     /* one base address, constant offsets */
#pragma unroll
  for (int i = 0; i < KD; i++) { cl[i] = lrow[i]; ch[i] = W[i]; }
  bool amb;
  mul_finish<K, L, false>(g, lds, cl, ch, mode, ea, eb, re, R, amb);
  if (kLockstep && act != FPA_A_MUL) {
    re = act;
#pragma unroll
    for (int i = 0; i < K; i++) R[i] = 0;
  }
}

/* Columns to the rounded product.  cl / ch: this lane's low and high chunk of column
 * sums (columns gl*KD + i and ND + gl*KD + i).  The columns are normalised to radix-2^27
 * digits, the mantissa words and round word are extracted and the result is rounded.
 * kShort: the columns hold a short product Q with Q <= A*B < Q + 2^(27*T + 36) (the
 * omitted digit products all lie in columns below T, mul_short); the rounding is then
 * decided by Q alone, with sticky 1, unless the 64 bits of Q just below the round word
 * are all zeros or all ones: amb reports that case (the same for every lane of the group),
 * and the caller recomputes the product in full.  LOWC0 >= 0: only the top lane holds a
 * low chunk, whose columns below LOWC0 are zero (no carry enters a low chunk). */
template <int K, int L, bool kShort, int LOWC0>
__device__ __forceinline__ void mul_finish(const Group<L> &g, uint32_t *lds,
                                           const uint64_t (&cl)[mul_digits_per_lane(K)],
                                           const uint64_t (&ch)[mul_digits_per_lane(K)], int mode, uint32_t ea,
                                           uint32_t eb, uint32_t &re, uint32_t (&R)[K], bool &amb) {
  constexpr int n = K * L;
  constexpr int KD = mul_digits_per_lane(K);
  constexpr int ND = KD * L;               /* digits per operand */
  constexpr uint32_t M = (1u << MUL_RADIX) - 1;

  /* normalise: chain of 2L chunks, low chunks 0..L-1 then high chunks 0..L-1 */
  uint32_t dl[KD], dh[KD];
  {
    wsync();   /* every lane has read its columns before the digits overwrite them */
    uint64_t co_l = normalize_columns<KD, (LOWC0 > 0 ? LOWC0 : 0)>(cl, dl, 0);
    uint64_t co_h = normalize_columns<KD>(ch, dh, 0);
    /* multi-bit carries from the preceding chunk */
    uint32_t ph_lo = g.from_prev_row0((uint32_t)co_h), ph_hi = g.from_prev_row0((uint32_t)(co_h >> 32));
    uint64_t top_l = ((uint64_t)g.bcast_top((uint32_t)(co_l >> 32)) << 32) | g.bcast_top((uint32_t)co_l);
    uint64_t in_h = g.gl == 0 ? top_l : (((uint64_t)ph_hi << 32) | ph_lo);
    /* chunk carries are below 2^38: a column sum is below 2^64 */
    uint32_t ol = 0;
    if constexpr (LOWC0 < 0) {
      uint32_t pl_lo = g.from_prev_row0((uint32_t)co_l), pl_hi = g.from_prev_row0((uint32_t)(co_l >> 32));
      ol = add_carry_digits_wave<KD>(dl, ((uint64_t)pl_hi << 32) | pl_lo);
    }
    uint32_t oh = add_carry_digits_wave<KD>(dh, in_h);
    /* 1-bit carries along the chain: bit t = low chunk t, bit L+t = high chunk t */
    bool pl = true, ph = true;
#pragma unroll
    for (int i = 0; i < KD; i++) { pl &= dl[i] == M; ph &= dh[i] == M; }
    uint64_t G = g.ballot(ol != 0) | (g.ballot(oh != 0) << L);
    uint64_t P = g.ballot(pl) | (g.ballot(ph) << L);
    uint64_t X = G | P;
    uint64_t S = X + G;                      /* 2L <= 64 bits; the value's top carry is zero */
    uint64_t C = S ^ X ^ G;
    const uint32_t cl1 = (uint32_t)(C >> g.gl) & 1u, ch1 = (uint32_t)(C >> (L + g.gl)) & 1u;
    /* a 1-bit carry needs a chunk of all-ones digits: rare, so test the whole wave */
    if (__builtin_expect(__ballot((cl1 | ch1) != 0) != 0, 0)) {
      add_carry_digits<KD>(dl, cl1);
      add_carry_digits<KD>(dh, ch1);
    }
  }
  /* P = A*B < 2^(2p); mantissa = top p bits from bit `base`, round word below, sticky */
  constexpr int p = 32 * n;
  constexpr int topdig = (2 * p - 1) / MUL_RADIX, topoff = (2 * p - 1) % MUL_RADIX;
  bool full;
  uint32_t H[K], rw;
  if constexpr (mul_extract_nb(K, L) > 0) {
    /* In registers: lane gl's mantissa words start at bit base + 32K*gl, at most
     * 27*NB - 1 bits below its high digits (27*KD bits from B = 27*(ND + KD*gl)), so the
     * lane packs its high digits and the NB digits below them (the predecessor lane's top
     * high digits; for lane 0 the low half's top digits) into 32-bit words and funnel-
     * shifts them. */
    constexpr int NB = mul_extract_nb(K, L), DL = mul_delta(K), NDG = NB + KD;
    constexpr int SMIN = MUL_RADIX * NB - 1 + DL * (2 * L - 1), OFF = SMIN / 32 * 32;
    constexpr int hd = topdig - ND;
    full = (g.template bcast_lane<hd / KD>(dh[hd % KD]) >> topoff) & 1u;
    uint32_t Dg[NDG];
    /* the predecessor lane's top high digits; lane 0 takes the low half's top digits, held
     * by lane L-1 */
#pragma unroll
    for (int m = 0; m < NB; m++) {
      const uint32_t fromtop = g.bcast_top(dl[KD - NB + m]);
      Dg[m] = g.from_prev_row(dh[KD - NB + m], fromtop);
    }
#pragma unroll
    for (int k = 0; k < KD; k++) Dg[NB + k] = dh[k];
    /* dense[w + 1] = bits [OFF + 32w, +32) of the string Dg (27 bits per digit), w = -1..K+1 */
    uint32_t dense[K + 3];
#pragma unroll
    for (int w = -1; w <= K + 1; w++) {
      const int t = OFF + 32 * w;
      uint32_t x = 0;
      if (t >= 0) {
        const int m = t / MUL_RADIX, sh = t % MUL_RADIX;
        if (m < NDG) x = Dg[m] >> sh;
        if (m + 1 < NDG) x |= Dg[m + 1] << (MUL_RADIX - sh);
        if (m + 2 < NDG && 2 * MUL_RADIX - sh < 32) x |= Dg[m + 2] << (2 * MUL_RADIX - sh);
      }
      dense[w + 1] = x;
    }
    /* this lane's first mantissa bit, relative to bit OFF of the string: 0 <= st < 64 */
    const int st = MUL_RADIX * NB - (full ? 0 : 1) + DL * (L + g.gl) - OFF;
    const bool j0 = st >= 32;
    const uint32_t sh = (uint32_t)st & 31u;
    uint32_t sel[K + 2];
#pragma unroll
    for (int w = 0; w < K + 2; w++) sel[w] = j0 ? dense[w + 1] : dense[w];
#pragma unroll
    for (int i = 0; i < K; i++) H[i] = __builtin_amdgcn_alignbit(sel[i + 2], sel[i + 1], sh);
    /* the round word: the 32 bits below group lane 0's first mantissa bit */
    rw = g.template bcast_lane<0>(__builtin_amdgcn_alignbit(sel[1], sel[0], sh));
  } else {
    /* digits to LDS: [0, ND) low, [ND, 2ND) high, then two zero words */
#pragma unroll
    for (int i = 0; i < KD; i++) { lds[g.gl * KD + i] = dl[i]; lds[ND + g.gl * KD + i] = dh[i]; }
    if (g.gl == L - 1) { lds[2 * ND] = 0; lds[2 * ND + 1] = 0; }
    wsync();
    full = (lds[topdig] >> topoff) & 1u;
    const int base0 = full ? p : p - 1;
#pragma unroll
    for (int i = 0; i < K; i++) H[i] = digits_bits(lds, base0 + 32 * (g.gl * K + i));
    rw = digits_bits(lds, base0 - 32);
  }
  const int base = full ? p : p - 1;
  uint32_t sticky;
  if constexpr (kShort) {
    /* Q = A*B - E with 0 <= E < 2^(27T+36) <= 2^(p-97).  If bits [p-97, p-33) of Q hold
     * both a zero and a one, Q + E has the same bits from p-33 up (no carry reaches them)
     * and a nonzero part below, so mantissa, round word and sticky = 1 are those of A*B
     * in either normalisation (round word at p-33 or p-32).  The window lies in the low
     * half's digits WD0..WD1, in at most two lanes. */
    constexpr int WB = p - 97, WE = p - 33;
    constexpr int WD0 = WB / MUL_RADIX, WD1 = (WE - 1) / MUL_RADIX;
    constexpr int WL0 = WD0 / KD, WL1 = WD1 / KD;
    static_assert(WB >= 0 && WL1 - WL0 <= 1 && WL1 < L, "check window inside the low half, two lanes at most");
    bool z = true, o = true;
    auto window_part = [&](int wl) {
      bool zz = true, oo = true;
#pragma unroll
      for (int k = 0; k < KD; k++) {
        const int d0 = MUL_RADIX * (wl * KD + k);
        const int lo = WB - d0 > 0 ? WB - d0 : 0, hi = WE - d0 < MUL_RADIX ? WE - d0 : MUL_RADIX;
        if (hi > lo) {
          const uint32_t m = ((1u << (hi - lo)) - 1u) << lo;
          zz &= (dl[k] & m) == 0u;
          oo &= (dl[k] & m) == m;
        }
      }
      if (g.gl == wl) { z = zz; o = oo; }
    };
    window_part(WL0);
    if constexpr (WL1 != WL0) window_part(WL1);
    amb = g.ballot(!z) == 0 || g.ballot(!o) == 0;
    sticky = 1u;
  } else {
    uint32_t loc;
    /* sticky covers bits [0, lim), lim = base - 32 in {p - 33, p - 32}: low digits below
     * digit D0 = (p - 33) / 27 are whole, digit D0 holds lim (p - 32 - 27*D0 <= 27), and
     * D0 sits in lane D0 / KD at index D0 % KD */
    constexpr int D0 = (p - 33) / MUL_RADIX, lane_b = D0 / KD, i_b = D0 % KD;
    const int lim = base - 32;
    uint32_t pre = 0, rest = 0;
#pragma unroll
    for (int i = 0; i < KD; i++) {
      if (i < i_b) pre |= dl[i];
      else rest |= dl[i];
    }
    const uint32_t part = dl[i_b] & ((1u << (lim - MUL_RADIX * D0)) - 1u);
    loc = g.gl < lane_b ? (pre | rest) : g.gl == lane_b ? (pre | part) : 0u;
    sticky = g.any(loc != 0) ? 1u : 0u;
    amb = false;
  }
  wsync();

  const uint32_t sign = (ea ^ eb) & 1u;
  int64_t E = (int64_t)(ea >> 1) + (int64_t)(eb >> 1) - FPA_BIAS_LL - (full ? 0 : 1);
  if (round_mantissa<K, L>(g, H, mode, sign, rw, sticky)) E += 1;
  re = pack_exp(E, sign, 5, 0x7FFFFFFFll);
  chunk_select(R, true, H, H);
}

/*
 * Short product.  The rounded product needs A*B only down to the round word and a
 * sticky bit; MPFR's method (mpfr_mulhigh_n, then mpfr_round_p) applies: sum only the
 * digit products of columns >= T, which leaves an error below 2^(27T+36) (at most 511
 * products per column, each below 2^54), round from that sum when its bits just below the
 * round word are not all equal, and form the full product only when they are.
 *
 * With chunks of KD digits, chunk pairs (alpha, beta) with alpha + beta >= F0 hold every
 * product of columns >= KD*F0; on the guard anti-diagonal alpha + beta = F0-1 only the
 * corner x + y >= G of each pair's KD x KD digit block reaches column T = KD*(F0-1) + G
 * (G >= KD, so the anti-diagonal below it holds no column >= T).  F0 = L-1 when T lies
 * above chunk L-1's start: L(L+1)/2 pairs plus L-1 small corners instead of L^2 pairs;
 * lower for 32-lane groups whose chunks are short against the 97 bits below the round
 * word (e.g. 4 limbs on 32 lanes at 4096 bits: F0 = 29, 589 pairs).  The pairs are not
 * spread evenly over the rounds of the product-scanning form (round r needs only lanes
 * gl >= F0-r), so every lane here runs its own list (fpa_mul_sched.h): two segments of a
 * fixed b chunk and a rising a row, the window of 2*KD column sums sliding one chunk per
 * round by register renaming, each finished chunk added into a column array in LDS
 * (64-bit LDS atomic adds: several lanes finish the same chunk).  Only the columns >= T are
 * kept (chunk F0 from column T on, and the chunks above it); what falls below is dropped.
 */
template <int K, int L>
struct MulShort {
  static constexpr int n = K * L, p = 32 * n;
  static constexpr int KD = mul_digits_per_lane(K), KDP = mul_digits_padded(K), ND = KD * L;
  /* the check window [p-97, p-33) must lie above the error: 27T + 36 <= p - 97 */
  static constexpr int TMAX = (p - 133) / MUL_RADIX;
  /* the lowest anti-diagonal of full chunk pairs: L-1, or lower for 32-lane groups whose
   * chunks are short against the 97 bits below the round word (with fewer lanes the full
   * product scan's DPP window shift beat the extra rounds' LDS atomics: 4x16 at 2048 bits
   * ran 2.2x slower on the short product, 8x16 at 4096 1.6x; 4x32 and 8x32, whose full scan
   * shifts through ds_bpermute, ran 25-30% faster) */
  static constexpr int F0 = L == 32 && TMAX / KD < L - 1 ? TMAX / KD : L - 1;
  /* corner threshold on anti-diagonal F0-1; 2KD-1 (no corner) once T clears the guard
   * pairs' top column */
  static constexpr int G = TMAX - KD * (F0 - 1) < 2 * KD - 1 ? TMAX - KD * (F0 - 1) : 2 * KD - 1;
  static constexpr int T = KD * (F0 - 1) + G;
  static constexpr bool kCorner = G <= 2 * KD - 2;
  /* a's digit chunks, then the 64-bit columns of chunks F0 .. 2L-1 */
  static constexpr int COL0 = L * KDP;
  static constexpr int NCOL = (2 * L - F0) * KD;
  static constexpr int words = COL0 + 2 * NCOL;
  /* T >= KD*F0: the chunks below F0 that are dropped hold only columns below T */
  static constexpr bool ok = F0 >= 1 && MulSched<L, F0>::ok && G >= KD && ND < 512 && 2 * L <= 64;
};

template <int K, int L>
__host__ __device__ constexpr bool mul_short_ok() { return MulShort<K, L>::ok; }

/* product rounds per lane: the short product's schedule, or L for the full scan */
template <int K, int L>
__host__ __device__ constexpr int mul_rounds() {
  if constexpr (mul_short_ok<K, L>()) return MulSched<L, MulShort<K, L>::F0>::R1 + MulSched<L, MulShort<K, L>::F0>::R2;
  else return L;
}

/* LDS words per value of op_mul: the short product's, and core_mul's for its fallback */
template <int K, int L>
__host__ __device__ constexpr int mul_op_lds_words() {
  return mul_short_ok<K, L>() && MulShort<K, L>::words > mul_lds_words(K, L) ? MulShort<K, L>::words
                                                                             : mul_lds_words(K, L);
}

/* one chunk (KD columns) of the window into the column array if the chunk is kept; of
 * chunk F0 only columns C0.. (those at or above T) */
template <int F0, int KD, int C0>
__device__ __forceinline__ void mul_emit(uint64_t *cols, int c, const uint64_t *w) {
  if (c >= F0) {
    /* an opaque offset keeps the base in one register and the columns as the
     * instructions' immediate offsets (folded, the chunk's negative start costs an add per
     * column) */
    int off = (c - F0) * KD;
    asm volatile("" : "+v"(off));
    uint64_t *dst = cols + off;
    if (c > F0) {
#pragma unroll
      for (int i = 0; i < C0; i++) __hip_atomic_fetch_add(dst + i, w[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
    }
#pragma unroll
    for (int i = C0; i < KD; i++) __hip_atomic_fetch_add(dst + i, w[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
  }
}

/* a's digit row alpha (KDP words from 16-byte-aligned LDS) */
template <int KD, int KDP>
__device__ __forceinline__ void mul_load_row(const uint32_t *adig, int alpha, uint32_t (&ar)[KDP]) {
  const uint32_t *row = static_cast<const uint32_t *>(__builtin_assume_aligned(adig + alpha * KDP, 16));
#pragma unroll
  for (int q = 0; q < KDP / 4; q++) {
    const uint4 t = reinterpret_cast<const uint4 *>(row)[q];
    ar[4 * q] = t.x; ar[4 * q + 1] = t.y; ar[4 * q + 2] = t.z; ar[4 * q + 3] = t.w;
  }
  /* keep the padding words' loads alive: the row then stays whole 16-byte reads instead
   * of being split into 4-byte pairs around the dead words */
  if constexpr (KD < KDP) {
#pragma unroll
    for (int i = KD; i < KDP; i++) asm volatile("" ::"v"(ar[i]));
  }
}

/* The lane's two segments: R1 rounds of b chunk beta1 with a rows alpha1.., then R2 rounds
 * of beta2 with rows alpha2...  Both b chunks are fetched up front and every round's a row
 * is read before the previous round's atomics are issued (which the compiler may not move
 * loads above), so the LDS latency overlaps a round of multiply-adds; chunk F0 keeps its
 * columns C0.. */
template <int R1, int R2, int L, int F0, int KD, int KDP, int C0>
__device__ __forceinline__ void mul_short_rounds(const Group<L> &g, const uint32_t *adig, uint64_t *cols,
                                                 const uint32_t (&B)[KD], int alpha1, int beta1, int alpha2,
                                                 int beta2) {
  constexpr int R = R1 + R2;
  uint32_t Bc1[KD], Bc2[KD];
#pragma unroll
  for (int j = 0; j < KD; j++) {
    Bc1[j] = g.bcast(B[j], beta1);
    Bc2[j] = g.bcast(B[j], beta2);
  }
  uint64_t W[2 * KD];
#pragma unroll
  for (int i = 0; i < 2 * KD; i++) W[i] = 0;
  uint32_t ar[KDP];
  mul_load_row<KD, KDP>(adig, alpha1, ar);
  /* straight-line rounds, so the window's slide is a register renaming */
  repeat<R>([&](int t) {
    const bool seg1 = t < R1;
    uint32_t nx[KDP];
    if (t + 1 < R) mul_load_row<KD, KDP>(adig, t + 1 < R1 ? alpha1 + t + 1 : alpha2 + (t + 1 - R1), nx);
#pragma unroll
    for (int i = 0; i < KD; i++) {
#pragma unroll
      for (int j = 0; j < KD; j++) W[i + j] += (uint64_t)ar[i] * (seg1 ? Bc1[j] : Bc2[j]);
    }
    const int c = seg1 ? alpha1 + t + beta1 : alpha2 + (t - R1) + beta2;
    mul_emit<F0, KD, C0>(cols, c, W);
    if (t == R1 - 1 || t == R - 1) {
      mul_emit<F0, KD, C0>(cols, c + 1, W + KD);
#pragma unroll
      for (int i = 0; i < 2 * KD; i++) W[i] = 0;
    } else {
#pragma unroll
      for (int i = 0; i < KD; i++) { W[i] = W[KD + i]; W[KD + i] = 0; }
    }
    if (t + 1 < R) {
#pragma unroll
      for (int i = 0; i < KDP; i++) ar[i] = nx[i];
    }
  });
}

/* (re, R) = a * b through the short product.  lds: max(MulShort::words, mul_lds_words)
 * words.  kLockstep: every group of the wave calls it together (the element-wise kernel),
 * and the whole wave falls back to core_mul when some group's rounding is not decided by
 * the short product.  Without it (dot chains, whose groups run chains of different
 * lengths) a group returns class results at once and falls back on its own; every step
 * here reads only the group's own lanes and LDS, so that stays exact. */
template <int K, int L, bool kLockstep = true>
__device__ __forceinline__ void core_mul_short(const Group<L> &g, uint32_t *lds, int group_stride, int mode,
                                               uint32_t ea, const uint32_t (&Aw)[K], uint32_t eb,
                                               const uint32_t (&Bw)[K], uint32_t &re, uint32_t (&R)[K]) {
  using S = MulShort<K, L>;
  constexpr int KD = S::KD, KDP = S::KDP;
  static_assert(S::ok, "shape without a short product");
  static_assert(K * L + (MUL_RADIX * KD * L + 31) / 32 + 1 - K * L <= S::COL0,
                "the recode's staging area lies below the column array");
  const uint32_t act = k_mul_tab[fpa_classify(ea) * 9 + fpa_classify(eb)];
  if (!kLockstep && act != FPA_A_MUL) {
    re = act;
#pragma unroll
    for (int i = 0; i < K; i++) R[i] = 0;
    return;
  }
  uint64_t *cols = reinterpret_cast<uint64_t *>(lds + S::COL0);
  {
    /* clear the column array, 16 bytes per store dealt over the group's lanes */
    constexpr int NV = (2 * S::NCOL + 3) / 4;
    const uint4 z = {0u, 0u, 0u, 0u};
#pragma unroll
    for (int q = 0; q < (NV + L - 1) / L; q++) {
      const int idx = q * L + g.gl;
      if (idx < NV) reinterpret_cast<uint4 *>(lds + S::COL0)[idx] = z;
    }
  }
  uint32_t A[KD], B[KD];
  if constexpr (recode_dpp_ok<K, L, KD>()) {
    recode_digits_dpp<K, L, KD>(g, Aw, A);
    recode_digits_dpp<K, L, KD>(g, Bw, B);
  } else {
    recode_digits<K, L, KD>(g, lds, Aw, A);
    recode_digits<K, L, KD>(g, lds, Bw, B);
  }
#pragma unroll
  for (int k = 0; k < KD; k++) lds[g.gl * KDP + k] = A[k];
  wsync();

  using Sched = MulSched<L, S::F0>;
  const uint32_t sc = Sched::lane(g.gl);
  constexpr int C0 = S::T - KD * S::F0;
  mul_short_rounds<Sched::R1, Sched::R2, L, S::F0, KD, KDP, C0>(
      g, lds, cols, B, (int)(sc & 31u), (int)((sc >> 5) & 31u), (int)((sc >> 10) & 31u), (int)((sc >> 15) & 31u));
  if constexpr (S::kCorner) {
    /* the corner x + y >= G of guard pair (ca, cb): columns KD*(F0-1) + x + y, i.e. column
     * array entries x + y - KD of chunk F0 */
    constexpr int X0 = S::G - (KD - 1), NC = KD - X0;
    const int ca = (int)((sc >> 20) & 31u), cb = (int)((sc >> 25) & 31u);
    uint32_t bc[NC];
#pragma unroll
    for (int y = 0; y < NC; y++) bc[y] = g.bcast(B[X0 + y], cb);   /* every lane takes part */
    if ((sc >> 30) & 1u) {
      const uint32_t *arow = lds + ca * KDP;
      uint64_t cs[2 * NC - 1];
#pragma unroll
      for (int t = 0; t < 2 * NC - 1; t++) cs[t] = 0;
#pragma unroll
      for (int x = 0; x < NC; x++) {
        const uint32_t ax = arow[X0 + x];
#pragma unroll
        for (int y = 0; y < NC; y++)
          if (x + y + 2 * X0 >= S::G) cs[x + y] += (uint64_t)ax * bc[y];
      }
#pragma unroll
      for (int t = 0; t < 2 * NC - 1; t++)
        if (t + 2 * X0 >= S::G)
          __hip_atomic_fetch_add(cols + (t + 2 * X0 - KD), cs[t], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
    }
  }
  wsync();

  /* lane gl: low chunk gl (kept only for gl >= F0), high chunk gl = chunk L+gl */
  uint64_t cl[KD], ch[KD];
  {
    const uint64_t *hrow = cols + (L - S::F0 + g.gl) * KD;
    const bool low = g.gl >= S::F0;
    const uint64_t *lrow = cols + (low ? g.gl - S::F0 : 0) * KD;
#pragma unroll
    for (int i = 0; i < KD; i++) {
      ch[i] = hrow[i];
      cl[i] = low ? lrow[i] : 0ull;
    }
  }
  bool amb;
  /* with F0 = L-1 only the top lane holds a low chunk, chunk L-1, from column C0 on */
  mul_finish<K, L, true, S::F0 == L - 1 ? C0 : -1>(g, lds, cl, ch, mode, ea, eb, re, R, amb);
  if constexpr (!kLockstep) {
    /* this group is undecided: it forms its product in full */
    if (__builtin_expect(amb, 0)) core_mul<K, L, false>(g, lds, group_stride, mode, ea, Aw, eb, Bw, re, R);
  } else if (__builtin_expect(__ballot(amb && act == FPA_A_MUL) != 0, 0)) {
    /* some group of the wave is undecided: the whole wave forms its products in full */
    core_mul<K, L, true>(g, lds, group_stride, mode, ea, Aw, eb, Bw, re, R);
  } else if (act != FPA_A_MUL) {
    re = act;
#pragma unroll
    for (int i = 0; i < K; i++) R[i] = 0;
  }
}

template <int K, int L>
__device__ __forceinline__ void op_mul(const fpa_arguments &a, const Slot &s, const Group<L> &g, uint32_t *lds, int group_stride) {
  uint32_t A[K], B[K], R[K], re;
  load_value<K, L>(g, a.a_limbs, s, A);
  load_value<K, L>(g, a.b_limbs, s, B);
  if constexpr (mul_short_ok<K, L>())
    core_mul_short<K, L>(g, lds, group_stride, (int)a.mode, a.a_exps[s.v], A, a.b_exps[s.v], B, re, R);
  else
    core_mul<K, L, true>(g, lds, group_stride, (int)a.mode, a.a_exps[s.v], A, a.b_exps[s.v], B, re, R);
  store_result(a, s, g.gl, re, R);
}

/* -------------------------------- div -------------------------------- */
/* ------------------- div with lazy radix-2^27 column sums ------------------- */
/*
 * Long division in radix beta = 2^27 with the remainder kept as signed 64-bit column sums
 * that are never carry-propagated inside the digit loop (the same idea as mul's column
 * sums).  Divisor and remainder are recoded to KD = ceil(32K/27) digits per lane, aligned
 * to the top (the divisor's top digit has its top bit set).  Each step:
 *   1. R <- R * beta: every column moves up one position (a register renaming inside the
 *      unrolled trip; one 64-bit column crosses each lane boundary); the column leaving
 *      the top is folded into the new top column (c_top * beta + c, exact modulo 2^64,
 *      and small in truth because the remainder stays below 2D);
 *   2. the group's top lane estimates the digit q = floor(V / Dt) in FP64 from its top two
 *      columns V = c1 * beta + c0 and the divisor's top two digits Dt in [2^53, 2^54), and
 *      broadcasts it.  Only c0's high word enters (one 32-bit conversion): that drops less
 *      than 2^32 of V, under 2^-21 of a digit, and the columns below c0 add less than 2^-17
 *      of one;
 *   3. every lane subtracts q * d from its columns: one v_mad_i64_i32 per column with the
 *      negated divisor digits.
 * The estimate lies within 2^-16 of the true quotient x, so q = floor of it leaves x - q in
 * (-2^-16, 1 + 2^-16): the digits are redundant (signed, |q| < beta + 2^12) and the
 * remainder stays in (-D, 2D); |q * d| < 2^55, so a column that lives ND <= 255 steps stays
 * below 2^63, and longer values renormalise every 128 steps by one carry level.  At the
 * end, the FP64 estimate of R/D decides (R in (0, D): no correction, remainder nonzero)
 * unless it lies within 2^-16 of 0 or 1; then R is normalised exactly
 * and corrected by one D (quotient -1 or +1).  The quotient digits (one per step, in LDS)
 * are carry-normalised only when some digit lies outside [0, beta).  The reference's
 * division (fpa_warp_div.h:127-201) runs 32-bit digits with exact remainders instead.
 */
constexpr int DIV_RADIX = 27;

__host__ __device__ constexpr int div_digits_per_lane(int K) { return (32 * K + DIV_RADIX - 1) / DIV_RADIX; }
/* quotient digits: enough for the p-bit mantissa, the 32-bit round word and a bit more */
__host__ __device__ constexpr int div_steps(int n) { return (32 * (n + 1) + DIV_RADIX - 1) / DIV_RADIX; }
/* zero words staged below a value so that the top-aligned digits' padding reads zeros */
__host__ __device__ constexpr int div_pad_words(int K, int L) {
  return (DIV_RADIX * div_digits_per_lane(K) * L - 32 * K * L + 31) / 32;
}

/* The short division's phases: TA trips of KD steps on the whole frame (JA steps), then S-JA
 * steps on the top FB = KB*L slots (B below them).  After step j no later digit and not the
 * end's estimate reads a slot below j - C1 (slot s reaches slot ND-2, where the estimates
 * read, only through ND-2-s shifts), so the frame of the second phase is the slots from
 * JA-2-C1 up (one slot of margin).  Its relayout passes through LDS words [0, FB), below the
 * quotient digits, which move up by X words to leave them free. */
struct DivShortShape {
  int KD, ND, S, TA, JA, C1, KB, FB, B, X;
  bool ok;
  /* three phases (long frames: ND >= 300, or ND >= 150 with lanes of at most 24 limbs; for
   * shorter frames and for 32-limb lanes of 152 slots the third frame's relayout costs more
   * than it saves): the whole frame for about S/3 steps, then KB columns until JB, then the
   * top FC = KC*L slots (BC below them) */
  bool T3;
  int JB, KC, FC, BC;
};
__host__ __device__ constexpr DivShortShape div_short_shape(int K, int L) {
  DivShortShape d{};
  d.KD = div_digits_per_lane(K);
  d.ND = d.KD * L;
  d.S = div_steps(K * L);
  d.C1 = d.S - d.ND + 1;
  const bool t3 = d.ND >= 300 || (d.ND >= 150 && K <= 24);
  d.TA = t3 ? (d.S + d.KD + d.KD / 2) / (3 * d.KD) : (d.S + d.KD) / (2 * d.KD);
  d.JA = d.TA * d.KD;
  d.KB = (d.ND - d.JA + 2 + d.C1 + L - 1) / L;
  d.FB = d.KB * L;
  d.B = d.ND - d.FB;
  d.X = d.FB > d.S - d.JA ? d.FB - (d.S - d.JA) : 0;
  d.ok = d.TA >= 1 && d.JA < d.S && d.KB >= 2 && d.KB < d.KD && d.B >= 0 && d.B <= d.JA - 2 - d.C1;
  d.JB = d.JA + (d.S - d.JA) / 2;
  d.KC = (d.ND - d.JB + 2 + d.C1 + L - 1) / L;
  d.FC = d.KC * L;
  d.BC = d.ND - d.FC;
  d.T3 = t3 && d.ok && d.KC >= 2 && d.KC < d.KB && d.BC >= d.B && d.BC <= d.JB - 2 - d.C1;
  if (d.T3 && d.FC > d.X + d.S - d.JB) d.X = d.FC - (d.S - d.JB);
  return d;
}

/* LDS words per value: recoding stage or the quotient digits (+ topbit, + two zeros; X words
 * up in the short division) */
__host__ __device__ constexpr int div_lds_words(int K, int L) {
  return div_pad_words(K, L) + K * L + 2 > div_short_shape(K, L).X + div_steps(K * L) + 3
             ? div_pad_words(K, L) + K * L + 2
             : div_short_shape(K, L).X + div_steps(K * L) + 3;
}

/* Columns per lane of the short sqrt's frame: the top FB = KB*L >= (F + 2) / 2 slots of the
 * F = KD*L root slots (ceil) */
__host__ __device__ constexpr int sqrt_short_cols(int KD, int L) { return (KD * L + 2 + 2 * L - 1) / (2 * L); }
template <int K, int L>
__host__ __device__ constexpr bool sqrt_short_ok() {
  constexpr int KD = div_digits_per_lane(K), F = KD * L, KB = sqrt_short_cols(KD, L), B0 = F - KB * L;
  return KB >= 2 && KB < KD && B0 >= 1 && 2 * B0 <= F - 2;
}

/* LDS words per value of the lazy sqrt: recoding stage, the radicand digits of the half-frame
 * phase (with L zeros below and KH above) and of the short frame (2L + 2 zeros below, KB
 * above), or the root digits (+ three zeros) */
template <int K, int L>
__host__ __device__ constexpr int sqrt_lazy_lds_words() {
  constexpr int KD = div_digits_per_lane(K), recode = div_pad_words(K, L) + K * L + 2, F = KD * L;
  constexpr int half = L + F + (KD + 1) / 2, root = F + 3;
  constexpr int shrt = sqrt_short_ok<K, L>() ? 2 * L + 2 + F + sqrt_short_cols(KD, L) : 0;
  constexpr int m1 = recode > half ? recode : half, m2 = root > shrt ? root : shrt;
  return m1 > m2 ? m1 : m2;
}

/* x as held by group lane src, 64 bits */
template <int L>
__device__ __forceinline__ int64_t bcast64(const Group<L> &g, int64_t x, int src) {
  const uint32_t lo = g.bcast((uint32_t)x, src), hi = g.bcast((uint32_t)((uint64_t)x >> 32), src);
  return (int64_t)(((uint64_t)hi << 32) | lo);
}

/* signed 64-bit value as a double (3 VALU: two conversions and one fma) */
__device__ __forceinline__ double i64_to_f64(int64_t x) {
  return __builtin_fma((double)(int32_t)((uint64_t)x >> 32), 4294967296.0, (double)(uint32_t)x);
}

/* A lazy frame's digit estimate floor(x), x = V * inv32 (|x| < 2^31), as x - 1/2 rounded to
 * nearest by adding 1.5 * 2^52, whose low word is then the integer: one add in place of a
 * floor and a conversion.  The fma's rounding (below 2^-23) can move the result by one where
 * x lies that close to an integer: the remainder then ends at most about 2^-22 of the divisor
 * outside [0, D), well inside the (-D, 2D) that the digit steps keep (the next digit comes
 * out one higher or lower), and the end correction treats such a remainder as unsure. */
__device__ __forceinline__ int32_t lazy_digit(double V, double inv32) {
  const double y = __builtin_fma(V, inv32, -0.5) + 0x1.8p52;
  return (int32_t)(uint32_t)__builtin_bit_cast(uint64_t, y);
}

/* The digit estimate from a lazy frame's top two columns c1 (top) and c0: floor(V * 2^32 / Dt)
 * for V = c1 * 2^-5 + floor(c0 / 2^32), i.e. floor(W * invs) for the integer W = 32 V =
 * c1 + 32 floor(c0 / 2^32) (one v_mad_i64_i32) and invs = 2^27 / Dt */
__device__ __forceinline__ int32_t lazy_top_digit(int64_t c1, int64_t c0, double invs) {
  int32_t k32;   /* 32 in an SGPR the compiler cannot see through: it would shift and add */
  asm("s_mov_b32 %0, 32" : "=s"(k32));
  const int64_t w = (int64_t)(int32_t)((uint64_t)c0 >> 32) * (int64_t)k32 + c1;
  return lazy_digit(i64_to_f64(w), invs);
}

/* U <- U * beta on a lazy frame of KC columns per lane: the columns move up one slot, the one
 * leaving a lane enters the next lane's bottom slot (zero in group lane 0), the one leaving
 * the top lane folds into the new top column.  nz0 (all ones except in group lane 0) is a
 * lane mask the caller keeps opaque, so that it stays an AND: the move from the previous
 * lane is then one VALU per word (the AND reads the DPP move directly). */
template <int KC, int L>
__device__ __forceinline__ void lazy_shift_up(const Group<L> &g, int64_t (&C)[KC], uint32_t nz0, bool top) {
  const uint64_t so = (uint64_t)C[KC - 1];
  const uint32_t lo = g.from_prev_row_and((uint32_t)so, nz0);
  const uint32_t hi = g.from_prev_row_and((uint32_t)(so >> 32), nz0);
#pragma unroll
  for (int i = KC - 1; i > 0; i--) C[i] = C[i - 1];
  C[0] = (int64_t)(((uint64_t)hi << 32) | lo);
  C[KC - 1] = (int64_t)((uint64_t)C[KC - 1] + (top ? so << DIV_RADIX : 0u));
}

/* lazy_shift_up with the fold as one v_mad_u64_u32 of the low word (ftop = 2^27 in the top
 * lane, 0 elsewhere) and a shifted add of the high word's low 5 bits (mtop: all ones in the
 * top lane), three VALU instead of a 64-bit shift, two selects and an add; both lane
 * constants opaque, so that they stay registers */
template <int KC, int L>
__device__ __forceinline__ void lazy_shift_up(const Group<L> &g, int64_t (&C)[KC], uint32_t nz0, uint32_t ftop,
                                              uint32_t mtop) {
  const uint64_t so = (uint64_t)C[KC - 1];
  const uint32_t lo = g.from_prev_row_and((uint32_t)so, nz0);
  const uint32_t hi = g.from_prev_row_and((uint32_t)(so >> 32), nz0);
#pragma unroll
  for (int i = KC - 1; i > 0; i--) C[i] = C[i - 1];
  C[0] = (int64_t)(((uint64_t)hi << 32) | lo);
  const uint64_t t = (uint64_t)(uint32_t)so * ftop + (uint64_t)C[KC - 1];
  const uint32_t th = (uint32_t)(t >> 32) + (((uint32_t)(so >> 32) & mtop) << DIV_RADIX);
  C[KC - 1] = (int64_t)(((uint64_t)th << 32) | (uint32_t)t);
}

/* One carry level over a lazy frame, in place from the top down (column i takes column
 * i-1's carry before that column is reduced): every column but the top lane's top one keeps
 * its low 27 bits and takes the carry of the column below (zero into group lane 0's bottom).
 * The value is unchanged.  The steps run it between trips, never inside one, so that its
 * code is not repeated at every unrolled step. */
template <int KC, int L>
__device__ __forceinline__ void lazy_renorm(const Group<L> &g, int64_t (&C)[KC], bool top) {
  constexpr int64_t MASK = ((int64_t)1 << DIV_RADIX) - 1;
  const int64_t ctop = C[KC - 1] >> DIV_RADIX;
  const uint32_t clo = g.from_prev_row0((uint32_t)ctop);
  const uint32_t chi = g.from_prev_row0((uint32_t)((uint64_t)ctop >> 32));
  const int64_t cbelow = (int64_t)(((uint64_t)chi << 32) | clo);
#pragma unroll
  for (int i = KC - 1; i >= 0; i--) {
    const int64_t keep = (top && i == KC - 1) ? C[i] : (C[i] & MASK);
    C[i] = keep + (i > 0 ? (C[i - 1] >> DIV_RADIX) : cbelow);
  }
}

/* One 32-bit word per slot moved from one lazy frame into another through LDS: the slots of
 * a frame of KC1 columns per lane whose bottom slot is B1 (lane gl holds B1 + gl*KC1 ..) go to
 * R[slot - SB] where slot >= SB; where the new frame reaches below B1 (SB < B1), 'pre' writes
 * R[0, B1 - SB); then lane gl takes the new frame's column i (slot SB + gl*KC2 + i) from
 * R[gl*KC2 + i].  get(i) gives the old column i's word, set(i, w) takes the new one's. */
template <int KC1, int KC2, int B1, int SB, int L, class Get, class Pre, class Set>
__device__ __forceinline__ void frame_relayout(const Group<L> &g, uint32_t *R, Get get, Pre pre, Set set) {
  wsync();
  if constexpr (SB <= B1) {
#pragma unroll
    for (int i = 0; i < KC1; i++) R[B1 - SB + g.gl * KC1 + i] = get(i);
    pre();
  } else {
    constexpr int LB = (SB - B1) / KC1, IB = (SB - B1) - LB * KC1;
    if (g.gl > LB) {
#pragma unroll
      for (int i = 0; i < KC1; i++) R[B1 - SB + g.gl * KC1 + i] = get(i);
    } else if (g.gl == LB) {
#pragma unroll
      for (int i = IB; i < KC1; i++) R[i - IB] = get(i);
    }
  }
  wsync();
#pragma unroll
  for (int i = 0; i < KC2; i++) set(i, R[g.gl * KC2 + i]);
}

/* Digits of X (n = K*L words) in radix 2^27 aligned to the top of KD*L digits: digit k of
 * the value = bits [27k - PAD + extra, +27) of X, PAD = 27*KD*L - 32n (bits below X's bit
 * 0 read as zero).  lds: ZW + n + 2 words, ZW = ceil(PAD / 32). */
template <int K, int L, int KD>
__device__ __forceinline__ void recode_digits_top(const Group<L> &g, uint32_t *lds, const uint32_t (&X)[K],
                                                  uint32_t (&D)[KD], int extra = 0) {
  constexpr int n = K * L, PAD = DIV_RADIX * KD * L - 32 * n, ZW = (PAD + 31) / 32;
  constexpr uint32_t M = (1u << DIV_RADIX) - 1;
  static_assert(PAD >= 0, "digits cover the value");
  lds_store(lds + ZW + g.gl * K, X);
  if (g.gl == 0) {
#pragma unroll
    for (int i = 0; i < ZW; i++) lds[i] = 0;
  }
  if (g.gl == L - 1) { lds[ZW + n] = 0; lds[ZW + n + 1] = 0; }
  wsync();
  /* the lane's digits are bits [b, b + 27*KD) of the staged string (its bit 32*ZW is X's
   * bit 0); the highest word read is floor(b/32) + NW <= ZW + n + 1.  extra = 1 takes the
   * digits of X / 2 (one bit less padding; extra <= PAD) */
  constexpr int NW = (DIV_RADIX * KD + 31) / 32;
  const int b = DIV_RADIX * KD * g.gl - PAD + extra + 32 * ZW;
  const uint32_t *src = lds + (b >> 5);
  const uint32_t sh = (uint32_t)b & 31u;
  uint32_t win[NW + 1];
  uint32_t lo = src[0];
#pragma unroll
  for (int i = 0; i < NW; i++) {
    const uint32_t hi = src[i + 1];
    win[i] = __builtin_amdgcn_alignbit(hi, lo, sh);
    lo = hi;
  }
  win[NW] = 0;
#pragma unroll
  for (int k = 0; k < KD; k++) {
    const int j = (DIV_RADIX * k) >> 5, s2 = (DIV_RADIX * k) & 31;
    D[k] = (s2 + DIV_RADIX <= 32 ? win[j] >> s2 : __builtin_amdgcn_alignbit(win[j + 1], win[j], (uint32_t)s2)) & M;
  }
  wsync();
}

/* recode_digits_top without LDS, for shapes whose lanes' digit windows drift by at most a
 * word over the group (16 and 32 limbs per lane): lane gl's digits start E*(gl - L) + extra
 * bits from its own word 0 (E = 27*KD - 32K, PAD = E*L), i.e. in the previous lane's top word,
 * and end at most one bit into the next lane's word 0; those two words come through DPP
 * (zero below and above the value).  The window starts c bits lower still where E = 1, so
 * that its start stays below word 0 for every lane and extra; the digits then sit c bits
 * into it. */
template <int K, int L, int KD>
__host__ __device__ constexpr bool recode_top_dpp_ok() {
  constexpr int E = DIV_RADIX * KD - 32 * K, c = E >= 2 ? 0 : 2 - E;
  return L >= 2 && L <= 32 && E >= 1 && E * L + c <= 32;
}
template <int K, int L, int KD>
__device__ __forceinline__ void recode_digits_top_dpp(const Group<L> &g, const uint32_t (&X)[K], uint32_t (&D)[KD],
                                                      int extra = 0) {
  constexpr int E = DIV_RADIX * KD - 32 * K, c = E >= 2 ? 0 : 2 - E;
  constexpr int NW = (DIV_RADIX * KD + c + 31) / 32;
  constexpr uint32_t M = (1u << DIV_RADIX) - 1;
  static_assert(recode_top_dpp_ok<K, L, KD>(), "digit window within the lane and its neighbours");
  static_assert(NW <= K + 1, "the window ends in the next lane's word 0");
  uint32_t w[K + 2];
  w[0] = g.from_prev_row(X[K - 1], 0u);
#pragma unroll
  for (int i = 0; i < K; i++) w[i + 1] = X[i];
  w[K + 1] = g.from_next_row0(X[0]);
  /* window start, bits from the lane's word 0: E*(gl-L) + extra - c, in [-32, -1] */
  const uint32_t sh = (uint32_t)(32 + E * (g.gl - L) + extra - c);
  uint32_t win[NW + 1];
#pragma unroll
  for (int i = 0; i < NW; i++) win[i] = __builtin_amdgcn_alignbit(w[i + 1], w[i], sh);
  win[NW] = 0;
#pragma unroll
  for (int k = 0; k < KD; k++) {
    const int j = (DIV_RADIX * k + c) >> 5, s2 = (DIV_RADIX * k + c) & 31;
    D[k] = (s2 + DIV_RADIX <= 32 ? win[j] >> s2 : __builtin_amdgcn_alignbit(win[j + 1], win[j], (uint32_t)s2)) & M;
  }
}

/* Exact normalisation, in place, of N signed columns per lane (lane gl holds positions
 * gl*N .. +N-1) into digits in [0, 2^27); returns the carry out of the top lane (the value's
 * top part): value = sum C_k beta^k + top * beta^(N*L).  Columns below 2^63 in magnitude. */
template <int N, int L>
__device__ __forceinline__ int64_t normalize_exact(const Group<L> &g, int64_t (&C)[N]) {
  constexpr int64_t M = ((int64_t)1 << DIV_RADIX) - 1;
  int64_t carry = 0;
#pragma unroll
  for (int i = 0; i < N; i++) {
    const int64_t v = C[i] + carry;
    C[i] = v & M;
    carry = v >> DIV_RADIX;
  }
  /* lane t's carry out moves into lane t+1, one lane at a time (the other lanes add zero
   * to digits that are already normalised, which changes nothing) */
  for (int t = 0; t < L - 1; t++) {
    const int64_t c = bcast64(g, carry, t);
    int64_t cin = g.gl == t + 1 ? c : 0;
#pragma unroll
    for (int i = 0; i < N; i++) {
      const int64_t v = C[i] + cin;
      C[i] = v & M;
      cin = v >> DIV_RADIX;
    }
    carry += cin;
  }
  return bcast64(g, carry, L - 1);
}

/* One digit step of the lazy division on a frame of KC columns per lane: R <- R * beta, the
 * digit from the top two columns (top lane) broadcast, R -= q * D, q to *qd */
template <int KC, int L>
__device__ __forceinline__ void div_lazy_step(const Group<L> &g, int64_t (&C)[KC], const int32_t (&nd)[KC], int32_t *qd,
                                              uint32_t nz0, uint32_t ftop, uint32_t mtop, double invs) {
  lazy_shift_up<KC, L>(g, C, nz0, ftop, mtop);
  const int32_t q = (int32_t)g.bcast_top((uint32_t)lazy_top_digit(C[KC - 1], C[KC - 2], invs));
#pragma unroll
  for (int i = 0; i < KC; i++) C[i] = (int64_t)q * nd[i] + C[i];
  *qd = q;
  /* keep the scheduler from running the digit chain of later steps ahead of this step's
   * multiply-adds: that lookahead holds several steps' columns live at once and costs
   * occupancy, while other waves hide the chain's latency anyway */
  __builtin_amdgcn_sched_barrier(0);
}

/* (re, Qr) = a / b; lds: div_lds_words(K, L) words.  kShort: the short division (second
 * phase on the top slots only); false where the end's estimate leaves some group of the wave
 * unsure, and the caller then runs the whole frame. */
template <int K, int L, bool kShort>
__device__ __forceinline__ bool core_div_lazy(const Group<L> &g, uint32_t *lds, int mode, uint32_t ea, const uint32_t (&A)[K],
                              uint32_t eb, const uint32_t (&B)[K], uint32_t &re, uint32_t (&Qr)[K]) {
  constexpr int n = K * L, p = 32 * n;
  constexpr int KD = div_digits_per_lane(K), ND = KD * L, S = div_steps(n);
  constexpr uint32_t MASK = (1u << DIV_RADIX) - 1;
  constexpr int64_t BETA = (int64_t)1 << DIV_RADIX;
  static_assert(KD >= 2, "the top lane holds the two top columns");
  static_assert(27 * S >= p + 32, "quotient digits cover the mantissa and the round word");
  const uint32_t act = k_div_tab[fpa_classify(ea) * 9 + fpa_classify(eb)];
  if (act != FPA_A_DIV) {
    re = act;
#pragma unroll
    for (int i = 0; i < K; i++) Qr[i] = 0;
    return true;
  }
  const uint32_t sign = (ea ^ eb) & 1u;
  int64_t E = (int64_t)(ea >> 1) - (int64_t)(eb >> 1) + FPA_BIAS_LL;
  /* leading quotient bit: A >= B; the remainder R0 = A or A - B is below B */
  uint32_t R0[K];
  const uint32_t below = g.sub(R0, A, B);
  const uint32_t topbit = below ? 0u : 1u;
  if (below) {
#pragma unroll
    for (int i = 0; i < K; i++) R0[i] = A[i];
  }
  uint32_t dd[KD], rd[KD];
  if constexpr (recode_top_dpp_ok<K, L, KD>()) {
    recode_digits_top_dpp<K, L, KD>(g, B, dd);
    recode_digits_top_dpp<K, L, KD>(g, R0, rd);
  } else {
    recode_digits_top<K, L, KD>(g, lds, B, dd);
    recode_digits_top<K, L, KD>(g, lds, R0, rd);
  }
  /* divisor's top two digits (meaningful in the top lane); invs = 2^27 / Dt scales the
   * step estimate */
  const double inv = 1.0 / __builtin_fma((double)dd[KD - 1], (double)BETA, (double)dd[KD - 2]);
  int32_t nd[KD];
  int64_t C[KD];
#pragma unroll
  for (int i = 0; i < KD; i++) {
    nd[i] = -(int32_t)dd[i];
    /* opaque, so the exact end path's -nd[i] is not taken from dd[i], which would keep the
     * divisor's digits live through the digit loop beside their negations (KD registers) */
    asm volatile("" : "+v"(nd[i]));
    C[i] = (int64_t)rd[i];
  }
  const double invs = inv * 0x1p27;
  const bool top = g.gl == L - 1;
  uint32_t nz0 = g.gl == 0 ? 0u : ~0u, mtop = top ? ~0u : 0u, ftop = mtop & (1u << DIV_RADIX);
  asm volatile("" : "+v"(nz0), "+v"(mtop), "+v"(ftop));
  constexpr DivShortShape SH = div_short_shape(K, L);
  static_assert(!kShort || SH.ok, "short division shape");
  /* quotient digit of step j at S-1-j (X words up in the short division) */
  int32_t *qd = reinterpret_cast<int32_t *>(lds) + (kShort ? SH.X : 0);
  /* one carry level at most every 128 steps keeps the columns far below 2^63 (frames of
   * more than 255 slots; a frame of at most 255 slots takes at most 255 products of < 2^55
   * per column) */
  constexpr int RT = 128 / KD;
  wsync();   /* the recoding reads of lds are done before the digits overwrite them */
  int j0 = 0;
  for (; j0 + KD <= (kShort ? SH.JA : S); j0 += KD) {
    if constexpr (ND > 255) {
      if (j0 > 0 && (j0 / KD) % RT == 0) lazy_renorm<KD, L>(g, C, top);
    }
    repeat<KD>([&](int u) __attribute__((always_inline)) {
      div_lazy_step<KD, L>(g, C, nd, qd + (S - 1 - j0 - u), nz0, ftop, mtop, invs);
    });
  }
  if constexpr (kShort) {
    /* the slots from B up into a frame of KB columns per lane (lane gl: B + gl*KB ..), one
     * 32-bit word per pass through LDS [0, FB); lane LB holds slot B at index IB */
    constexpr int KB = SH.KB, FB = SH.FB, SB = SH.B;
    int64_t Cb[KB];
    int32_t ndb[KB];
    auto none = [] {};
    frame_relayout<KD, KB, 0, SB, L>(g, lds, [&](int i) { return (uint32_t)C[i]; }, none,
                                     [&](int i, uint32_t w) { Cb[i] = (int64_t)w; });
    frame_relayout<KD, KB, 0, SB, L>(g, lds, [&](int i) { return (uint32_t)((uint64_t)C[i] >> 32); }, none,
                                     [&](int i, uint32_t w) { Cb[i] = (int64_t)(((uint64_t)w << 32) | (uint32_t)Cb[i]); });
    frame_relayout<KD, KB, 0, SB, L>(g, lds, [&](int i) { return (uint32_t)nd[i]; }, none,
                                     [&](int i, uint32_t w) { ndb[i] = (int32_t)w; });
    static_assert(FB <= SH.X + S - SH.JA, "the relayout stays below the quotient digits");
    int since = 0;
    auto renorm_before = [&](int n) __attribute__((always_inline)) {
      if constexpr (ND > 255) {
        if (since + n > 128) {
          lazy_renorm<KB, L>(g, Cb, top);
          since = 0;
        }
        since += n;
      }
    };
    constexpr int JE = SH.T3 ? SH.JB : S;   /* end of the second phase */
    for (; j0 + KB <= JE; j0 += KB) {
      renorm_before(KB);
      repeat<KB>([&](int u) __attribute__((always_inline)) {
        div_lazy_step<KB, L>(g, Cb, ndb, qd + (S - 1 - j0 - u), nz0, ftop, mtop, invs);
      });
    }
    renorm_before((JE - SH.JA) % KB);
    repeat<(JE - SH.JA) % KB>([&](int u) __attribute__((always_inline)) {
      div_lazy_step<KB, L>(g, Cb, ndb, qd + (S - 1 - j0 - u), nz0, ftop, mtop, invs);
    });
    j0 = JE;
    /* the end's estimate as in the whole frame; sure: that frame would keep the digits and
     * find a nonzero remainder */
    double t;
    if constexpr (SH.T3) {
      /* the third phase on the top FC slots, relaid out like the second */
      constexpr int KC = SH.KC, BC = SH.BC;
      static_assert(SH.FC <= SH.X + S - SH.JB, "the relayout stays below the quotient digits");
      int64_t Cc[KC];
      int32_t ndc[KC];
      frame_relayout<KB, KC, SB, BC, L>(g, lds, [&](int i) { return (uint32_t)Cb[i]; }, none,
                                        [&](int i, uint32_t w) { Cc[i] = (int64_t)w; });
      frame_relayout<KB, KC, SB, BC, L>(g, lds, [&](int i) { return (uint32_t)((uint64_t)Cb[i] >> 32); }, none,
                                        [&](int i, uint32_t w) { Cc[i] = (int64_t)(((uint64_t)w << 32) | (uint32_t)Cc[i]); });
      frame_relayout<KB, KC, SB, BC, L>(g, lds, [&](int i) { return (uint32_t)ndb[i]; }, none,
                                        [&](int i, uint32_t w) { ndc[i] = (int32_t)w; });
      auto renorm_c = [&](int n) __attribute__((always_inline)) {
        if constexpr (ND > 255) {
          if (since + n > 128) {
            lazy_renorm<KC, L>(g, Cc, top);
            since = 0;
          }
          since += n;
        }
      };
      for (; j0 + KC <= S; j0 += KC) {
        renorm_c(KC);
        repeat<KC>([&](int u) __attribute__((always_inline)) {
          div_lazy_step<KC, L>(g, Cc, ndc, qd + (S - 1 - j0 - u), nz0, ftop, mtop, invs);
        });
      }
      renorm_c((S - SH.JB) % KC);
      repeat<(S - SH.JB) % KC>([&](int u) __attribute__((always_inline)) {
        div_lazy_step<KC, L>(g, Cc, ndc, qd + (S - 1 - j0 - u), nz0, ftop, mtop, invs);
      });
      t = __builtin_fma(i64_to_f64(Cc[KC - 1]), (double)BETA, i64_to_f64(Cc[KC - 2])) * (invs * 0x1p-27);
    } else {
      t = __builtin_fma(i64_to_f64(Cb[KB - 1]), (double)BETA, i64_to_f64(Cb[KB - 2])) * (invs * 0x1p-27);
    }
    const bool unsure = !(t > 0x1p-16 && t < 1.0 - 0x1p-16);
    if (__builtin_expect(__ballot(top && unsure) != 0, 0)) return false;
  } else {
    repeat<S % KD>([&](int u) __attribute__((always_inline)) {
      div_lazy_step<KD, L>(g, C, nd, qd + (S - 1 - j0 - u), nz0, ftop, mtop, invs);
    });
  }
  uint32_t rem_nz = 1;
  int32_t delta = 0;
  if constexpr (!kShort) {
    /* final remainder R = sum C_k beta^k lies in (-D, 2D); its estimate t ~ R/D */
    const double t = __builtin_fma(i64_to_f64(C[KD - 1]), (double)BETA, i64_to_f64(C[KD - 2])) * (invs * 0x1p-27);
    const bool unsure = !(t > 0x1p-16 && t < 1.0 - 0x1p-16);
    if (__builtin_expect(__ballot(top && unsure) != 0, 0)) {
      /* exact: normalise R; add or subtract D once; the quotient moves by -1 / +1 */
      const int64_t T = normalize_exact<KD, L>(g, C);
      int64_t sgn = 0;                       /* +1: subtract D, -1: add D */
      if (T < 0) {
        sgn = -1;
      } else if (T > 0) {
        sgn = 1;
      } else {
        /* compare R with D: the highest lane with a differing digit decides */
        int c = 0;
#pragma unroll
        for (int i = KD - 1; i >= 0; i--)
          if (c == 0 && C[i] != (int64_t)-nd[i]) c = C[i] > (int64_t)-nd[i] ? 1 : -1;
        const uint64_t diff = g.ballot(c != 0);
        const int cmp = diff ? (int)g.bcast((uint32_t)c, 63 - __clzll(diff)) : 0;
        sgn = cmp >= 0 ? 1 : 0;
      }
      delta = (int32_t)sgn;
#pragma unroll
      for (int i = 0; i < KD; i++) C[i] += sgn * (int64_t)nd[i];
      normalize_exact<KD, L>(g, C);
      uint32_t nz = 0;
#pragma unroll
      for (int i = 0; i < KD; i++) nz |= (uint32_t)C[i];
      rem_nz = g.any(nz != 0) ? 1u : 0u;
    }
  }
  wsync();
  if (delta != 0 && g.gl == 0) qd[0] += delta;
  if (g.gl == 0) { qd[S] = (int32_t)topbit; qd[S + 1] = 0; qd[S + 2] = 0; }
  wsync();
  /* quotient digits outside [0, beta) (a corrected estimate): carry-normalise them */
  {
    bool bad = false;
    for (int k = g.gl; k < S; k += L) bad |= (uint32_t)qd[k] > MASK;
    if (__builtin_expect(__ballot(bad) != 0, 0)) {
      constexpr int CS = (S + L - 1) / L;
      int64_t X[CS];
#pragma unroll
      for (int i = 0; i < CS; i++) {
        const int k = g.gl * CS + i;
        X[i] = k < S ? (int64_t)qd[k] : 0;
      }
      normalize_exact<CS, L>(g, X);   /* the quotient is below beta^S: no carry out */
      wsync();
#pragma unroll
      for (int i = 0; i < CS; i++) {
        const int k = g.gl * CS + i;
        if (k < S) qd[k] = (int32_t)X[i];
      }
      wsync();
    }
  }
  /* F = topbit * beta^S + Q; the mantissa is its top p bits from the leading one */
  const uint32_t *qu = reinterpret_cast<const uint32_t *>(qd);
  const int base = topbit ? 27 * S - p + 1 : 27 * S - p;
  uint32_t Mw[K];
#pragma unroll
  for (int i = 0; i < K; i++) Mw[i] = digits_bits(qu, base + 32 * (g.gl * K + i));
  const uint32_t rw = digits_bits(qu, base - 32);
  const int lim = base - 32;   /* sticky bits [0, lim) lie in digit 0: lim <= 27 */
  const uint32_t sticky = rem_nz | ((lim > 0 && (qu[0] & ((1u << lim) - 1u)) != 0) ? 1u : 0u);
  wsync();
  if (topbit) E += 1;
  if (round_mantissa<K, L>(g, Mw, mode, sign, rw, sticky)) E += 1;
  re = pack_exp(E, sign, 10, 0x7FFFFFFEll);
  chunk_select(Qr, true, Mw, Mw);
  return true;
}

template <int K, int L>
__device__ __forceinline__ void op_div(const fpa_arguments &a, const Slot &s, const Group<L> &g, uint32_t *lds) {
  uint32_t A[K], B[K], R[K], re;
  load_value<K, L>(g, a.a_limbs, s, A);
  load_value<K, L>(g, a.b_limbs, s, B);
  if constexpr (div_short_shape(K, L).ok) {
    if (!core_div_lazy<K, L, true>(g, lds, (int)a.mode, a.a_exps[s.v], A, a.b_exps[s.v], B, re, R)) {
      /* some group of the wave was left unsure: every group of the wave but the special
       * classes (already done) runs the whole frame, from the operands reloaded (the clobber
       * keeps the compiler from holding the first loads' registers through the short run
       * instead) */
      asm volatile("" ::: "memory");
      load_value<K, L>(g, a.a_limbs, s, A);
      load_value<K, L>(g, a.b_limbs, s, B);
      core_div_lazy<K, L, false>(g, lds, (int)a.mode, a.a_exps[s.v], A, a.b_exps[s.v], B, re, R);
    }
  } else {
    core_div_lazy<K, L, false>(g, lds, (int)a.mode, a.a_exps[s.v], A, a.b_exps[s.v], B, re, R);
  }
  store_result(a, s, g.gl, re, R);
}

/* -------------------------------- sqrt -------------------------------- */
template <int K, int L, bool kShort>
__device__ __forceinline__ bool core_sqrt_lazy(const Group<L> &g, uint32_t *lds, int mode, uint32_t ea,
                                               const uint32_t (&A)[K], uint32_t &re, uint32_t (&Sr)[K]);

template <int K, int L>
__device__ __forceinline__ void op_sqrt(const fpa_arguments &a, const Slot &s, const Group<L> &g, uint32_t *lds) {
  uint32_t A[K], R[K], re;
  load_value<K, L>(g, a.a_limbs, s, A);
  if constexpr (sqrt_short_ok<K, L>()) {
    if (!core_sqrt_lazy<K, L, true>(g, lds, (int)a.mode, a.a_exps[s.v], A, re, R)) {
      /* the short frames left some group's root undecided: every group of the wave but the
       * special classes (already done) runs the full frame, from the operand reloaded (the
       * clobber keeps the compiler from holding the first load's registers through the
       * short run instead) */
      asm volatile("" ::: "memory");
      load_value<K, L>(g, a.a_limbs, s, A);
      core_sqrt_lazy<K, L, false>(g, lds, (int)a.mode, a.a_exps[s.v], A, re, R);
    }
  } else {
    core_sqrt_lazy<K, L, false>(g, lds, (int)a.mode, a.a_exps[s.v], A, re, R);
  }
  store_result(a, s, g.gl, re, R);
}

/* ------------------- sqrt with lazy radix-2^27 column sums ------------------- */
/*
 * Digit-by-digit square root in radix beta = 2^27 on div's lazy frame.  With F = KD*L root
 * digits and t = 27F - p >= 1 extra root bits, the radicand is M = A * 2^(54F - p - o)
 * (o = 1 for an odd exponent: the radicand is A/2) and Z = floor(sqrt(M)) has exactly 27F
 * bits: the mantissa is Z >> t, the round word the 32 bits below it, the sticky bit the
 * rest of Z and the remainder M - Z^2.  M's low F digits are zero, so the frame starts as
 * U = M / beta^F (A's digits, top-aligned) and only zeros enter it from below.  T = 2S sits
 * in the same F slots (root digit k at slot pos = F-1-k, kept negated in nT).  Step k:
 *   1. U <- U * beta: columns move up one slot, the top one folds into the new top column;
 *   2. the digit: k = 0: q = isqrt of the top column (exact, below 2^54); k = 1: T's second
 *      digit is not there yet and the q^2 term is 2^-27 of q*T, so q solves
 *      q^2 + 2 q0 beta q = V as V / (q0 beta + sqrt((q0 beta)^2 + V)) in FP64 (nothing
 *      cancels); k >= 2: q = V / (T_{F-1} beta + T_{F-2}) as in div (the q^2 term and T's
 *      lower digits are below 2^-53 of it), V from the top two columns;
 *   3. U -= q * T (one v_mad_i64_i32 per column), the owner of slot pos subtracts q^2 and
 *      sets T's slot to 2q.
 The short root (kShort) runs all F steps on the top (F+2)/2 slots, since no later digit
 * and not the end's estimate reads a slot below k - 1 after step k (long roots: on the top
 * quarter for the first and last quarter of the steps); where its end is unsure for some
 * group, the wave's groups run the full frame instead.  The full frame: in the first
 * HF = ceil(KD/2)*L steps T fits the frame's top HF slots and the slots below only move
 * radicand digits upwards, so those steps run on a half frame of the top HF slots
 * (ceil(KD/2) columns per lane), the radicand digits entering at its bottom one per step,
 * and the frame is then laid out whole again through LDS (about 3/4 of the multiply-adds
 * of F whole-frame steps).
 * The digits are redundant as in div and |q * 2q_j| < 2^57, so one carry level every 64
 * steps keeps the columns below 2^63.  At the end the estimate of R/T decides S unless it
 * lies within 2^-16 of 0 or 1; then R and T are normalised exactly and S moves by one
 * (R -> R + T - 1 or R - T - 1).  The root digits (-nT/2) go to LDS, are carry-normalised
 * only when one lies outside [0, beta), and the mantissa, round word and sticky bit are
 * read from them.  The reference's root (fpa_warp_sqrt.h:195-254) runs 32-bit digits
 * with exact remainders instead.
 */
/* One digit step of the lazy sqrt on a frame of KC columns per lane (the whole frame, or its
 * top part).  U <- U * beta (kFill: 'fill', nonzero only in the group's lane 0, enters the
 * bottom slot), digit q from the top two columns, U -= q * T, the owner lane's slot idx of
 * T (zero so far) becomes 2q (kOwn: the step's slot lies in the frame).  Returns q. */
template <int KC, int L, bool kFill, bool kOwn = true>
__device__ __forceinline__ int32_t sqrt_lazy_step(const Group<L> &g, int64_t (&C)[KC], int32_t (&nT)[KC], int idx,
                                                  uint32_t ownm, int kind, uint32_t ftop, uint32_t mtop, uint32_t nz0,
                                                  uint32_t fill, int32_t &q0, double &invs) {
  constexpr double BETA_D = 134217728.0;   /* 2^27 */
  /* U <- U * beta */
  lazy_shift_up<KC, L>(g, C, nz0, ftop, mtop);
  if constexpr (kFill) C[0] = (int64_t)((uint64_t)C[0] | fill);
  int32_t q;
  if (kind == 0) {
    q = (int32_t)fpa_isqrt64((uint64_t)C[KC - 1]);
  } else if (kind == 1) {
    const double Vs = __builtin_fma(i64_to_f64(C[KC - 1]), 0x1p-5, (double)(int32_t)((uint64_t)C[KC - 2] >> 32));
    const double a = (double)q0 * BETA_D, V = Vs * 0x1p32;
    q = (int32_t)__builtin_floor(V / (a + __builtin_sqrt(__builtin_fma(a, a, V))));
  } else {
    q = lazy_top_digit(C[KC - 1], C[KC - 2], invs);
  }
  q = (int32_t)g.bcast_top((uint32_t)q);
  /* slot idx is still zero in the owner lane (ownm all ones; the lanes above it hold
   * earlier digits there, and take 0): the owner sets it to -q for the multiply-adds, so
   * they subtract q^2 there too, then subtracts q again for -2q (no branch on the owner) */
  const int32_t qm = (int32_t)((uint32_t)q & ownm);
  if constexpr (kOwn) {
#pragma unroll
    for (int i = 0; i < KC; i++)
      if (i == idx) nT[i] -= qm;
  }
#pragma unroll
  for (int i = 0; i < KC; i++) C[i] = (int64_t)q * nT[i] + C[i];
  if constexpr (kOwn) {
#pragma unroll
    for (int i = 0; i < KC; i++)
      if (i == idx) nT[i] -= qm;
  }
  if (kind == 0) q0 = q;
  if (kind == 1) {
    const double tt = __builtin_fma((double)(2 * q0), BETA_D, (double)(2 * q));
    invs = 0x1p27 / tt;
  }
  /* keep the scheduler from running later steps' digit chains ahead of this step's
   * multiply-adds (that holds several steps' columns live and costs occupancy) */
  __builtin_amdgcn_sched_barrier(0);
  return q;
}

template <int K, int L, bool kShort>
__device__ __forceinline__ bool core_sqrt_lazy(const Group<L> &g, uint32_t *lds, int mode, uint32_t ea,
                                               const uint32_t (&A)[K], uint32_t &re, uint32_t (&Sr)[K]) {
  constexpr int n = K * L, p = 32 * n;
  constexpr int KD = div_digits_per_lane(K), F = KD * L, TX = DIV_RADIX * F - p;
  constexpr uint32_t MASK = (1u << DIV_RADIX) - 1;
  constexpr double BETA_D = 134217728.0;   /* 2^27 */
  static_assert(KD >= 2, "the top lane holds the two top columns");
  static_assert(TX >= 1, "at least one root bit below the mantissa");
  const uint32_t act = k_sqrt_tab[fpa_classify(ea)];
  if (act != FPA_A_SQRT) {
    re = act;
#pragma unroll
    for (int i = 0; i < K; i++) Sr[i] = 0;
    return true;
  }
  int64_t e = (int64_t)(ea >> 1) - FPA_BIAS_LL;
  const int o = (int)(e & 1);
  e += o;
  const int64_t er = e / 2;
  uint32_t ad[KD];
  if constexpr (recode_top_dpp_ok<K, L, KD>())
    recode_digits_top_dpp<K, L, KD>(g, A, ad, o);
  else
    recode_digits_top<K, L, KD>(g, lds, A, ad, o);
  constexpr int KH = (KD + 1) / 2, HF = KH * L, P2 = F - HF;
  constexpr bool kRen = F > 64;
  static_assert(KH >= 2, "the first trip holds the two special steps");
  const bool top = g.gl == L - 1;
  uint32_t nz0 = g.gl == 0 ? 0u : ~0u;     /* opaque: the column moves stay ANDs */
  uint32_t mtop = top ? ~0u : 0u, ftop = mtop & (1u << DIV_RADIX);
  asm volatile("" : "+v"(nz0), "+v"(mtop), "+v"(ftop));
  double invs = 0.0;                       /* 2^27 / (T_{F-1} beta + T_{F-2}), from step 2 on */
  int32_t q0 = 0;
  /* |q * 2q_j| < 2^57: one carry level at most every 64 steps keeps the columns below 2^63;
   * it runs before a trip of n steps that would take the count past 64 */
  int since = 0;
  auto renorm_before = [&](auto &Cf, int n) __attribute__((always_inline)) {
    if constexpr (kRen) {
      if (since + n > 64) {
        lazy_renorm(g, Cf, top);
        since = 0;
      }
      since += n;
    }
  };

  int32_t *sd = reinterpret_cast<int32_t *>(lds);
  uint32_t rem_nz = 1;
  if constexpr (kShort) {
    /* Short frame: no step's digit and no end decision reads a slot below s = k - 1 after
     * step k (the digits come from the top two columns, the end's estimate too, and slot s
     * reaches slot F-2 only through F-2-s shifts), so those slots need not be kept.  The
     * frame holds the top FB = KB*L >= (F+2)/2 slots, B0 = F - FB below it, for all F steps:
     * until step F-1-B0 T fits in it and the slots below it only carry radicand digits
     * upwards (they enter at its bottom lane, as in the half-frame phase of the full frame);
     * after that the products and root digits below it land where nothing is read again,
     * since 2*B0 <= F-2.  The digits are those of the full frame, and where the end's
     * estimate is sure so is the result (S unchanged, a nonzero remainder); the root digits
     * of the slots below the frame go to LDS as they are made.  Where some group of the
     * wave is unsure, false: the caller runs the full frame.
     * LDS: [0, ZB) zeros, [ZB, ZB+F) the radicand's digits, [ZB+F, ZB+F+KB) zeros; the
     * root digits at [0, F) once the digits have entered. */
    constexpr int KB = sqrt_short_cols(KD, L), FB = KB * L, B0 = F - FB, ZB = 2 * L + 2;
    static_assert(sqrt_short_ok<K, L>(), "short frame shape");
    /* Where KB and L are even and the frame is long (F >= 300, or F >= 150 with lanes of at
     * most 24 limbs: for shorter frames, and for 32-limb lanes of 152 slots, the two more
     * relayouts cost more than they save), the first and last steps run on frames of KB/2
     * columns per lane (the top FB/2 slots): steps 0 .. FB/2-1 on the top FB/2 slots (T fits them, the
     * slots below only carry radicand digits), and from step K2 = B3 + 3 on the top FB/2 slots
     * again (B3 = F - FB/2: no later digit reads a slot below k - 1).  Between the frames the
     * slots pass through LDS: first into dig[B0, F), whose radicand digits have all entered the
     * frame by then, with the digits still to enter below the first frame taken from
     * dig[B0 - FB/2, B0); then into the root digits' LDS words [B3, F), which are written
     * last. */
    constexpr bool kQ = KB % 2 == 0 && L % 2 == 0 && KB / 2 >= 2 && (F >= 300 || (F >= 150 && K <= 24));
    constexpr int KQ = kQ ? KB / 2 : KB, FQ = KQ * L, B1 = F - FQ, B3 = F - FQ, K1 = kQ ? FQ : 0;
    constexpr int K2 = kQ ? B3 + 3 : F;
    static_assert(!kQ || (K1 % KB == 0 && K2 >= FB && 2 * FQ < F + 4), "quarter frames");
    uint32_t *dig = lds + ZB;
    wsync();   /* the recoding reads of lds are done before the digits overwrite them */
    for (int w = g.gl; w < ZB; w += L) lds[w] = 0;
#pragma unroll
    for (int i = 0; i < KD; i++) dig[g.gl * KD + i] = ad[i];
    for (int w = g.gl; w < KB; w += L) dig[F + w] = 0;
    wsync();
    int64_t Ch[KB];
    int32_t nTh[KB];
    {
      /* digit B-1-k enters the bottom lane of a frame with bottom slot B at step k (zeros once
       * k >= B; the lowest read, B0 - FB >= -2L-1, is in [0, ZB)); the other lanes read zeros */
      uint32_t fl[KB];
      auto fills = [&](int tr, auto kc, int b) __attribute__((always_inline)) {
        constexpr int KC = decltype(kc)::value;
        const uint32_t *src = g.gl == 0 ? dig + (b - KC - tr * KC) : dig + F;
#pragma unroll
        for (int u = 0; u < KC; u++) fl[u] = src[KC - 1 - u];
      };
      int tr0 = 0;
      if constexpr (kQ) {
        int64_t Cq[KQ];
        int32_t nTq[KQ];
#pragma unroll
        for (int i = 0; i < KQ; i++) { Cq[i] = (int64_t)dig[B1 + g.gl * KQ + i]; nTq[i] = 0; }
        using KQc = std::integral_constant<int, KQ>;
        fills(0, KQc{}, B1);
        uint32_t ownm = top ? ~0u : 0u;
        asm volatile("" : "+v"(ownm));
        renorm_before(Cq, KQ);
        repeat<KQ>([&](int u) __attribute__((always_inline)) {
          sqrt_lazy_step<KQ, L, true>(g, Cq, nTq, KQ - 1 - u, ownm, u == 0 ? 0 : u == 1 ? 1 : 2, ftop, mtop, nz0, fl[u],
                                      q0, invs);
        });
        for (int tr = 1; tr < L; tr++) {
          fills(tr, KQc{}, B1);
          ownm = g.gl == L - 1 - tr ? ~0u : 0u;
          asm volatile("" : "+v"(ownm));
          renorm_before(Cq, KQ);
          repeat<KQ>([&](int u) __attribute__((always_inline)) {
            sqrt_lazy_step<KQ, L, true>(g, Cq, nTq, KQ - 1 - u, ownm, 2, ftop, mtop, nz0, fl[u], q0, invs);
          });
        }
        /* into the frame of the top FB slots: below B1 the radicand digits not yet entered
         * (slot s holds the digit first at slot s - K1) */
        uint32_t *R = dig + B0;
        frame_relayout<KQ, KB, B1, B0, L>(
            g, R, [&](int i) { return (uint32_t)Cq[i]; },
            [&] {
              for (int w = g.gl; w < B1 - B0; w += L) R[w] = dig[B0 - K1 + w];
            },
            [&](int i, uint32_t w) { Ch[i] = (int64_t)w; });
        frame_relayout<KQ, KB, B1, B0, L>(
            g, R, [&](int i) { return (uint32_t)((uint64_t)Cq[i] >> 32); },
            [&] {
              for (int w = g.gl; w < B1 - B0; w += L) R[w] = 0;
            },
            [&](int i, uint32_t w) { Ch[i] = (int64_t)(((uint64_t)w << 32) | (uint32_t)Ch[i]); });
        frame_relayout<KQ, KB, B1, B0, L>(
            g, R, [&](int i) { return (uint32_t)nTq[i]; },
            [&] {
              for (int w = g.gl; w < B1 - B0; w += L) R[w] = 0;
            },
            [&](int i, uint32_t w) { nTh[i] = (int32_t)w; });
        tr0 = K1 / KB;
      } else {
#pragma unroll
        for (int i = 0; i < KB; i++) { Ch[i] = (int64_t)dig[B0 + g.gl * KB + i]; nTh[i] = 0; }
      }
      using KBc = std::integral_constant<int, KB>;
      if (tr0 == 0) {
        fills(0, KBc{}, B0);
        uint32_t ownm = top ? ~0u : 0u;
        asm volatile("" : "+v"(ownm));
        renorm_before(Ch, KB);
        repeat<KB>([&](int u) __attribute__((always_inline)) {
          sqrt_lazy_step<KB, L, true>(g, Ch, nTh, KB - 1 - u, ownm, u == 0 ? 0 : u == 1 ? 1 : 2, ftop, mtop, nz0, fl[u],
                                      q0, invs);
        });
        tr0 = 1;
      }
      for (int tr = tr0; tr < L; tr++) {
        fills(tr, KBc{}, B0);
        uint32_t ownm = g.gl == L - 1 - tr ? ~0u : 0u;
        asm volatile("" : "+v"(ownm));
        renorm_before(Ch, KB);
        repeat<KB>([&](int u) __attribute__((always_inline)) {
          sqrt_lazy_step<KB, L, true>(g, Ch, nTh, KB - 1 - u, ownm, 2, ftop, mtop, nz0, fl[u], q0, invs);
        });
      }
    }
    /* steps FB .. K2-1: slot F-1-k lies below the frame; its digit goes to LDS */
    int k0 = FB;
    for (; k0 + KB <= K2; k0 += KB) {
      renorm_before(Ch, KB);
      repeat<KB>([&](int u) __attribute__((always_inline)) {
        const int32_t q = sqrt_lazy_step<KB, L, false, false>(g, Ch, nTh, 0, 0u, 2, ftop, mtop, nz0, 0u, q0, invs);
        sd[F - 1 - k0 - u] = q;
      });
    }
    renorm_before(Ch, (K2 - FB) % KB);
    repeat<(K2 - FB) % KB>([&](int u) __attribute__((always_inline)) {
      const int32_t q = sqrt_lazy_step<KB, L, false, false>(g, Ch, nTh, 0, 0u, 2, ftop, mtop, nz0, 0u, q0, invs);
      sd[F - 1 - k0 - u] = q;
    });
    if constexpr (kQ) {
      /* steps K2 .. F-1 on the top FQ slots: the root digits of slots B0 .. B3-1 go to LDS
       * first, then the frame moves through the root digits' words [B3, F) */
      int64_t Cq[KQ];
      int32_t nTq[KQ];
      {
        wsync();
        constexpr int LQ = (B3 - B0) / KB, IQ = (B3 - B0) - LQ * KB;
        if (g.gl < LQ) {
#pragma unroll
          for (int i = 0; i < KB; i++) sd[B0 + g.gl * KB + i] = -nTh[i] / 2;
        } else if (g.gl == LQ) {
#pragma unroll
          for (int i = 0; i < IQ; i++) sd[B0 + g.gl * KB + i] = -nTh[i] / 2;
        }
        uint32_t *R = reinterpret_cast<uint32_t *>(sd) + B3;
        auto none = [] {};
        frame_relayout<KB, KQ, B0, B3, L>(g, R, [&](int i) { return (uint32_t)Ch[i]; }, none,
                                          [&](int i, uint32_t w) { Cq[i] = (int64_t)w; });
        frame_relayout<KB, KQ, B0, B3, L>(g, R, [&](int i) { return (uint32_t)((uint64_t)Ch[i] >> 32); }, none,
                                          [&](int i, uint32_t w) { Cq[i] = (int64_t)(((uint64_t)w << 32) | (uint32_t)Cq[i]); });
        frame_relayout<KB, KQ, B0, B3, L>(g, R, [&](int i) { return (uint32_t)nTh[i]; }, none,
                                          [&](int i, uint32_t w) { nTq[i] = (int32_t)w; });
      }
      int k3 = K2;
      for (; k3 + KQ <= F; k3 += KQ) {
        renorm_before(Cq, KQ);
        repeat<KQ>([&](int u) __attribute__((always_inline)) {
          const int32_t q = sqrt_lazy_step<KQ, L, false, false>(g, Cq, nTq, 0, 0u, 2, ftop, mtop, nz0, 0u, q0, invs);
          sd[F - 1 - k3 - u] = q;
        });
      }
      renorm_before(Cq, (F - K2) % KQ);
      repeat<(F - K2) % KQ>([&](int u) __attribute__((always_inline)) {
        const int32_t q = sqrt_lazy_step<KQ, L, false, false>(g, Cq, nTq, 0, 0u, 2, ftop, mtop, nz0, 0u, q0, invs);
        sd[F - 1 - k3 - u] = q;
      });
      const double r = __builtin_fma(i64_to_f64(Cq[KQ - 1]), BETA_D, i64_to_f64(Cq[KQ - 2])) * invs * 0x1p-27;
      const bool unsure = !(r > 0x1p-16 && r < 1.0 - 0x1p-16);
      if (__builtin_expect(__ballot(top && unsure) != 0, 0)) return false;
      wsync();
#pragma unroll
      for (int i = 0; i < KQ; i++) sd[B3 + g.gl * KQ + i] = -nTq[i] / 2;
      if (top) { sd[F] = 0; sd[F + 1] = 0; sd[F + 2] = 0; }
      wsync();
    } else {
      const double r = __builtin_fma(i64_to_f64(Ch[KB - 1]), BETA_D, i64_to_f64(Ch[KB - 2])) * invs * 0x1p-27;
      const bool unsure = !(r > 0x1p-16 && r < 1.0 - 0x1p-16);
      if (__builtin_expect(__ballot(top && unsure) != 0, 0)) return false;
      wsync();
#pragma unroll
      for (int i = 0; i < KB; i++) sd[B0 + g.gl * KB + i] = -nTh[i] / 2;
      if (top) { sd[F] = 0; sd[F + 1] = 0; sd[F + 2] = 0; }
      wsync();
    }
  } else {
    /* Phase 1, steps 0 .. HF-1: T fits the frame's top HF slots, and U's slots below them only
     * carry radicand digits upwards, one per step.  These steps run on a half frame of the top
     * HF = KH*L slots, KH columns per lane, the digits below entering at its bottom lane.
     * LDS: [0, L) zeros, [L, L+F) the radicand's digits, [L+F, L+F+KH) zeros. */
    uint32_t *dig = lds + L;
    wsync();   /* the recoding reads of lds are done before the digits overwrite them */
    lds[g.gl] = 0;
  #pragma unroll
    for (int i = 0; i < KD; i++) dig[g.gl * KD + i] = ad[i];
  #pragma unroll
    for (int q = 0; q < (KH + L - 1) / L; q++)
      if (q * L + g.gl < KH) dig[F + q * L + g.gl] = 0;
    wsync();
    int64_t Ch[KH];
    int32_t nTh[KH];
  #pragma unroll
    for (int i = 0; i < KH; i++) { Ch[i] = (int64_t)dig[P2 + g.gl * KH + i]; nTh[i] = 0; }
    {
      /* digit P2-1-k enters the bottom lane at step k (zeros once k >= P2; they come from
       * [0, L)); the other lanes read zeros */
      uint32_t fl[KH];
      auto fills = [&](int tr) __attribute__((always_inline)) {
        const uint32_t *b = g.gl == 0 ? dig + (P2 - KH - tr * KH) : dig + F;
  #pragma unroll
        for (int u = 0; u < KH; u++) fl[u] = b[KH - 1 - u];
      };
      /* trips of KH steps: step k's half-frame slot HF-1-k is index KH-1-u of lane L-1-trip;
       * ownm: all ones in the lane that owns the trip's slots, opaque so that it is formed
       * once per trip rather than again at every step */
      fills(0);
      uint32_t ownm = top ? ~0u : 0u;
      asm volatile("" : "+v"(ownm));
      renorm_before(Ch, KH);
      repeat<KH>([&](int u) __attribute__((always_inline)) {
        sqrt_lazy_step<KH, L, true>(g, Ch, nTh, KH - 1 - u, ownm, u == 0 ? 0 : u == 1 ? 1 : 2, ftop, mtop, nz0, fl[u],
                                    q0, invs);
      });
      for (int tr = 1; tr < L; tr++) {
        fills(tr);
        ownm = g.gl == L - 1 - tr ? ~0u : 0u;
        asm volatile("" : "+v"(ownm));
        renorm_before(Ch, KH);
        repeat<KH>([&](int u) __attribute__((always_inline)) {
          sqrt_lazy_step<KH, L, true>(g, Ch, nTh, KH - 1 - u, ownm, 2, ftop, mtop, nz0, fl[u], q0, invs);
        });
      }
    }
    /* the half frame's slots P2 .. F-1 into the whole frame (KD columns per lane, zeros below
     * P2), one 32-bit word per pass through LDS */
    int64_t C[KD];
    int32_t nT[KD];
    {
      wsync();
  #pragma unroll
      for (int i = 0; i < KH; i++) lds[P2 + g.gl * KH + i] = (uint32_t)Ch[i];
  #pragma unroll
      for (int q = 0; q < (P2 + L - 1) / L; q++)
        if (q * L + g.gl < P2) lds[q * L + g.gl] = 0;
      wsync();
      uint32_t lo[KD];
  #pragma unroll
      for (int i = 0; i < KD; i++) lo[i] = lds[g.gl * KD + i];
      wsync();
  #pragma unroll
      for (int i = 0; i < KH; i++) lds[P2 + g.gl * KH + i] = (uint32_t)((uint64_t)Ch[i] >> 32);
      wsync();
  #pragma unroll
      for (int i = 0; i < KD; i++) C[i] = (int64_t)(((uint64_t)lds[g.gl * KD + i] << 32) | lo[i]);
      wsync();
  #pragma unroll
      for (int i = 0; i < KH; i++) lds[P2 + g.gl * KH + i] = (uint32_t)nTh[i];
      wsync();
  #pragma unroll
      for (int i = 0; i < KD; i++) nT[i] = (int32_t)lds[g.gl * KD + i];
    }
    /* Phase 2, steps HF .. F-1 on the whole frame: slot F-1-k runs from P2-1 down to 0; the
     * first trip covers the RT slots of lane LT, then whole trips of KD for lanes LT-1 .. 0 */
    {
      constexpr int LT = (P2 - 1) / KD, RT = P2 - LT * KD;
      uint32_t ownm = g.gl == LT ? ~0u : 0u;
      asm volatile("" : "+v"(ownm));
      renorm_before(C, RT);
      repeat<RT>([&](int u) __attribute__((always_inline)) {
        sqrt_lazy_step<KD, L, false>(g, C, nT, RT - 1 - u, ownm, 2, ftop, mtop, nz0, 0u, q0, invs);
      });
      for (int tr = LT - 1; tr >= 0; tr--) {
        ownm = g.gl == tr ? ~0u : 0u;
        asm volatile("" : "+v"(ownm));
        renorm_before(C, KD);
        repeat<KD>([&](int u) __attribute__((always_inline)) {
          sqrt_lazy_step<KD, L, false>(g, C, nT, KD - 1 - u, ownm, 2, ftop, mtop, nz0, 0u, q0, invs);
        });
      }
    }

    /* R = sum C_k beta^k should lie in [0, T]; its estimate r ~ R / T */
    const double r = __builtin_fma(i64_to_f64(C[KD - 1]), BETA_D, i64_to_f64(C[KD - 2])) * invs * 0x1p-27;
    const bool unsure = !(r > 0x1p-16 && r < 1.0 - 0x1p-16);
    int32_t delta = 0;
    if (__builtin_expect(__ballot(top && unsure) != 0, 0)) {
      int64_t Tn[KD];
  #pragma unroll
      for (int i = 0; i < KD; i++) Tn[i] = -(int64_t)nT[i];
      const int64_t Rtop = normalize_exact<KD, L>(g, C);
      const int64_t Ttop = normalize_exact<KD, L>(g, Tn);
      int cmp;                               /* sign of R - T */
      if (Rtop != Ttop) {
        cmp = Rtop > Ttop ? 1 : -1;
      } else {
        int c = 0;
  #pragma unroll
        for (int i = KD - 1; i >= 0; i--)
          if (c == 0 && C[i] != Tn[i]) c = C[i] > Tn[i] ? 1 : -1;
        const uint64_t diff = g.ballot(c != 0);
        cmp = diff ? (int)g.bcast((uint32_t)c, 63 - __clzll(diff)) : 0;
      }
      /* S += delta: R -> R - delta * T - 1 for delta = +-1 */
      delta = Rtop < 0 ? -1 : cmp > 0 ? 1 : 0;
      int64_t Rhi = Rtop;
      if (delta != 0) {
  #pragma unroll
        for (int i = 0; i < KD; i++) C[i] -= (int64_t)delta * Tn[i];
        if (g.gl == 0) C[0] -= 1;
        Rhi = normalize_exact<KD, L>(g, C);
      }
      /* R lies in [0, 2S], which reaches past beta^F: its part above the F digits counts */
      uint32_t nz = 0;
  #pragma unroll
      for (int i = 0; i < KD; i++) nz |= (uint32_t)C[i];
      rem_nz = (Rhi != 0 || g.any(nz != 0)) ? 1u : 0u;
    }
    /* root digits S_k = -nT/2 at slot k, then three zero words */
    wsync();
  #pragma unroll
    for (int i = 0; i < KD; i++) sd[g.gl * KD + i] = -nT[i] / 2 + (g.gl == 0 && i == 0 ? delta : 0);
    if (top) { sd[F] = 0; sd[F + 1] = 0; sd[F + 2] = 0; }
    wsync();
  }
  {
    bool bad = false;
#pragma unroll
    for (int i = 0; i < KD; i++) bad |= (uint32_t)sd[g.gl * KD + i] > MASK;
    if (__builtin_expect(__ballot(bad) != 0, 0)) {
      int64_t X[KD];
#pragma unroll
      for (int i = 0; i < KD; i++) X[i] = (int64_t)sd[g.gl * KD + i];
      normalize_exact<KD, L>(g, X);   /* S is below beta^F: no carry out */
      wsync();
#pragma unroll
      for (int i = 0; i < KD; i++) sd[g.gl * KD + i] = (int32_t)X[i];
      wsync();
    }
  }
  /* Z has 27F bits, its top bit set: mantissa = bits [TX, TX + p), round word below */
  const uint32_t *su = reinterpret_cast<const uint32_t *>(sd);
  uint32_t Mw[K];
#pragma unroll
  for (int i = 0; i < K; i++) Mw[i] = digits_bits(su, TX + 32 * (g.gl * K + i));
  uint32_t rw, sticky = rem_nz;
  if constexpr (TX >= 32) {
    rw = digits_bits(su, TX - 32);
    /* sticky bits [0, TX - 32): whole digits below digit D0, part of digit D0 */
    constexpr int LIM = TX - 32, D0 = LIM / DIV_RADIX, REM = LIM % DIV_RADIX;
    uint32_t loc = 0;
    for (int k = g.gl; k < D0; k += L) loc |= su[k];
    if (REM > 0 && g.gl == 0) loc |= su[D0] & ((1u << REM) - 1u);
    sticky |= g.any(loc != 0) ? 1u : 0u;
  } else {
    rw = digits_bits(su, 0) << (32 - TX);
  }
  wsync();
  int64_t E = er + FPA_BIAS_LL;
  if (round_mantissa<K, L>(g, Mw, mode, 0u, rw, sticky)) E += 1;
  re = pack_exp(E, 0u, 0, 0x7FFFFFFFll);
  chunk_select(Sr, true, Mw, Mw);
  return true;
}

}  // namespace fpa

#endif
