/*
 * Scalar 32/64-bit helpers shared by the HIP kernels and the host unit tests.
 *
 *   fpa_round_inc   -- rounding decision of fpa_round (reference fpa_internal.h:62-88)
 *   fpa_recip_3by2  -- v = floor((2^96-1)/(d1*2^32+d0)) - 2^32 for a normalised
 *                      two-word divisor (Moller & Granlund, "Improved division by
 *                      invariant integers", IEEE TC 2011, eq. for the 3/2 reciprocal)
 *   fpa_div3by2     -- floor((u2,u1,u0)/(d1,d0)) for (u2,u1) < (d1,d0), the same
 *                      paper's Algorithm 5, used as the quotient-digit estimate of the
 *                      long division and square root (the reference instead iterates
 *                      DIV32/APPROX32, fpa_math.h:1-50)
 *   fpa_isqrt64     -- floor(sqrt(u)) for u < 2^64 (first root digit; the reference's
 *                      SQRT64, fpa_math.h:68-110, computes the same value bitwise)
 *   fpa_encode_scalar -- exact fpa encoding of a 64-bit integer or an IEEE binary32 /
 *                      binary64 value (fpa_set_ui/si/float/double, reference
 *                      fpa_warp_set.h:1-132)
 *
 * Compiled by hipcc for the device and by gcc/g++ for the host tests.
 */
#ifndef FPA_SCALAR_H
#define FPA_SCALAR_H

#include <stdint.h>
#include <math.h>

#if defined(__HIPCC__)
#define FPA_HD __host__ __device__ __forceinline__
#else
#define FPA_HD static inline
#endif

#define FPA_BIAS_LL 0x3FFFFFF3ll

/* exponent-word class codes (values <= 8) and operation actions */
#define FPA_C_POS_VALUE 0u
#define FPA_C_NEG_VALUE 1u
#define FPA_C_ZERO      2u
#define FPA_C_NEG_TINY  3u
#define FPA_C_POS_TINY  4u
#define FPA_C_NEG_INF   5u
#define FPA_C_POS_INF   6u
#define FPA_C_NAN       7u
#define FPA_C_SNAN      8u

#define FPA_A_MUL    16u
#define FPA_A_DIV    17u
#define FPA_A_SQRT   18u
#define FPA_A_NEG    19u
#define FPA_A_ADD    20u
#define FPA_A_SUB    21u
#define FPA_A_COPY_A 22u
#define FPA_A_COPY_B 23u
#define FPA_A_NEG_B  24u

FPA_HD uint32_t fpa_classify(uint32_t w) { return w <= FPA_C_SNAN ? w : (w & 1u); }

/* Returns 1 when the magnitude must be incremented.  round_word = the 32 bits just
 * below the kept mantissa, sticky = nonzero iff any lower bit is set. */
FPA_HD uint32_t fpa_round_inc(int mode, uint32_t sign, uint32_t lsb, uint32_t round_word, uint32_t sticky) {
  uint32_t inexact = (round_word | sticky) != 0;
  switch (mode) {
    case 0: return 0;                                   /* truncate */
    case 1: return sign ? 0u : inexact;                 /* toward +inf */
    case 2: return sign ? inexact : 0u;                 /* toward -inf */
    case 4: return round_word >= 0x80000000u;           /* ties away from zero */
    default:                                            /* ties to even */
      if (round_word != 0x80000000u) return round_word > 0x80000000u;
      return (sticky != 0) | (lsb & 1u);
  }
}

FPA_HD uint32_t fpa_recip_3by2(uint32_t d1, uint32_t d0) {
  uint64_t d = ((uint64_t)d1 << 32) | d0;      /* d >= 2^63 */
  uint64_t rem = 0x7FFFFFFFFFFFFFFFull;        /* the top 63 one-bits of 2^96-1, below d */
  uint64_t q = 0;
  for (int i = 0; i < 33; i++) {
    uint64_t top = rem >> 63;
    rem = (rem << 1) | 1u;
    q <<= 1;
    if (top || rem >= d) { rem -= d; q |= 1u; }
  }
  return (uint32_t)(q - (1ull << 32));
}

FPA_HD uint32_t fpa_div3by2(uint32_t u2, uint32_t u1, uint32_t u0, uint32_t d1, uint32_t d0, uint32_t v) {
  uint64_t q = (uint64_t)v * u2 + (((uint64_t)u2 << 32) | u1);
  uint32_t q1 = (uint32_t)(q >> 32), q0 = (uint32_t)q;
  uint32_t r1 = u1 - q1 * d1;
  uint64_t d = ((uint64_t)d1 << 32) | d0;
  uint64_t r = (((uint64_t)r1 << 32) | u0) - (uint64_t)d0 * q1 - d;
  q1 += 1u;
  if ((uint32_t)(r >> 32) >= q0) { q1 -= 1u; r += d; }
  if (r >= d) { q1 += 1u; }
  return q1;
}

FPA_HD uint32_t fpa_isqrt64(uint64_t u) {
  uint64_t r = (uint64_t)sqrt((double)u);
  if (r > 0xFFFFFFFFull) r = 0xFFFFFFFFull;
  while (r * r > u) r--;
  while (r < 0xFFFFFFFFull && (r + 1) * (r + 1) <= u) r++;
  return (uint32_t)r;
}

/* Source kinds of fpa_encode_scalar (the four fpa_set_* conversions) */
#define FPA_SRC_U64 0
#define FPA_SRC_S64 1
#define FPA_SRC_F32 2
#define FPA_SRC_F64 3

/* Exact fpa encoding of one source value: returns the exponent word, or a class code for
 * zero (either sign), infinities and NaN.  For a value, *top receives its magnitude's
 * significand left-aligned (bit 63 set), so the mantissa is top * 2^(p-64): the top two
 * limbs, all others zero.  The value is |x| = top * 2^(E - BIAS - 64) with E the word's
 * biased exponent.  bits holds the integer, or the IEEE bit pattern (binary32 in the low
 * 32 bits); subnormal floats are exact too.  The reference forms the same words with
 * frexp/ldexp and a count of leading zeros (fpa_warp_set.h:1-132). */
FPA_HD uint32_t fpa_encode_scalar(int kind, uint64_t bits, uint64_t *top) {
  uint32_t sign = 0;
  uint64_t m;
  int64_t s = 0;                               /* |x| = m * 2^s */
  *top = 0;
  if (kind == FPA_SRC_U64) {
    m = bits;
  } else if (kind == FPA_SRC_S64) {
    sign = (uint32_t)(bits >> 63);
    m = sign ? (uint64_t)0 - bits : bits;      /* INT64_MIN -> 2^63 */
  } else if (kind == FPA_SRC_F32) {
    const uint32_t f = (uint32_t)bits, e = (f >> 23) & 0xFFu;
    sign = f >> 31;
    m = f & 0x7FFFFFu;
    if (e == 0xFFu) return m ? FPA_C_NAN : FPA_C_POS_INF - sign;
    if (e) { m |= 1u << 23; s = (int64_t)e - 150; }
    else s = -149;
  } else {
    const uint32_t e = (uint32_t)(bits >> 52) & 0x7FFu;
    sign = (uint32_t)(bits >> 63);
    m = bits & ((1ull << 52) - 1u);
    if (e == 0x7FFu) return m ? FPA_C_NAN : FPA_C_POS_INF - sign;
    if (e) { m |= 1ull << 52; s = (int64_t)e - 1075; }
    else s = -1074;
  }
  if (m == 0) return FPA_C_ZERO;
  const int lz = __builtin_clzll(m);
  *top = m << lz;
  return (uint32_t)(2 * (FPA_BIAS_LL + (64 - lz) + s) + sign);
}

#endif
