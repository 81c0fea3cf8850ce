/*
 * ORACLE -- test infrastructure only.
 *
 * Exact-integer CPU restatement of cuBigFloat's element-wise operations
 * (neg/add/sub/mul/div/sqrt) on the fpa number format.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this code,
 * and only as the checker; the product (cubigfloat_amd/) never calls it.
 *
 * Number format (fpa_internal.h:1-17, tests.c:114-159):
 *   exponent word  <= 8 : class code (2 ZERO, 3 NEG_TINY, 4 POS_TINY, 5 NEG_INF,
 *                          6 POS_INF, 7 NAN, 8 SNAN)
 *   exponent word  >= 10: bit 0 = sign, E = word>>1, value = (-1)^s * M * 2^(E-BIAS-p),
 *                          M = p-bit integer with its top bit set, p = 32*n,
 *                          little-endian 32-bit limbs.
 *
 * Every function computes the exact result with big-integer arithmetic and rounds
 * it once, so finite results equal MPFR's correctly rounded ones.  Where the
 * reference's GPU code departs from correct rounding or leaves an action
 * unhandled, the oracle follows MPFR (the reference's own gold, tests.c:756-918):
 *   - sub with a==0: the reference table says NEG_B (fpa_warp_add_sub.h:32-34) but its
 *     switch has no NEG_B case (fpa_warp_add_sub.h:357-379); the oracle returns -b.
 *   - div special classes that MPFR can represent (values, zero, +-inf, nan) follow
 *     MPFR: x/0 and inf/0 -> +-INF (reference table: NAN, fpa_warp_div.h:3-9),
 *     inf/x -> INF with the quotient's sign (the reference's inf rows,
 *     fpa_warp_div.h:8-9, carry the opposite sign), x/inf -> ZERO (reference: TINY).
 *   - mul and sub-diff "critical" all-ones paths (fpa_warp_mul.h:99-113,
 *     fpa_warp_add_sub.h:187-200) round at the unshifted position; the oracle rounds
 *     the normalised value (e.g. 0x8000..0 * 0xFFFF..F is exact).
 *   - sqrt round-up past all-ones (directed modes) raises the exponent by one binade;
 *     the reference adds 1 to the exponent word (fpa_warp_sqrt.h:238-243).
 *
 * Range rules restated from the reference's exponent checks:
 *   add/sub/mul : E < 5 -> +-TINY, E > 0x7FFFFFFF -> +-INF
 *                 (fpa_warp_add_sub.h:118, 226, 272; fpa_warp_mul.h:125-128)
 *   div         : E < 10 -> +-TINY, E >= 0x7FFFFFFF -> +-INF (fpa_warp_div.h:190-193)
 *   sqrt        : no range check (fpa_warp_sqrt.h:247-250)
 * Special-value action tables are restated as data from fpa_warp_add_sub.h:3-39,
 * fpa_warp_mul.h:1-12, fpa_warp_div.h:1-12, fpa_warp_sqrt.h:1-11.
 * Rounding decision restated from fpa_round (fpa_internal.h:62-88).
 *
 * set/sgn/cmp (fpa.h:57-61): sgn and cmp follow the reference's class tables
 * (fpa_threaded.h:1-24; a table "error" entry yields 0, fpa_threaded.h:40-43, 69-70) and
 * compare two same-sign values by exponent word, then mantissa (fpa_threaded.h:60-88).
 * set (precision conversion) is unfinished in the reference (fpa_warp_set.h:134-199);
 * its gold is mpfr_set (tests.c:580-605): the source rounded once to the target
 * precision, special classes copied, and the add/sub/mul range rule on overflow.
 * set_ui/si/float/double (fpa_warp_set.h:1-132): exact.  An integer's significand is the
 * integer shifted left by its leading zeros, exponent word 2*(BIAS + bit length) + sign;
 * a float or double goes through frexp to f in [1/2, 1) and ldexp(f, 64), as the
 * reference kernels do (frexp, ldexp by 24 or 53), with NaN -> NAN, +-inf -> +-INF and
 * zero of either sign -> ZERO.  Gold: mpfr_set_ui/si/flt/d (tests.c:487-578).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_BIAS 0x3FFFFFF3ll

enum { C_POSV = 0, C_NEGV = 1, C_ZERO = 2, C_NTINY = 3, C_PTINY = 4, C_NINF = 5, C_PINF = 6, C_NAN = 7, C_SNAN = 8 };
enum { A_MUL = 16, A_DIV = 17, A_SQRT = 18, A_NEG = 19, A_ADD = 20, A_SUB = 21, A_COPY_A = 22, A_COPY_B = 23, A_NEG_B = 24 };
enum { M_RZ = 0, M_RU = 1, M_RD = 2, M_RNE = 3, M_RNA = 4 };

/* ---- action tables (data restated from the reference, see header) ---- */
static const uint8_t neg_tab[9] = { A_NEG, A_NEG, C_ZERO, C_PTINY, C_NTINY, C_PINF, C_NINF, C_NAN, C_SNAN };

#define V C_POSV
static const uint8_t add_tab[81] = {
  A_ADD,    A_SUB,    A_COPY_A, A_COPY_A, A_COPY_A, C_NINF, C_PINF, C_NAN, C_SNAN,
  A_SUB,    A_ADD,    A_COPY_A, A_COPY_A, A_COPY_A, C_NINF, C_PINF, C_NAN, C_SNAN,
  A_COPY_B, A_COPY_B, C_ZERO,   C_NTINY,  C_PTINY,  C_NINF, C_PINF, C_NAN, C_SNAN,
  A_COPY_B, A_COPY_B, C_NTINY,  C_NTINY,  C_NAN,    C_NINF, C_PINF, C_NAN, C_SNAN,
  A_COPY_B, A_COPY_B, C_PTINY,  C_NAN,    C_PTINY,  C_NINF, C_PINF, C_NAN, C_SNAN,
  C_NINF,   C_NINF,   C_NINF,   C_NINF,   C_NINF,   C_NINF, C_NAN,  C_NAN, C_SNAN,
  C_PINF,   C_PINF,   C_PINF,   C_PINF,   C_PINF,   C_NAN,  C_PINF, C_NAN, C_SNAN,
  C_NAN,    C_NAN,    C_NAN,    C_NAN,    C_NAN,    C_NAN,  C_NAN,  C_NAN, C_SNAN,
  C_SNAN,   C_SNAN,   C_SNAN,   C_SNAN,   C_SNAN,   C_SNAN, C_SNAN, C_SNAN, C_SNAN,
};
static const uint8_t sub_tab[81] = {
  A_SUB,    A_ADD,    A_COPY_A, A_COPY_A, A_COPY_A, C_PINF, C_NINF, C_NAN, C_SNAN,
  A_ADD,    A_SUB,    A_COPY_A, A_COPY_A, A_COPY_A, C_PINF, C_NINF, C_NAN, C_SNAN,
  A_NEG_B,  A_NEG_B,  C_ZERO,   C_PTINY,  C_NTINY,  C_PINF, C_NINF, C_NAN, C_SNAN,
  A_NEG_B,  A_NEG_B,  C_NTINY,  C_NAN,    C_NINF,   C_PINF, C_NINF, C_NAN, C_SNAN,
  A_NEG_B,  A_NEG_B,  C_PTINY,  C_PTINY,  C_NAN,    C_PINF, C_NINF, C_NAN, C_SNAN,
  C_NINF,   C_NINF,   C_NINF,   C_NINF,   C_NINF,   C_NAN,  C_NINF, C_NAN, C_SNAN,
  C_PINF,   C_PINF,   C_PINF,   C_PINF,   C_PINF,   C_PINF, C_NAN,  C_NAN, C_SNAN,
  C_NAN,    C_NAN,    C_NAN,    C_NAN,    C_NAN,    C_NAN,  C_NAN,  C_NAN, C_SNAN,
  C_SNAN,   C_SNAN,   C_SNAN,   C_SNAN,   C_SNAN,   C_SNAN, C_SNAN, C_SNAN, C_SNAN,
};
static const uint8_t mul_tab[81] = {
  A_MUL,   A_MUL,   C_ZERO, C_NTINY, C_PTINY, C_NINF, C_PINF, C_NAN, C_SNAN,
  A_MUL,   A_MUL,   C_ZERO, C_PTINY, C_NTINY, C_PINF, C_NINF, C_NAN, C_SNAN,
  C_ZERO,  C_ZERO,  C_ZERO, C_ZERO,  C_ZERO,  C_NAN,  C_NAN,  C_NAN, C_SNAN,
  C_NTINY, C_PTINY, C_ZERO, C_PTINY, C_NTINY, C_NAN,  C_NAN,  C_NAN, C_SNAN,
  C_PTINY, C_NTINY, C_ZERO, C_NTINY, C_PTINY, C_NAN,  C_NAN,  C_NAN, C_SNAN,
  C_NINF,  C_PINF,  C_NAN,  C_NAN,   C_NAN,   C_PINF, C_NINF, C_NAN, C_SNAN,
  C_PINF,  C_NINF,  C_NAN,  C_NAN,   C_NAN,   C_NINF, C_PINF, C_NAN, C_SNAN,
  C_NAN,   C_NAN,   C_NAN,  C_NAN,   C_NAN,   C_NAN,  C_NAN,  C_NAN, C_SNAN,
  C_SNAN,  C_SNAN,  C_SNAN, C_SNAN,  C_SNAN,  C_SNAN, C_SNAN, C_SNAN, C_SNAN,
};
/* Rows/columns MPFR can represent (values, zero, +-inf, nan) follow MPFR:
 * x/0 -> +-INF (reference: NAN), inf/0 -> +-INF (reference: NAN), inf/x has the sign
 * of the quotient (the reference's inf rows have it flipped), x/inf -> ZERO (MPFR +-0;
 * reference: TINY).  These are 12 entries.  Entries with a TINY or SNAN operand are the
 * reference's, e.g. value/TINY -> INF with the quotient's sign (fpa_warp_div.h:4-5). */
static const uint8_t div_tab[81] = {
  A_DIV,   A_DIV,   C_PINF, C_NINF,  C_PINF,  C_ZERO,  C_ZERO,  C_NAN, C_SNAN,
  A_DIV,   A_DIV,   C_NINF, C_PINF,  C_NINF,  C_ZERO,  C_ZERO,  C_NAN, C_SNAN,
  C_ZERO,  C_ZERO,  C_NAN,  C_ZERO,  C_ZERO,  C_ZERO,  C_ZERO,  C_NAN, C_SNAN,
  C_NTINY, C_PTINY, C_NAN,  C_NAN,   C_NAN,   C_PTINY, C_NTINY, C_NAN, C_SNAN,
  C_PTINY, C_NTINY, C_NAN,  C_NAN,   C_NAN,   C_NTINY, C_PTINY, C_NAN, C_SNAN,
  C_NINF,  C_PINF,  C_NINF, C_PINF,  C_NINF,  C_NAN,   C_NAN,   C_NAN, C_SNAN,
  C_PINF,  C_NINF,  C_PINF, C_NINF,  C_PINF,  C_NAN,   C_NAN,   C_NAN, C_SNAN,
  C_NAN,   C_NAN,   C_NAN,  C_NAN,   C_NAN,   C_NAN,   C_NAN,   C_NAN, C_SNAN,
  C_SNAN,  C_SNAN,  C_SNAN, C_SNAN,  C_SNAN,  C_SNAN,  C_SNAN,  C_SNAN, C_SNAN,
};
#undef V
static const uint8_t sqrt_tab[9] = { A_SQRT, C_NAN, C_ZERO, C_NAN, C_PTINY, C_NAN, C_PINF, C_NAN, C_SNAN };

static uint32_t classify(uint32_t w) { return w <= C_SNAN ? w : (w & 1u); }

/* expose the tables to the tests (restated data is checked against the reference text) */
int orc_action(int op, uint32_t a_exp, uint32_t b_exp) {
  uint32_t ca = classify(a_exp), cb = classify(b_exp);
  switch (op) {
    case 0: return neg_tab[ca];
    case 1: return add_tab[ca * 9 + cb];
    case 2: return sub_tab[ca * 9 + cb];
    case 3: return mul_tab[ca * 9 + cb];
    case 4: return div_tab[ca * 9 + cb];
    case 5: return sqrt_tab[ca];
  }
  return -1;
}

/* ---------------------------- big integers ---------------------------- */
static int clz32(uint32_t x) { return x ? __builtin_clz(x) : 32; }

static int bn_len(const uint32_t *x, int n) { while (n > 0 && x[n - 1] == 0) n--; return n; }

static int64_t bn_bitlen(const uint32_t *x, int n) {
  n = bn_len(x, n);
  if (n == 0) return 0;
  return (int64_t)32 * n - clz32(x[n - 1]);
}

static int bn_bit(const uint32_t *x, int n, int64_t i) {
  if (i < 0 || i >= (int64_t)32 * n) return 0;
  return (x[i >> 5] >> (i & 31)) & 1;
}

/* any bit of x below position i set? */
static int bn_any_below(const uint32_t *x, int n, int64_t i) {
  if (i <= 0) return 0;
  if (i >= (int64_t)32 * n) return bn_len(x, n) != 0;
  int64_t w = i >> 5;
  for (int64_t k = 0; k < w; k++) if (x[k]) return 1;
  int b = (int)(i & 31);
  return b && (x[w] & ((1u << b) - 1)) != 0;
}

/* r[0..nr) = x >> s (s >= 0), bits above x are zero */
static void bn_shr(uint32_t *r, int nr, const uint32_t *x, int nx, int64_t s) {
  for (int i = 0; i < nr; i++) {
    int64_t bit = (int64_t)i * 32 + s;
    int64_t w = bit >> 5;
    int b = (int)(bit & 31);
    uint32_t lo = (w < nx) ? x[w] : 0, hi = (w + 1 < nx) ? x[w + 1] : 0;
    r[i] = b ? ((lo >> b) | (hi << (32 - b))) : lo;
  }
}

/* r[0..nr) = x << s (s >= 0), truncated to nr words */
static void bn_shl(uint32_t *r, int nr, const uint32_t *x, int nx, int64_t s) {
  for (int i = 0; i < nr; i++) {
    int64_t bit = (int64_t)i * 32 - s;
    int64_t w = bit >= 0 ? (bit >> 5) : -((-bit + 31) >> 5);
    int b = (int)(bit - w * 32);
    uint32_t lo = (w >= 0 && w < nx) ? x[w] : 0, hi = (w + 1 >= 0 && w + 1 < nx) ? x[w + 1] : 0;
    r[i] = b ? ((lo >> b) | (hi << (32 - b))) : lo;
  }
}

static int bn_cmp(const uint32_t *a, const uint32_t *b, int n) {
  for (int i = n - 1; i >= 0; i--) if (a[i] != b[i]) return a[i] > b[i] ? 1 : -1;
  return 0;
}

/* r = a + b, returns carry */
static uint32_t bn_add(uint32_t *r, const uint32_t *a, const uint32_t *b, int n) {
  uint64_t c = 0;
  for (int i = 0; i < n; i++) { c += (uint64_t)a[i] + b[i]; r[i] = (uint32_t)c; c >>= 32; }
  return (uint32_t)c;
}

/* r = a - b, returns borrow */
static uint32_t bn_sub(uint32_t *r, const uint32_t *a, const uint32_t *b, int n) {
  int64_t c = 0;
  for (int i = 0; i < n; i++) { c += (int64_t)a[i] - b[i]; r[i] = (uint32_t)c; c >>= 32; }
  return (uint32_t)(c != 0);
}

/* r[0..na+nb) = a * b */
static void bn_mul(uint32_t *r, const uint32_t *a, int na, const uint32_t *b, int nb) {
  memset(r, 0, sizeof(uint32_t) * (na + nb));
  for (int i = 0; i < na; i++) {
    uint64_t c = 0;
    for (int j = 0; j < nb; j++) {
      uint64_t t = (uint64_t)a[i] * b[j] + r[i + j] + c;
      r[i + j] = (uint32_t)t;
      c = t >> 32;
    }
    r[i + nb] = (uint32_t)c;
  }
}

/*
 * Knuth algorithm D: q[0..nu-nv] = floor(u / v), rem[0..nv) = u mod v.
 * nu >= nv >= 1, v[nv-1] != 0.
 */
static void bn_divmod(uint32_t *q, uint32_t *rem, const uint32_t *u, int nu, const uint32_t *v, int nv) {
  if (nv == 1) {
    uint64_t r = 0;
    for (int i = nu - 1; i >= 0; i--) {
      uint64_t cur = (r << 32) | u[i];
      q[i] = (uint32_t)(cur / v[0]);
      r = cur % v[0];
    }
    rem[0] = (uint32_t)r;
    return;
  }
  int s = clz32(v[nv - 1]);
  uint32_t *vn = (uint32_t *)malloc(sizeof(uint32_t) * nv);
  uint32_t *un = (uint32_t *)malloc(sizeof(uint32_t) * (nu + 1));
  for (int i = nv - 1; i > 0; i--) vn[i] = (v[i] << s) | (s ? v[i - 1] >> (32 - s) : 0);
  vn[0] = v[0] << s;
  un[nu] = s ? u[nu - 1] >> (32 - s) : 0;
  for (int i = nu - 1; i > 0; i--) un[i] = (u[i] << s) | (s ? u[i - 1] >> (32 - s) : 0);
  un[0] = u[0] << s;
  for (int j = nu - nv; j >= 0; j--) {
    uint64_t num = ((uint64_t)un[j + nv] << 32) | un[j + nv - 1];
    uint64_t qh = num / vn[nv - 1], rh = num % vn[nv - 1];
    while (qh >> 32 || qh * vn[nv - 2] > ((rh << 32) | un[j + nv - 2])) {
      qh--;
      rh += vn[nv - 1];
      if (rh >> 32) break;
    }
    /* un[j..j+nv] -= qh * vn */
    int64_t borrow = 0;
    uint64_t carry = 0;
    for (int i = 0; i < nv; i++) {
      uint64_t p = qh * vn[i] + carry;
      carry = p >> 32;
      int64_t t = (int64_t)un[i + j] - (int64_t)(uint32_t)p + borrow;
      un[i + j] = (uint32_t)t;
      borrow = t >> 32;
    }
    int64_t t = (int64_t)un[j + nv] - (int64_t)carry + borrow;
    un[j + nv] = (uint32_t)t;
    if (t < 0) { /* add back */
      qh--;
      uint64_t c = 0;
      for (int i = 0; i < nv; i++) { c += (uint64_t)un[i + j] + vn[i]; un[i + j] = (uint32_t)c; c >>= 32; }
      un[j + nv] += (uint32_t)c;
    }
    q[j] = (uint32_t)qh;
  }
  for (int i = 0; i < nv; i++) rem[i] = (un[i] >> s) | (s ? un[i + 1] << (32 - s) : 0);
  free(vn);
  free(un);
}

/* r[0..nx/2+1) = floor(sqrt(x)), returns 1 if x is not a perfect square. Newton from above. */
static int bn_isqrt(uint32_t *r, const uint32_t *x, int nx) {
  int nr = nx / 2 + 1;
  int64_t bl = bn_bitlen(x, nx);
  uint32_t *cur = (uint32_t *)calloc(nr + 1, 4), *nxt = (uint32_t *)calloc(nr + 1, 4);
  uint32_t *qq = (uint32_t *)calloc(nx + 1, 4), *rr = (uint32_t *)calloc(nr + 2, 4);
  if (bl == 0) { memset(r, 0, 4 * nr); free(cur); free(nxt); free(qq); free(rr); return 0; }
  int64_t e = (bl + 1) / 2;  /* 2^e >= sqrt(x) */
  cur[e >> 5] |= 1u << (e & 31);
  for (;;) {
    int lc = bn_len(cur, nr + 1);
    /* nxt = (cur + x / cur) / 2 */
    memset(qq, 0, 4 * (nx + 1));
    bn_divmod(qq, rr, x, nx, cur, lc);
    uint32_t *sum = (uint32_t *)calloc(nx + 2, 4);
    uint32_t *cc = (uint32_t *)calloc(nx + 2, 4);
    memcpy(cc, cur, 4 * lc);
    uint32_t c = bn_add(sum, qq, cc, nx + 1);
    sum[nx + 1] = c;
    memset(nxt, 0, 4 * (nr + 1));
    bn_shr(nxt, nr + 1, sum, nx + 2, 1);
    free(sum); free(cc);
    if (bn_cmp(nxt, cur, nr + 1) >= 0) break;
    memcpy(cur, nxt, 4 * (nr + 1));
  }
  memcpy(r, cur, 4 * nr);
  /* inexact iff cur*cur != x */
  uint32_t *sq = (uint32_t *)calloc(2 * (nr + 1), 4);
  bn_mul(sq, cur, nr, cur, nr);
  int inexact = 0;
  for (int i = 0; i < 2 * nr; i++) {
    uint32_t xi = i < nx ? x[i] : 0;
    if (sq[i] != xi) { inexact = 1; break; }
  }
  free(sq); free(cur); free(nxt); free(qq); free(rr);
  return inexact;
}

/* ------------------------------ rounding ------------------------------ */

/* fpa_round (fpa_internal.h:62-88) expressed with the first dropped bit and
 * the OR of the rest. Returns 1 to increment the magnitude. */
int orc_round_decision(int mode, int sign, int lsb, int round_bit, int sticky) {
  switch (mode) {
    case M_RZ: return 0;
    case M_RU: return !sign && (round_bit || sticky);
    case M_RD: return sign && (round_bit || sticky);
    case M_RNA: return round_bit;
    case M_RNE: return round_bit && (sticky || lsb);
  }
  return 0;
}

/*
 * Round |value| = (X + frac) * 2^scale, 0 <= frac < 1 with frac > 0 iff sticky_in,
 * to p = 32n bits.  X must be nonzero.  Writes the mantissa, returns the MPFR-style
 * exponent e (value = 0.M * 2^e).
 */
static int64_t round_to(uint32_t *m, int n, const uint32_t *X, int nx, int64_t scale, int sticky_in, int sign, int mode) {
  int64_t p = (int64_t)32 * n;
  int64_t L = bn_bitlen(X, nx);
  int64_t e = L + scale;
  int rbit, sticky;
  if (L > p) {
    int64_t s = L - p;
    bn_shr(m, n, X, nx, s);
    rbit = bn_bit(X, nx, s - 1);
    sticky = bn_any_below(X, nx, s - 1) || sticky_in;
  } else {
    bn_shl(m, n, X, nx, p - L);
    rbit = 0;
    sticky = sticky_in;
  }
  if (orc_round_decision(mode, sign, m[0] & 1, rbit, sticky)) {
    uint64_t c = 1;
    for (int i = 0; i < n; i++) { c += m[i]; m[i] = (uint32_t)c; c >>= 32; }
    if (c) { /* carried out of the top: 2^p -> 2^(p-1), one binade up */
      memset(m, 0, 4 * n);
      m[n - 1] = 0x80000000u;
      e++;
    }
  }
  return e;
}

/* Map (sign, e) to an exponent word with the op's range rule. */
static uint32_t pack_exp(int64_t e, int sign, int64_t emin_E, int64_t emax_E) {
  int64_t E = e + ORC_BIAS;
  if (E < emin_E) return (uint32_t)(C_PTINY - sign);
  if (E > emax_E) return (uint32_t)(C_PINF - sign);
  return (uint32_t)(2 * E + sign);
}

static void set_special(uint32_t *r_exp, uint32_t *r_limbs, int n, uint32_t code) {
  *r_exp = code;
  memset(r_limbs, 0, 4 * n);
}

/* -------------------------------- ops --------------------------------- */

void orc_neg(uint32_t *r_exp, uint32_t *r_limbs, uint32_t a_exp, const uint32_t *a, int n) {
  int act = neg_tab[classify(a_exp)];
  if (act == A_NEG) {
    memmove(r_limbs, a, 4 * n);
    *r_exp = a_exp ^ 1u;
  } else set_special(r_exp, r_limbs, n, act);
}

/* magnitude add/sub with signs already resolved: result sign 'sa' for |a| >= |b| logic below */
static void addsub_values(uint32_t *r_exp, uint32_t *r_limbs, uint32_t a_exp, const uint32_t *a,
                          uint32_t b_exp, const uint32_t *b, int n, int mode) {
  int64_t p = 32 * (int64_t)n;
  int sa = a_exp & 1, sb = b_exp & 1;
  int64_t Ea = a_exp >> 1, Eb = b_exp >> 1;
  /* scale of the integer mantissa: value = M * 2^(E - BIAS - p) */
  const uint32_t *X = a, *Y = b;
  int64_t Ex = Ea, Ey = Eb;
  int sx = sa, sy = sb;
  if (Ea < Eb || (Ea == Eb && bn_cmp(a, b, n) < 0)) { X = b; Y = a; Ex = Eb; Ey = Ea; sx = sb; sy = sa; }
  int64_t d = Ex - Ey;
  int64_t dc = d > p + 64 ? p + 64 : d;       /* clamping keeps every rounding outcome */
  int nw = n + (int)((dc + 31) / 32) + 2;
  uint32_t *Xs = (uint32_t *)calloc(nw, 4), *Ys = (uint32_t *)calloc(nw, 4), *T = (uint32_t *)calloc(nw, 4);
  bn_shl(Xs, nw, X, n, dc);
  memcpy(Ys, Y, 4 * n);
  int64_t scale = Ey + (d - dc) - ORC_BIAS - p;  /* = Ex - dc - BIAS - p */
  int negative_result = sx;
  if (sx == sy) bn_add(T, Xs, Ys, nw);
  else {
    bn_sub(T, Xs, Ys, nw);  /* |X| >= |Y| by construction */
    if (bn_len(T, nw) == 0) { set_special(r_exp, r_limbs, n, C_ZERO); goto out; }
  }
  {
    int64_t e = round_to(r_limbs, n, T, nw, scale, 0, negative_result, mode);
    *r_exp = pack_exp(e, negative_result, 5, 0x7FFFFFFFll);
    if (*r_exp <= C_SNAN) memset(r_limbs, 0, 4 * n);
  }
out:
  free(Xs); free(Ys); free(T);
}

void orc_addsub(uint32_t *r_exp, uint32_t *r_limbs, uint32_t a_exp, const uint32_t *a,
                uint32_t b_exp, const uint32_t *b, int n, int mode, int is_sub) {
  uint32_t ca = classify(a_exp), cb = classify(b_exp);
  int act = is_sub ? sub_tab[ca * 9 + cb] : add_tab[ca * 9 + cb];
  uint32_t *tmp = (uint32_t *)malloc(4 * n);
  switch (act) {
    case A_ADD: case A_SUB:
      addsub_values(r_exp, tmp, a_exp, a, is_sub ? (b_exp ^ 1u) : b_exp, b, n, mode);
      memcpy(r_limbs, tmp, 4 * n);
      break;
    case A_COPY_A: memmove(r_limbs, a, 4 * n); *r_exp = a_exp; break;
    case A_COPY_B: memmove(r_limbs, b, 4 * n); *r_exp = b_exp; break;
    case A_NEG_B: memmove(r_limbs, b, 4 * n); *r_exp = b_exp ^ 1u; break;
    default: set_special(r_exp, r_limbs, n, act);
  }
  free(tmp);
}

void orc_mul(uint32_t *r_exp, uint32_t *r_limbs, uint32_t a_exp, const uint32_t *a,
             uint32_t b_exp, const uint32_t *b, int n, int mode) {
  int act = mul_tab[classify(a_exp) * 9 + classify(b_exp)];
  if (act != A_MUL) { set_special(r_exp, r_limbs, n, act); return; }
  int64_t p = 32 * (int64_t)n;
  int sign = (a_exp ^ b_exp) & 1;
  uint32_t *P = (uint32_t *)malloc(8 * n);
  bn_mul(P, a, n, b, n);
  int64_t scale = ((int64_t)(a_exp >> 1) - ORC_BIAS - p) + ((int64_t)(b_exp >> 1) - ORC_BIAS - p);
  int64_t e = round_to(r_limbs, n, P, 2 * n, scale, 0, sign, mode);
  *r_exp = pack_exp(e, sign, 5, 0x7FFFFFFFll);
  if (*r_exp <= C_SNAN) memset(r_limbs, 0, 4 * n);
  free(P);
}

void orc_div(uint32_t *r_exp, uint32_t *r_limbs, uint32_t a_exp, const uint32_t *a,
             uint32_t b_exp, const uint32_t *b, int n, int mode) {
  int act = div_tab[classify(a_exp) * 9 + classify(b_exp)];
  if (act != A_DIV) { set_special(r_exp, r_limbs, n, act); return; }
  int64_t p = 32 * (int64_t)n;
  int sign = (a_exp ^ b_exp) & 1;
  /* Q = floor(A * 2^(p+64) / B), sticky = remainder != 0 */
  int nu = 2 * n + 2;
  uint32_t *U = (uint32_t *)calloc(nu, 4), *Q = (uint32_t *)calloc(nu, 4), *R = (uint32_t *)calloc(n, 4);
  memcpy(U + n + 2, a, 4 * n);
  bn_divmod(Q, R, U, nu, b, bn_len(b, n));
  int sticky = bn_len(R, n) != 0;
  int64_t scale = ((int64_t)(a_exp >> 1) - ORC_BIAS - p) - ((int64_t)(b_exp >> 1) - ORC_BIAS - p) - (p + 64);
  int64_t e = round_to(r_limbs, n, Q, nu, scale, sticky, sign, mode);
  *r_exp = pack_exp(e, sign, 10, 0x7FFFFFFEll);
  if (*r_exp <= C_SNAN) memset(r_limbs, 0, 4 * n);
  free(U); free(Q); free(R);
}

void orc_sqrt(uint32_t *r_exp, uint32_t *r_limbs, uint32_t a_exp, const uint32_t *a, int n, int mode) {
  int act = sqrt_tab[classify(a_exp)];
  if (act != A_SQRT) { set_special(r_exp, r_limbs, n, act); return; }
  int64_t p = 32 * (int64_t)n;
  /* value = A * 2^(E - BIAS - p).  Make the scaled exponent even: sqrt(A * 2^(2k)) * 2^t. */
  int64_t sc = (int64_t)(a_exp >> 1) - ORC_BIAS - p;   /* value = A * 2^sc */
  int64_t extra = 2 * p + 64;                          /* root gets >= p+32 bits */
  if ((sc - extra) & 1) extra++;                       /* sc - extra even */
  int nx = (int)((p + extra) / 32) + 2;
  uint32_t *X = (uint32_t *)calloc(nx, 4), *S = (uint32_t *)calloc(nx / 2 + 1, 4);
  bn_shl(X, nx, a, n, extra);
  int inexact = bn_isqrt(S, X, nx);
  int64_t scale = (sc - extra) / 2;
  int64_t e = round_to(r_limbs, n, S, nx / 2 + 1, scale, inexact, 0, mode);
  /* no range check in the reference's sqrt: |e| shrinks */
  *r_exp = (uint32_t)(2 * (e + ORC_BIAS));
  free(X); free(S);
}

/* ------------------------- array entry points ------------------------- */
/* op: 0 neg, 1 add, 2 sub, 3 mul, 4 div, 5 sqrt.  Natural layout, n = precision/32. */
void orc_op_array(int op, uint32_t count, uint32_t precision, int mode,
                  uint32_t *r_exps, uint32_t *r_limbs,
                  const uint32_t *a_exps, const uint32_t *a_limbs,
                  const uint32_t *b_exps, const uint32_t *b_limbs) {
  int n = (int)(precision / 32);
  for (uint32_t i = 0; i < count; i++) {
    uint32_t *rl = r_limbs + (size_t)i * n;
    const uint32_t *al = a_limbs + (size_t)i * n;
    const uint32_t *bl = b_limbs ? b_limbs + (size_t)i * n : NULL;
    switch (op) {
      case 0: orc_neg(&r_exps[i], rl, a_exps[i], al, n); break;
      case 1: orc_addsub(&r_exps[i], rl, a_exps[i], al, b_exps[i], bl, n, mode, 0); break;
      case 2: orc_addsub(&r_exps[i], rl, a_exps[i], al, b_exps[i], bl, n, mode, 1); break;
      case 3: orc_mul(&r_exps[i], rl, a_exps[i], al, b_exps[i], bl, n, mode); break;
      case 4: orc_div(&r_exps[i], rl, a_exps[i], al, b_exps[i], bl, n, mode); break;
      case 5: orc_sqrt(&r_exps[i], rl, a_exps[i], al, n, mode); break;
    }
  }
}

/* big-integer self-test helpers exported for brute-force checks in tests */
void orc_bn_divmod(uint32_t *q, uint32_t *rem, const uint32_t *u, int nu, const uint32_t *v, int nv) {
  bn_divmod(q, rem, u, nu, v, nv);
}
int orc_bn_isqrt(uint32_t *r, const uint32_t *x, int nx) { return bn_isqrt(r, x, nx); }
void orc_bn_mul(uint32_t *r, const uint32_t *a, int na, const uint32_t *b, int nb) { bn_mul(r, a, na, b, nb); }

/* ------------------------- set / sgn / cmp ------------------------- */
#define T_CMP 2
#define T_ERR 3
static const int8_t sgn_tab[9] = { +1, -1, 0, -1, +1, -1, +1, T_ERR, T_ERR };
static const int8_t cmp_tab[81] = {
  /* a=+value */ T_CMP, +1, +1, +1, +1, +1, -1, T_ERR, T_ERR,
  /*   -value */ -1, T_CMP, -1, -1, -1, +1, -1, T_ERR, T_ERR,
  /*     zero */ -1, +1, 0, +1, -1, +1, -1, T_ERR, T_ERR,
  /*   -1/inf */ -1, +1, -1, T_ERR, -1, +1, -1, T_ERR, T_ERR,
  /*   +1/inf */ -1, +1, +1, +1, T_ERR, +1, -1, T_ERR, T_ERR,
  /*     -inf */ -1, -1, -1, -1, -1, T_ERR, -1, T_ERR, T_ERR,
  /*     +inf */ +1, +1, +1, +1, +1, +1, T_ERR, T_ERR, T_ERR,
  /*      nan */ T_ERR, T_ERR, T_ERR, T_ERR, T_ERR, T_ERR, T_ERR, T_ERR, T_ERR,
  /*     snan */ T_ERR, T_ERR, T_ERR, T_ERR, T_ERR, T_ERR, T_ERR, T_ERR, T_ERR,
};

/* the tables, for the tests (entry codes: -1, 0, +1, 2 = compare the values, 3 = error) */
int orc_sgn_table(uint32_t cls) { return cls < 9 ? sgn_tab[cls] : -9; }
int orc_cmp_table(uint32_t ca, uint32_t cb) { return ca < 9 && cb < 9 ? cmp_tab[ca * 9 + cb] : -9; }

void orc_sgn_array(uint32_t count, const uint32_t *x_exps, int32_t *r) {
  for (uint32_t i = 0; i < count; i++) {
    int t = sgn_tab[classify(x_exps[i])];
    r[i] = t == T_ERR ? 0 : t;
  }
}

void orc_cmp_array(uint32_t count, uint32_t precision, const uint32_t *a_exps, const uint32_t *a_limbs,
                   const uint32_t *b_exps, const uint32_t *b_limbs, int32_t *r) {
  int n = (int)(precision / 32);
  for (uint32_t i = 0; i < count; i++) {
    uint32_t ea = a_exps[i], eb = b_exps[i];
    int t = cmp_tab[classify(ea) * 9 + classify(eb)];
    if (t == T_CMP) {
      int sgn = (ea & 1u) ? -1 : 1;
      if (ea != eb) t = ea > eb ? sgn : -sgn;
      else t = sgn * bn_cmp(a_limbs + (size_t)i * n, b_limbs + (size_t)i * n, n);
    }
    r[i] = t == T_ERR ? 0 : t;
  }
}

/* r (precision pr) = x (precision px) rounded in `mode` */
void orc_set_array(uint32_t count, uint32_t r_precision, uint32_t x_precision, int mode, uint32_t *r_exps,
                   uint32_t *r_limbs, const uint32_t *x_exps, const uint32_t *x_limbs) {
  int nr = (int)(r_precision / 32), nx = (int)(x_precision / 32);
  for (uint32_t i = 0; i < count; i++) {
    uint32_t ex = x_exps[i];
    uint32_t *rl = r_limbs + (size_t)i * nr;
    if (ex <= C_SNAN) { set_special(&r_exps[i], rl, nr, ex); continue; }
    int sign = (int)(ex & 1u);
    int64_t scale = (int64_t)(ex >> 1) - ORC_BIAS - 32ll * nx;
    int64_t e = round_to(rl, nr, x_limbs + (size_t)i * nx, nx, scale, 0, sign, mode);
    r_exps[i] = pack_exp(e, sign, 5, 0x7FFFFFFFll);
    if (r_exps[i] <= C_SNAN) memset(rl, 0, 4 * (size_t)nr);
  }
}

/* r = the fpa_set_* conversion of values (kind 0 uint64, 1 int64, 2 float, 3 double) */
void orc_set_scalar_array(int kind, uint32_t count, uint32_t precision, const void *values, uint32_t *r_exps,
                          uint32_t *r_limbs) {
  int n = (int)(precision / 32);
  for (uint32_t i = 0; i < count; i++) {
    uint32_t *rl = r_limbs + (size_t)i * n;
    memset(rl, 0, 4 * (size_t)n);
    uint64_t mag;
    int sign;
    int64_t e;
    if (kind <= 1) {
      uint64_t u = ((const uint64_t *)values)[i];
      sign = kind == 1 && (int64_t)u < 0;
      mag = sign ? ~u + 1u : u;
      if (mag == 0) { r_exps[i] = C_ZERO; continue; }
      int bl = 64 - __builtin_clzll(mag);
      e = bl;
      mag <<= 64 - bl;
    } else {
      double x = kind == 2 ? (double)((const float *)values)[i] : ((const double *)values)[i];
      if (isnan(x)) { r_exps[i] = C_NAN; continue; }
      if (isinf(x)) { r_exps[i] = x < 0 ? C_NINF : C_PINF; continue; }
      if (x == 0) { r_exps[i] = C_ZERO; continue; }
      int ex;
      double f = frexp(fabs(x), &ex);           /* |x| = f * 2^ex, f in [1/2, 1) */
      mag = (uint64_t)ldexp(f, 64);             /* exact: f has at most 53 significant bits */
      sign = x < 0;
      e = ex;
    }
    rl[n - 1] = (uint32_t)(mag >> 32);
    rl[n - 2] = (uint32_t)mag;
    r_exps[i] = (uint32_t)(2 * (ORC_BIAS + e) + sign);
  }
}

/* Dot chains: r_j = fold over k of acc <- RN(acc + RN(x_k * y_k)) from an exact zero,
 * the same per-element op sequence MPFR gold uses (mpfr_mul then mpfr_add). */
void orc_dot_array(uint32_t count, uint32_t precision, int mode, const uint32_t *lengths,
                   const uint32_t *x_exps, const uint32_t *x_limbs, const uint32_t *y_exps,
                   const uint32_t *y_limbs, uint32_t *r_exps, uint32_t *r_limbs) {
  int n = (int)(precision / 32);
  uint32_t *p = (uint32_t *)malloc(4 * (size_t)n), *acc = (uint32_t *)malloc(4 * (size_t)n);
  uint64_t k = 0;
  for (uint32_t j = 0; j < count; j++) {
    uint32_t acc_e = C_ZERO, pe;
    memset(acc, 0, 4 * (size_t)n);
    for (uint32_t t = 0; t < lengths[j]; t++, k++) {
      orc_mul(&pe, p, x_exps[k], x_limbs + k * n, y_exps[k], y_limbs + k * n, n, mode);
      orc_addsub(&acc_e, acc, acc_e, acc, pe, p, n, mode, 0);
    }
    r_exps[j] = acc_e;
    memcpy(r_limbs + (size_t)j * n, acc, 4 * (size_t)n);
    if (acc_e <= C_SNAN) memset(r_limbs + (size_t)j * n, 0, 4 * (size_t)n);
  }
  free(p);
  free(acc);
}
