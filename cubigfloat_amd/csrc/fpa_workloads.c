/*
 * Deterministic synthetic operands for the benchmark configurations, in the natural
 * host layout.  Element i draws from its own counter-seeded xoshiro256** stream
 * (seeded by splitmix64 of (seed, i)), so any range of elements can be regenerated
 * independently and identically by the GPU bench and the CPU checks.
 *
 * Operand model (mantissas: uniform random limbs with the top bit set; exponent word
 * 2*(BIAS + e) + sign):
 *   kind 1 -- e ~ U[-64, 64], random signs, no zeros.
 *   kind 2 -- e ~ U[-1024, 1024], random signs, each operand an exact zero with
 *             probability 1/100; 5/100 of pairs are near-cancellation pairs
 *             b = -(a*(1+2^-k)) with k ~ U[1, min(4000, p-1)], the product truncated
 *             to p bits (for sub: b = +(a*(1+2^-k)) so a-b cancels); for sqrt the
 *             operand is |a|.
 */
#include <stdint.h>
#include <string.h>

#define W_BIAS 0x3FFFFFF3ll

typedef struct { uint64_t s[4]; } xo_t;

static uint64_t splitmix64(uint64_t *x) {
  uint64_t z = (*x += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

static uint64_t xo_next(xo_t *r) {
  uint64_t *s = r->s;
  uint64_t res = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
  s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
  s[2] ^= t; s[3] = rotl(s[3], 45);
  return res;
}

static void xo_seed(xo_t *r, uint64_t seed, uint64_t index) {
  uint64_t x = seed * 0xD1B54A32D192ED03ull ^ (index + 0x632BE59BD9B4E019ull);
  for (int i = 0; i < 4; i++) r->s[i] = splitmix64(&x);
}

/* uniform integer in [lo, hi] */
static int64_t xo_range(xo_t *r, int64_t lo, int64_t hi) {
  uint64_t span = (uint64_t)(hi - lo) + 1;
  return lo + (int64_t)(xo_next(r) % span);
}

static void rand_mantissa(xo_t *r, uint32_t *m, int n) {
  for (int i = 0; i < n; i += 2) {
    uint64_t x = xo_next(r);
    m[i] = (uint32_t)x;
    if (i + 1 < n) m[i + 1] = (uint32_t)(x >> 32);
  }
  m[n - 1] |= 0x80000000u;
}

static uint32_t exp_word(int64_t e, uint32_t sign) { return (uint32_t)(2 * (W_BIAS + e) + sign); }

/* b = a * (1 + 2^-k) truncated to p bits; returns the exponent increment (0 or 1) */
static int grow(uint32_t *b, const uint32_t *a, int n, int k) {
  uint64_t c = 0;
  int ws = k >> 5, bs = k & 31;
  for (int i = 0; i < n; i++) {
    int j = i + ws;
    uint32_t lo = j < n ? a[j] : 0, hi = j + 1 < n ? a[j + 1] : 0;
    uint32_t sh = bs ? ((lo >> bs) | (hi << (32 - bs))) : lo;
    c += (uint64_t)a[i] + sh;
    b[i] = (uint32_t)c;
    c >>= 32;
  }
  if (c) {
    for (int i = 0; i < n - 1; i++) b[i] = (b[i] >> 1) | (b[i + 1] << 31);
    b[n - 1] = (b[n - 1] >> 1) | 0x80000000u;
    return 1;
  }
  return 0;
}

/*
 * Generate elements [first, first+count) of a workload.  op: 0 neg, 1 add, 2 sub,
 * 3 mul, 4 div, 5 sqrt (shapes the cancellation pairs and sqrt operands).
 * b_exps/b_limbs may be NULL for unary ops.  Returns 0, or -1 for an unknown kind.
 */
int fpa_workload_generate(int kind, uint64_t seed, int op, uint32_t precision, uint64_t first, uint32_t count,
                          uint32_t *a_exps, uint32_t *a_limbs, uint32_t *b_exps, uint32_t *b_limbs) {
  if (kind != 1 && kind != 2) return -1;
  int n = (int)(precision / 32);
  int p = 32 * n;
  for (uint32_t t = 0; t < count; t++) {
    xo_t r;
    xo_seed(&r, seed, first + t);
    uint32_t *a = a_limbs + (size_t)t * n;
    uint32_t *b = b_limbs ? b_limbs + (size_t)t * n : NULL;
    int64_t elo = kind == 1 ? -64 : -1024, ehi = kind == 1 ? 64 : 1024;
    rand_mantissa(&r, a, n);
    int64_t ea = xo_range(&r, elo, ehi);
    uint32_t sa = (uint32_t)(xo_next(&r) & 1);
    if (op == 5) sa = 0;
    a_exps[t] = exp_word(ea, sa);
    if (b) {
      rand_mantissa(&r, b, n);
      int64_t eb = xo_range(&r, elo, ehi);
      uint32_t sb = (uint32_t)(xo_next(&r) & 1);
      b_exps[t] = exp_word(eb, sb);
    }
    if (kind == 2) {
      uint64_t u = xo_next(&r);
      uint32_t pz_a = (uint32_t)(u % 100), pz_b = (uint32_t)((u >> 16) % 100), pc = (uint32_t)((u >> 32) % 100);
      if (b && pc < 5) {
        int kmax = p - 1 < 4000 ? p - 1 : 4000;
        int k = (int)xo_range(&r, 1, kmax);
        int up = grow(b, a, n, k);
        uint32_t sb = op == 2 ? sa : sa ^ 1u;
        b_exps[t] = exp_word(ea + up, sb);
      }
      if (pz_a == 0) { a_exps[t] = 2; memset(a, 0, 4 * (size_t)n); }
      if (b && pz_b == 0) { b_exps[t] = 2; memset(b, 0, 4 * (size_t)n); }
    }
  }
  return 0;
}
