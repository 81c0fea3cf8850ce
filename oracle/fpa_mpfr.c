/*
 * ORACLE -- test infrastructure only (tests/, __graft_entry__.smoke(), bench.py cpu_baseline).
 *
 * 1. MPFR gold: the reference's tests compute every expected result with
 *    mpfr_add/sub/mul/div/sqrt in round-to-nearest (tests.c:756, 794, 832, 876,
 *    905/918) and convert with tests_get_mpfr/tests_set_mpfr (tests.c:114-159).
 *    This file restates that conversion (plus NaN/Inf classes, which the reference's
 *    harness does not map) and evaluates the ops over a widened exponent range, then
 *    applies the format's range rules (same as fpa_oracle.c).
 * 2. "Reference CPU build": fpa.h declares an MPFR-backed struct under -DCPU
 *    (fpa.h:28-34) with no implementation in the reference.  fpa_cpu_* below is that
 *    backend (array of mpfr_t per fpa_t), parallelised over elements with OpenMP.
 *    bench.py times it on the host cores as the CPU baseline.
 *
 * Mapping of the reference's rounding modes to MPFR (tests.c:55-73): truncate->RNDZ,
 * up->RNDU, down->RNDD, ties_to_even->RNDN; ties-away has no MPFR mode (callers skip it).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <gmp.h>
#include <mpfr.h>
#include <omp.h>

#define BIAS 0x3FFFFFF3ll
enum { C_ZERO = 2, C_NTINY = 3, C_PTINY = 4, C_NINF = 5, C_PINF = 6, C_NAN = 7, C_SNAN = 8 };

static mpfr_rnd_t rnd_of(int mode) {
  switch (mode) {
    case 0: return MPFR_RNDZ;
    case 1: return MPFR_RNDU;
    case 2: return MPFR_RNDD;
    default: return MPFR_RNDN;
  }
}

static void wide_range(void) {
  mpfr_set_emin(mpfr_get_emin_min());
  mpfr_set_emax(mpfr_get_emax_max());
}

/* format -> mpfr.  Returns 0 on success, -1 for classes MPFR cannot hold (TINY, SNAN). */
int mpf_from_fpa(mpfr_t x, uint32_t w, const uint32_t *limbs, int n) {
  if (w <= C_SNAN && w >= C_ZERO) {
    switch (w) {
      case C_ZERO: mpfr_set_zero(x, 1); return 0;
      case C_NINF: mpfr_set_inf(x, -1); return 0;
      case C_PINF: mpfr_set_inf(x, 1); return 0;
      case C_NAN: mpfr_set_nan(x); return 0;
      default: return -1;
    }
  }
  mpz_t z;
  mpz_init(z);
  mpz_import(z, (size_t)n, -1, 4, 0, 0, limbs);
  mpfr_set_z_2exp(x, z, (mpfr_exp_t)((int64_t)(w >> 1) - BIAS - 32ll * n), MPFR_RNDZ);
  if (w & 1) mpfr_neg(x, x, MPFR_RNDZ);
  mpz_clear(z);
  return 0;
}

/* mpfr -> format with the op's exponent window [emin_E, emax_E] for biased E. */
void mpf_to_fpa(uint32_t *w, uint32_t *limbs, int n, mpfr_t x, int64_t emin_E, int64_t emax_E) {
  memset(limbs, 0, 4 * (size_t)n);
  if (mpfr_nan_p(x)) { *w = C_NAN; return; }
  if (mpfr_inf_p(x)) { *w = mpfr_signbit(x) ? C_NINF : C_PINF; return; }
  if (mpfr_zero_p(x)) { *w = C_ZERO; return; }
  int sign = mpfr_signbit(x) ? 1 : 0;
  mpz_t z;
  mpz_init(z);
  mpfr_exp_t ex = mpfr_get_z_2exp(z, x);   /* x = z * 2^ex, |z| has exactly prec bits */
  mpz_abs(z, z);
  int64_t e = (int64_t)ex + 32ll * n;       /* x = 0.M * 2^e */
  int64_t E = e + BIAS;
  if (E < emin_E) *w = (uint32_t)(C_PTINY - sign);
  else if (E > emax_E) *w = (uint32_t)(C_PINF - sign);
  else {
    size_t cnt = 0;
    mpz_export(limbs, &cnt, -1, 4, 0, 0, z);
    *w = (uint32_t)(2 * E + sign);
  }
  mpz_clear(z);
}

/*
 * Gold for one element.  op: 0 neg, 1 add, 2 sub, 3 mul, 4 div, 5 sqrt.
 * Returns 0 if MPFR produced the result, -1 if an input class has no MPFR meaning.
 */
static int gold_one(int op, int n, int mode, uint32_t *rw, uint32_t *rl, uint32_t aw, const uint32_t *al,
                    uint32_t bw, const uint32_t *bl, mpfr_t r, mpfr_t a, mpfr_t b) {
  mpfr_rnd_t rnd = rnd_of(mode);
  if (mpf_from_fpa(a, aw, al, n)) return -1;
  if (op >= 1 && op <= 4 && mpf_from_fpa(b, bw, bl, n)) return -1;
  int64_t lo = 5, hi = 0x7FFFFFFFll;
  switch (op) {
    case 0: mpfr_neg(r, a, rnd); break;
    case 1: mpfr_add(r, a, b, rnd); break;
    case 2: mpfr_sub(r, a, b, rnd); break;
    case 3: mpfr_mul(r, a, b, rnd); break;
    case 4: mpfr_div(r, a, b, rnd); lo = 10; hi = 0x7FFFFFFEll; break;
    case 5: mpfr_sqrt(r, a, rnd); lo = 0; hi = 0x7FFFFFFFll; break;
  }
  mpf_to_fpa(rw, rl, n, r, lo, hi);
  return 0;
}

/* ok[i] = 1 where MPFR produced the gold for element i */
void mpfr_gold_array(int op, uint32_t count, uint32_t precision, int mode,
                     uint32_t *r_exps, uint32_t *r_limbs, uint8_t *ok,
                     const uint32_t *a_exps, const uint32_t *a_limbs,
                     const uint32_t *b_exps, const uint32_t *b_limbs, int threads) {
  int n = (int)(precision / 32);
  if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel num_threads(threads)
  {
    wide_range();
    mpfr_t r, a, b;
    mpfr_inits2((mpfr_prec_t)precision, r, a, b, (mpfr_ptr)0);
#pragma omp for schedule(dynamic, 64)
    for (int64_t i = 0; i < (int64_t)count; i++) {
      int rc = gold_one(op, n, mode, &r_exps[i], r_limbs + (size_t)i * n, a_exps[i], a_limbs + (size_t)i * n,
                        b_exps ? b_exps[i] : 0, b_limbs ? b_limbs + (size_t)i * n : NULL, r, a, b);
      ok[i] = rc == 0;
    }
    mpfr_clears(r, a, b, (mpfr_ptr)0);
  }
}

/* --------------------- MPFR-backed fpa.h CPU backend --------------------- */
typedef struct {
  uint32_t count;
  uint32_t precision;
  mpfr_t *mpfr;
} fpa_cpu_s;

static int cpu_threads = 0;
void fpa_cpu_set_threads(int t) { cpu_threads = t; }
static int nthreads(void) { return cpu_threads > 0 ? cpu_threads : omp_get_max_threads(); }

int fpa_cpu_init(fpa_cpu_s *f, uint32_t count, uint32_t precision) {
  f->count = count;
  f->precision = precision;
  f->mpfr = (mpfr_t *)malloc(sizeof(mpfr_t) * (count ? count : 1));
  for (uint32_t i = 0; i < count; i++) mpfr_init2(f->mpfr[i], (mpfr_prec_t)precision);
  return 0;
}

int fpa_cpu_clear(fpa_cpu_s *f) {
  for (uint32_t i = 0; i < f->count; i++) mpfr_clear(f->mpfr[i]);
  free(f->mpfr);
  f->mpfr = NULL;
  return 0;
}

int fpa_cpu_copy_in(fpa_cpu_s *f, const uint32_t *exps, const uint32_t *limbs) {
  int n = (int)(f->precision / 32);
#pragma omp parallel num_threads(nthreads())
  {
    wide_range();
#pragma omp for schedule(static)
    for (int64_t i = 0; i < (int64_t)f->count; i++)
      if (mpf_from_fpa(f->mpfr[i], exps[i], limbs + (size_t)i * n, n)) mpfr_set_nan(f->mpfr[i]);
  }
  return 0;
}

int fpa_cpu_copy_out(fpa_cpu_s *f, uint32_t *exps, uint32_t *limbs, int op) {
  int n = (int)(f->precision / 32);
  int64_t lo = op == 4 ? 10 : 5, hi = op == 4 ? 0x7FFFFFFEll : 0x7FFFFFFFll;
  if (op == 5) lo = 0;
#pragma omp parallel num_threads(nthreads())
  {
    wide_range();
#pragma omp for schedule(static)
    for (int64_t i = 0; i < (int64_t)f->count; i++) mpf_to_fpa(&exps[i], limbs + (size_t)i * n, n, f->mpfr[i], lo, hi);
  }
  return 0;
}

/* op: 1 add, 2 sub, 3 mul, 4 div, 5 sqrt.  Timed region of the CPU baseline. */
int fpa_cpu_op(int op, fpa_cpu_s *r, fpa_cpu_s *a, fpa_cpu_s *b, int mode) {
  if (r->count != a->count || (b && r->count != b->count)) return 3;
  mpfr_rnd_t rnd = rnd_of(mode);
#pragma omp parallel num_threads(nthreads())
  {
    wide_range();
#pragma omp for schedule(static)
    for (int64_t i = 0; i < (int64_t)r->count; i++) {
      switch (op) {
        case 1: mpfr_add(r->mpfr[i], a->mpfr[i], b->mpfr[i], rnd); break;
        case 2: mpfr_sub(r->mpfr[i], a->mpfr[i], b->mpfr[i], rnd); break;
        case 3: mpfr_mul(r->mpfr[i], a->mpfr[i], b->mpfr[i], rnd); break;
        case 4: mpfr_div(r->mpfr[i], a->mpfr[i], b->mpfr[i], rnd); break;
        case 5: mpfr_sqrt(r->mpfr[i], a->mpfr[i], rnd); break;
      }
    }
  }
  return 0;
}

size_t fpa_cpu_struct_size(void) { return sizeof(fpa_cpu_s); }
int fpa_cpu_max_threads(void) { return omp_get_max_threads(); }
const char *fpa_cpu_mpfr_version(void) { return mpfr_get_version(); }

/* MPFR gold for dot chains: acc <- RN(acc + RN(x_k * y_k)) from +0, chains in parallel.
 * Also the config-4 CPU baseline (timed by the caller). */
void mpfr_dot_array(uint32_t count, uint32_t precision, int mode, const uint32_t *lengths, const uint64_t *offsets,
                    const uint32_t *x_exps, const uint32_t *x_limbs, const uint32_t *y_exps, const uint32_t *y_limbs,
                    uint32_t *r_exps, uint32_t *r_limbs, int threads) {
  int n = (int)(precision / 32);
  mpfr_rnd_t rnd = rnd_of(mode);
  if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel num_threads(threads)
  {
    wide_range();
    mpfr_t acc, p, x, y;
    mpfr_inits2((mpfr_prec_t)precision, acc, p, x, y, (mpfr_ptr)0);
#pragma omp for schedule(dynamic, 16)
    for (int64_t j = 0; j < (int64_t)count; j++) {
      mpfr_set_zero(acc, 1);
      for (uint64_t k = offsets[j]; k < offsets[j] + lengths[j]; k++) {
        mpf_from_fpa(x, x_exps[k], x_limbs + k * n, n);
        mpf_from_fpa(y, y_exps[k], y_limbs + k * n, n);
        mpfr_mul(p, x, y, rnd);
        mpfr_add(acc, acc, p, rnd);
      }
      mpf_to_fpa(&r_exps[j], r_limbs + (size_t)j * n, n, acc, 5, 0x7FFFFFFFll);
    }
    mpfr_clears(acc, p, x, y, (mpfr_ptr)0);
  }
}

/* ------------------------ set / sgn / cmp gold ------------------------ */
/* r = mpfr_set(x) at r_precision (tests.c:580-605 draws its gold this way); ok[i] = 0
 * for inputs MPFR cannot hold (TINY, SNAN). */
void mpfr_set_array(uint32_t count, uint32_t r_precision, uint32_t x_precision, int mode, uint32_t *r_exps,
                    uint32_t *r_limbs, uint8_t *ok, const uint32_t *x_exps, const uint32_t *x_limbs) {
  int nr = (int)(r_precision / 32), nx = (int)(x_precision / 32);
  wide_range();
  mpfr_t r, x;
  mpfr_init2(r, (mpfr_prec_t)r_precision);
  mpfr_init2(x, (mpfr_prec_t)x_precision);
  for (uint32_t i = 0; i < count; i++) {
    ok[i] = mpf_from_fpa(x, x_exps[i], x_limbs + (size_t)i * nx, nx) == 0;
    if (!ok[i]) continue;
    mpfr_set(r, x, rnd_of(mode));
    mpf_to_fpa(&r_exps[i], r_limbs + (size_t)i * nr, nr, r, 5, 0x7FFFFFFFll);
  }
  mpfr_clear(r);
  mpfr_clear(x);
}

/* r[i] = mpfr_sgn(a) / mpfr_cmp(a, b) clamped to -1/0/+1; NaN operands give 0 (MPFR sets
 * its erange flag and returns 0); ok[i] = 0 where an input is TINY or SNAN. */
void mpfr_cmp_array(uint32_t count, uint32_t precision, const uint32_t *a_exps, const uint32_t *a_limbs,
                    const uint32_t *b_exps, const uint32_t *b_limbs, int32_t *r, uint8_t *ok) {
  int n = (int)(precision / 32);
  wide_range();
  mpfr_t a, b;
  mpfr_init2(a, (mpfr_prec_t)precision);
  mpfr_init2(b, (mpfr_prec_t)precision);
  for (uint32_t i = 0; i < count; i++) {
    ok[i] = mpf_from_fpa(a, a_exps[i], a_limbs + (size_t)i * n, n) == 0;
    if (b_exps) ok[i] &= mpf_from_fpa(b, b_exps[i], b_limbs + (size_t)i * n, n) == 0;
    if (!ok[i]) continue;
    int c = b_exps ? mpfr_cmp(a, b) : mpfr_sgn(a);
    r[i] = c > 0 ? 1 : (c < 0 ? -1 : 0);
  }
  mpfr_clear(a);
  mpfr_clear(b);
}
