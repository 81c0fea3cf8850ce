/*
 * fpa.h -- public C API of the MI355X fpa library.
 *
 * Drop-in for the reference's fpa.h (nemmart/CUBigFloat fpa.h:1-69): the same type
 * names, enum values, struct layout and function signatures, so code written against
 * the reference compiles and links unchanged.  An fpa_t is an array of `count` values
 * of one precision held in device memory (HBM):
 *
 *   gpu_exps  : count 32-bit exponent words
 *               (<= 8: class code; otherwise sign = bit 0, biased exponent = word >> 1,
 *               bias 0x3FFFFFF3)
 *   gpu_limbs : count * precision/32 little-endian 32-bit limbs, value i at
 *               [i*precision/32, (i+1)*precision/32) -- the same natural layout the
 *               host uses, so the copy functions are plain memcpys (the reference
 *               transposes on the host, fpa_management.h:17-47).
 *
 * Supported precisions: 256, 512, 1024, 2048, ..., 8192 (steps of 1024) and 16384 bits
 * (the reference accepts 1024..8192 in steps of 1024, fpa_management.h:2-3).
 *
 * Semantics follow the reference's action tables and exponent range rules with results
 * correctly rounded in the requested mode (bit-exact vs MPFR for the four MPFR modes).
 * Deliberate differences from the reference's GPU code: value/0 returns +-INF (MPFR)
 * rather than NAN; 0-b returns -b; corner-case roundings that the reference gets
 * wrong are correctly rounded; special results write zero limbs.
 *
 * Operations are asynchronous on the stream set by fpa_set_stream() (default: the null
 * stream, as the reference's <<<>>> launches); fpa_copy_array_to_host synchronises.
 * HIP API failures print a message and exit(1), as the reference's $GPU macro does.
 */
#ifndef FPA_H
#define FPA_H

#include <stdint.h>

#define GPU_FPA

typedef enum {
  fpa_round_truncate,              /* 0: toward zero           (MPFR_RNDZ) */
  fpa_round_up,                    /* 1: toward +infinity      (MPFR_RNDU) */
  fpa_round_down,                  /* 2: toward -infinity      (MPFR_RNDD) */
  fpa_round_ties_to_even,          /* 3: nearest, ties to even (MPFR_RNDN) */
  fpa_round_ties_away_from_zero    /* 4: nearest, ties away    (no MPFR mode) */
} fpa_rounding_mode;

typedef enum {
  fpa_success,                     /* 0 */
  fpa_math_error,                  /* 1 */
  fpa_invalid_precisions,          /* 2 */
  fpa_invalid_counts,              /* 3 */
} fpa_result;

struct fpa_s {
  uint32_t  count;
  uint32_t  precision;
  uint32_t *gpu_exps;
  uint32_t *gpu_limbs;
};

typedef struct fpa_s fpa_t[1];
typedef struct fpa_s *fpa_ptr;

#ifdef __cplusplus
extern "C" {
#endif

/* allocation and transfer -- replaces fpa_management.h:1-47 */
fpa_result fpa_init(fpa_t fpa, uint32_t count, uint32_t precision);
fpa_result fpa_clear(fpa_t fpa);
fpa_result fpa_copy_array_to_device(fpa_t fpa, uint32_t *exps, uint32_t *limbs);
fpa_result fpa_copy_array_to_host(fpa_t fpa, uint32_t *exps, uint32_t *limbs);

/* conversions -- replaces fpa_dispatch.h:8-107 (fpa_warp_set.h) */
fpa_result fpa_set_ui(fpa_t r, uint64_t *ui);
fpa_result fpa_set_si(fpa_t r, int64_t *si);
fpa_result fpa_set_float(fpa_t r, float *floats);
fpa_result fpa_set_double(fpa_t r, double *doubles);
/* r = x rounded to r's precision in `mode` (any two supported precisions; special
 * classes copied; E > 0x7FFFFFFF after rounding up -> INF).  Same results as mpfr_set,
 * the reference harness's gold (tests.c:580-605); the reference's own kernel is
 * unfinished (fpa_warp_set.h:134-199).  fpa_invalid_counts if the counts differ. */
fpa_result fpa_set(fpa_t r, fpa_t x, fpa_rounding_mode mode);

/* r[i] = sign / three-way compare (-1, 0, +1) into a device int32 array, by the
 * reference's class tables (fpa_threaded.h:1-93): NaN, SNaN and the tables' error
 * entries (equal infinities, equal-sign TINY pairs) give 0 -- fpa_dispatch.h:109-153 */
fpa_result fpa_sgn(int32_t *r, fpa_t x);
fpa_result fpa_cmp(int32_t *r, fpa_t a, fpa_t b);

/* element-wise arithmetic -- fpa_dispatch.h:155-421 */
fpa_result fpa_neg(fpa_t r, fpa_t a);
fpa_result fpa_add(fpa_t r, fpa_t a, fpa_t b, fpa_rounding_mode mode);
fpa_result fpa_sub(fpa_t r, fpa_t a, fpa_t b, fpa_rounding_mode mode);
fpa_result fpa_mul(fpa_t r, fpa_t a, fpa_t b, fpa_rounding_mode mode);
fpa_result fpa_div(fpa_t r, fpa_t a, fpa_t b, fpa_rounding_mode mode);
fpa_result fpa_sqrt(fpa_t r, fpa_t x, fpa_rounding_mode mode);

#ifdef __cplusplus
}
#endif

#endif
