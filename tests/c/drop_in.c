/*
 * Drop-in check in C: the reference's own test harness (tests.c: generators, MPFR gold,
 * tests_compare_results) drives this library through the unmodified fpa.h API, the way
 * the reference's fpa_test.c test_<op> functions do (fpa_test.c:207-327).
 * tests_compare_results exits(1) on the first mismatch, so the exit status is the verdict.
 *
 * usage: drop_in <op> <count> <precision> [source_precision]
 *        op = add|sub|mul|div|sqrt|neg|sgn|cmp|set_ui|set_si|set_float|set_double, or set
 *        (needs source_precision)
 * sgn/cmp write into a device int32 array, as fpa_test.c does with cudaMalloc
 * (fpa_test.c:144-155); here hipMalloc.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>
#include <hip/hip_runtime_api.h>
#include "fpa.h"
#include "tests.h"

static int sgn_cmp(const char *op, uint32_t count, uint32_t precision) {
  tests_t *t = !strcmp(op, "sgn") ? tests_generate_sgn(count, precision) : tests_generate_cmp(count, precision);
  if (!t) return 3;
  fpa_t a, b;
  int32_t *dr;
  if (fpa_init(a, count, precision) != fpa_success || hipMalloc((void **)&dr, 4 * (size_t)count) != hipSuccess) return 4;
  fpa_copy_array_to_device(a, t->a_exps, t->a_limbs);
  fpa_result rc;
  if (!strcmp(op, "sgn")) rc = fpa_sgn(dr, a);
  else {
    fpa_init(b, count, precision);
    fpa_copy_array_to_device(b, t->b_exps, t->b_limbs);
    rc = fpa_cmp(dr, a, b);
  }
  if (rc != fpa_success) { fprintf(stderr, "fpa_%s returned %d\n", op, rc); return 5; }
  if (hipMemcpy(t->r_exps, dr, 4 * (size_t)count, hipMemcpyDeviceToHost) != hipSuccess) return 6;
  tests_compare_results(t);
  printf("%s tests passed: %u values at %u bits\n", op, count, precision);
  tests_free(t);
  fpa_clear(a);
  if (strcmp(op, "sgn")) fpa_clear(b);
  hipFree(dr);
  return 0;
}

static int set(uint32_t count, uint32_t precision, uint32_t source) {
  tests_t *t = tests_generate_set(count, precision, source);
  if (!t) return 3;
  fpa_t r, x;
  if (fpa_init(r, count, precision) != fpa_success || fpa_init(x, count, source) != fpa_success) return 4;
  fpa_copy_array_to_device(x, t->a_exps, t->a_limbs);
  fpa_result rc = fpa_set(r, x, fpa_round_ties_to_even);
  if (rc != fpa_success) { fprintf(stderr, "fpa_set returned %d\n", rc); return 5; }
  fpa_copy_array_to_host(r, t->r_exps, t->r_limbs);
  tests_compare_results(t);
  printf("set tests passed: %u values, %u -> %u bits\n", count, source, precision);
  tests_free(t);
  fpa_clear(r);
  fpa_clear(x);
  return 0;
}

/* fpa_set_ui / si / float / double from the generator's host values (fpa_test.c:12-110) */
static int set_scalar(const char *op, uint32_t count, uint32_t precision) {
  tests_t *t;
  if (!strcmp(op, "set_ui")) t = tests_generate_set_ui(count, precision);
  else if (!strcmp(op, "set_si")) t = tests_generate_set_si(count, precision);
  else if (!strcmp(op, "set_float")) t = tests_generate_set_float(count, precision);
  else t = tests_generate_set_double(count, precision);
  if (!t) return 3;
  fpa_t r;
  if (fpa_init(r, count, precision) != fpa_success) return 4;
  fpa_result rc;
  if (!strcmp(op, "set_ui")) rc = fpa_set_ui(r, (uint64_t *)t->values);
  else if (!strcmp(op, "set_si")) rc = fpa_set_si(r, (int64_t *)t->values);
  else if (!strcmp(op, "set_float")) rc = fpa_set_float(r, (float *)t->values);
  else rc = fpa_set_double(r, (double *)t->values);
  if (rc != fpa_success) { fprintf(stderr, "fpa_%s returned %d\n", op, rc); return 5; }
  fpa_copy_array_to_host(r, t->r_exps, t->r_limbs);
  tests_compare_results(t);
  printf("%s tests passed: %u values at %u bits\n", op, count, precision);
  free(t->values);
  tests_free(t);
  fpa_clear(r);
  return 0;
}

int main(int argc, char **argv) {
  if (argc != 4 && argc != 5) { fprintf(stderr, "usage: %s op count precision [source]\n", argv[0]); return 2; }
  const char *op = argv[1];
  uint32_t count = (uint32_t)atoi(argv[2]), precision = (uint32_t)atoi(argv[3]);
  if (!strcmp(op, "sgn") || !strcmp(op, "cmp")) return sgn_cmp(op, count, precision);
  if (!strncmp(op, "set_", 4)) return set_scalar(op, count, precision);
  if (!strcmp(op, "set")) {
    if (argc != 5) { fprintf(stderr, "set needs a source precision\n"); return 2; }
    return set(count, precision, (uint32_t)atoi(argv[4]));
  }
  tests_t *t;
  if (!strcmp(op, "add")) t = tests_generate_add(count, precision);
  else if (!strcmp(op, "sub")) t = tests_generate_sub(count, precision);
  else if (!strcmp(op, "mul")) t = tests_generate_mul(count, precision);
  else if (!strcmp(op, "div")) t = tests_generate_div(count, precision);
  else if (!strcmp(op, "sqrt")) t = tests_generate_sqrt(count, precision);
  else if (!strcmp(op, "neg")) t = tests_generate_neg(count, precision);
  else { fprintf(stderr, "unknown op %s\n", op); return 2; }
  if (!t) return 3;
  fpa_t r, a, b;
  if (fpa_init(r, count, precision) != fpa_success || fpa_init(a, count, precision) != fpa_success) return 4;
  fpa_copy_array_to_device(a, t->a_exps, t->a_limbs);
  fpa_result rc;
  if (!strcmp(op, "sqrt")) rc = fpa_sqrt(r, a, fpa_round_ties_to_even);
  else if (!strcmp(op, "neg")) rc = fpa_neg(r, a);
  else {
    fpa_init(b, count, precision);
    fpa_copy_array_to_device(b, t->b_exps, t->b_limbs);
    if (!strcmp(op, "add")) rc = fpa_add(r, a, b, fpa_round_ties_to_even);
    else if (!strcmp(op, "sub")) rc = fpa_sub(r, a, b, fpa_round_ties_to_even);
    else if (!strcmp(op, "mul")) rc = fpa_mul(r, a, b, fpa_round_ties_to_even);
    else rc = fpa_div(r, a, b, fpa_round_ties_to_even);
  }
  if (rc != fpa_success) { fprintf(stderr, "fpa_%s returned %d\n", op, rc); return 5; }
  fpa_copy_array_to_host(r, t->r_exps, t->r_limbs);
  tests_compare_results(t);
  printf("%s tests passed: %u values at %u bits\n", op, count, precision);
  tests_free(t);
  fpa_clear(r);
  fpa_clear(a);
  if (strcmp(op, "sqrt") && strcmp(op, "neg")) fpa_clear(b);
  return 0;
}
