/*
 * C host side of fpa.h / fpa_kernels.h: allocation, transfers, argument validation
 * and dispatch to the HIP kernel layer.
 *
 * Replaces the reference's fpa_management.h:1-47 (allocation + host transposes) and
 * fpa_dispatch.h:1-456 (validation + precision switch + <<<>>>).  Sizes are computed in
 * size_t (the reference's uint32 products overflow at count*precision >= 2^32), the
 * device layout equals the host layout (no transpose), and each launch is checked.
 */
#define _GNU_SOURCE
#include <dlfcn.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <hip/hip_runtime_api.h>

#include "../../include/fpa.h"
#include "../../include/fpa_kernels.h"
#include "fpa_scalar.h"
#include "fpa_shapes.h"

/* HIP API failure: print and exit(1), as the reference's $GPU macro (fpa.cu:34) */
#define FPA_HIP(call)                                                                      \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess) {                                                                \
      fprintf(stderr, "\nCall \"%s\" failed from %s, line %d: %s\n", #call, __FILE__,     \
              __LINE__, hipGetErrorString(e_));                                            \
      exit(1);                                                                             \
    }                                                                                      \
  } while (0)

/* per-op launchers defined in k_<op>*.hip, one per compilation part: each launches the
 * shapes its part compiles and returns -1 for the others */
typedef int (*launch_fn)(const fpa_arguments *, int, int, void *);
#define FPA_PART_DECL(name, i) int fpa_impl_launch_##name##_part##i(const fpa_arguments *a, int K, int L, void *stream);
FPA_PART_DECL(neg, 0)
FPA_PART_DECL(add, 0)
FPA_PART_DECL(sub, 0)
FPA_PART_DECL(mul, 0)
FPA_PART_DECL(div, 0) FPA_PART_DECL(div, 1)
FPA_PART_DECL(sqrt, 0) FPA_PART_DECL(sqrt, 1) FPA_PART_DECL(sqrt, 2)
FPA_PART_DECL(sqrt, 3) FPA_PART_DECL(sqrt, 4) FPA_PART_DECL(sqrt, 5)
#undef FPA_PART_DECL
int fpa_impl_launch_sgn(uint32_t count, const uint32_t *x_exps, int32_t *r, void *stream);
int fpa_impl_launch_setcmp(int op, const fpa_arguments *args, int K, int L, void *stream);

#define FPA_MAX_PARTS 6
static const launch_fn launch_parts[FPA_OP_COUNT][FPA_MAX_PARTS] = {
  { fpa_impl_launch_neg_part0 },
  { fpa_impl_launch_add_part0 },
  { fpa_impl_launch_sub_part0 },
  { fpa_impl_launch_mul_part0 },
  { fpa_impl_launch_div_part0, fpa_impl_launch_div_part1 },
  { fpa_impl_launch_sqrt_part0, fpa_impl_launch_sqrt_part1, fpa_impl_launch_sqrt_part2,
    fpa_impl_launch_sqrt_part3, fpa_impl_launch_sqrt_part4, fpa_impl_launch_sqrt_part5 },
};

static int launch_op(int op, const fpa_arguments *a, int K, int L, void *stream) {
  for (int i = 0; i < FPA_MAX_PARTS && launch_parts[op][i]; i++) {
    int rc = launch_parts[op][i](a, K, L, stream);
    if (rc != -1) return rc;
  }
  return -1;
}

/* compiled shape table */
typedef struct { int n, K, L; } shape_t;
#define FPA_SHAPE_ENTRY(n_, K_, L_) { n_, K_, L_ },
static const shape_t shapes[] = { FPA_SHAPES(FPA_SHAPE_ENTRY) };
#define NSHAPES ((int)(sizeof(shapes) / sizeof(shapes[0])))

/* supported limb counts, in table order */
static const int prec_n[] = { 8, 16, 32, 64, 96, 128, 160, 192, 224, 256, 512 };
#define NPREC ((int)(sizeof(prec_n) / sizeof(prec_n[0])))

/* dispatch table: variant index per (op, precision); 0 = first compiled shape */
static int variant_tab[FPA_OP_COUNT + 1][NPREC];   /* + FPA_KERNEL_DOT */
/* optional shapes for small launches: tier t serves launches of at most tier_max values
 * (tiers sorted by bound; a launch takes the first tier it fits, else variant_tab).  Small
 * launches cannot fill 256 CUs, and the shape with the shortest per-value latency at that
 * size wins. */
#define FPA_SMALL_TIERS 4
static int tier_var[FPA_OP_COUNT + 1][NPREC][FPA_SMALL_TIERS];
static uint32_t tier_max[FPA_OP_COUNT + 1][NPREC][FPA_SMALL_TIERS];
static int ntier[FPA_OP_COUNT + 1][NPREC];
static void *cur_stream = NULL;

static int prec_index(uint32_t precision) {
  if (precision % 32 != 0) return -1;
  int n = (int)(precision / 32);
  for (int i = 0; i < NPREC; i++) if (prec_n[i] == n) return i;
  return -1;
}

int fpa_precision_supported(uint32_t precision) { return prec_index(precision) >= 0; }

/* shapes an op may use: every compiled one, except that mul exists only for
 * K <= FPA_MUL_MAX_K and the dot kernel only for K <= FPA_DOT_MAX_K */
static int shape_allowed(int op, int s) {
  if (op == FPA_KERNEL_DOT) return shapes[s].K <= FPA_DOT_MAX_K;
  if (op == FPA_OP_MUL) return shapes[s].K <= FPA_MUL_MAX_K;
  return 1;
}

/* variant v of (op, limb count n) -> index into shapes[], or -1 */
static int shape_of(int op, int n, int v) {
  int k = 0;
  for (int i = 0; i < NSHAPES; i++) {
    if (shapes[i].n != n || !shape_allowed(op, i)) continue;
    if (k == v) return i;
    k++;
  }
  return -1;
}

int fpa_kernel_variants(int op, uint32_t precision) {
  if (op < 0 || op > FPA_KERNEL_DOT || prec_index(precision) < 0) return 0;
  int n = (int)(precision / 32), k = 0;
  for (int i = 0; i < NSHAPES; i++) k += shapes[i].n == n && shape_allowed(op, i);
  return k;
}

int fpa_kernel_variant_shape(int op, uint32_t precision, int variant, int *K, int *L) {
  if (fpa_kernel_variants(op, precision) <= variant || variant < 0) return -1;
  int s = shape_of(op, (int)(precision / 32), variant);
  if (K) *K = shapes[s].K;
  if (L) *L = shapes[s].L;
  return 0;
}

int fpa_kernel_get_variant(int op, uint32_t precision) {
  int p = prec_index(precision);
  if (op < 0 || op > FPA_KERNEL_DOT || p < 0) return -1;
  return variant_tab[op][p];
}

static int set_large_variant(int op, uint32_t precision, int variant) {
  int p = prec_index(precision);
  if (op < 0 || op > FPA_KERNEL_DOT || p < 0) return -1;
  if (variant < 0 || variant >= fpa_kernel_variants(op, precision)) return -1;
  variant_tab[op][p] = variant;
  return 0;
}

/* an explicit choice applies to every launch size: it also drops the small-launch tiers */
int fpa_kernel_set_variant(int op, uint32_t precision, int variant) {
  if (set_large_variant(op, precision, variant) != 0) return -1;
  ntier[op][prec_index(precision)] = 0;
  return 0;
}

int fpa_kernel_set_small_variant(int op, uint32_t precision, int variant, uint32_t max_count) {
  int p = prec_index(precision);
  if (op < 0 || op > FPA_KERNEL_DOT || p < 0) return -1;
  if (max_count == 0) {                  /* removes every tier */
    ntier[op][p] = 0;
    return 0;
  }
  if (variant < 0 || variant >= fpa_kernel_variants(op, precision)) return -1;
  int n = ntier[op][p], i;
  for (i = 0; i < n && tier_max[op][p][i] < max_count; i++) {}
  if (i < n && tier_max[op][p][i] == max_count) {   /* same bound: replace */
    tier_var[op][p][i] = variant;
    return 0;
  }
  if (n == FPA_SMALL_TIERS) return -1;
  for (int j = n; j > i; j--) {          /* keep the tiers sorted by bound */
    tier_max[op][p][j] = tier_max[op][p][j - 1];
    tier_var[op][p][j] = tier_var[op][p][j - 1];
  }
  tier_max[op][p][i] = max_count;
  tier_var[op][p][i] = variant;
  ntier[op][p] = n + 1;
  return 0;
}

int fpa_kernel_get_variant_for_count(int op, uint32_t precision, uint64_t count) {
  int p = prec_index(precision);
  if (op < 0 || op > FPA_KERNEL_DOT || p < 0) return -1;
  for (int i = 0; i < ntier[op][p]; i++)
    if (count <= tier_max[op][p][i]) return tier_var[op][p][i];
  return variant_tab[op][p];
}

int fpa_kernel_load_tuning(const char *path) {
  FILE *f = fopen(path, "r");
  if (!f) return -1;
  char line[256];
  int applied = 0;
  /* a file gives the whole tier set of each (op, precision) it names: the first line of a
   * pair drops the tiers it had */
  unsigned char seen[FPA_OP_COUNT + 1][NPREC];
  memset(seen, 0, sizeof seen);
  while (fgets(line, sizeof line, f)) {
    int op, K, L;
    unsigned prec, maxc = 0;
    if (line[0] == '#') continue;
    int got = sscanf(line, "%d %u %d %d %u", &op, &prec, &K, &L, &maxc);
    if (got != 4 && got != 5) continue;
    int pi = prec_index(prec);
    if (op >= 0 && op <= FPA_KERNEL_DOT && pi >= 0 && !seen[op][pi]) {
      seen[op][pi] = 1;
      ntier[op][pi] = 0;
    }
    int nv = fpa_kernel_variants(op, prec);
    for (int v = 0; v < nv; v++) {
      int k, l;
      fpa_kernel_variant_shape(op, prec, v, &k, &l);
      if (k != K || l != L) continue;
      if (got == 5 && maxc > 0) fpa_kernel_set_small_variant(op, prec, v, maxc);
      else set_large_variant(op, prec, v);
      applied++;
      break;
    }
  }
  fclose(f);
  return applied;
}

int fpa_kernel_launch(int op, const fpa_arguments *args, int variant, void *stream) {
  if (op < 0 || op >= FPA_OP_COUNT || !args) return -1;
  int s = prec_index(args->precision) < 0 ? -1 : shape_of(op, (int)(args->precision / 32), variant);
  if (s < 0) return -1;
  if (args->count == 0) return 0;
  return launch_op(op, args, shapes[s].K, shapes[s].L, stream);
}

int fpa_kernel_time(int op, const fpa_arguments *args, int variant, int reps, void *stream, float *ms) {
  if (reps < 1 || !ms) return -1;
  int rc = fpa_kernel_launch(op, args, variant, stream);
  if (rc != 0) return rc;
  hipEvent_t e0, e1;
  FPA_HIP(hipEventCreate(&e0));
  FPA_HIP(hipEventCreate(&e1));
  FPA_HIP(hipEventRecord(e0, (hipStream_t)stream));
  for (int i = 0; i < reps && rc == 0; i++) rc = fpa_kernel_launch(op, args, variant, stream);
  FPA_HIP(hipEventRecord(e1, (hipStream_t)stream));
  FPA_HIP(hipEventSynchronize(e1));
  float t = 0.0f;
  FPA_HIP(hipEventElapsedTime(&t, e0, e1));
  FPA_HIP(hipEventDestroy(e0));
  FPA_HIP(hipEventDestroy(e1));
  *ms = t / (float)reps;
  return rc;
}

/* tuning table applied when the library is loaded (host-only work: no HIP call) */
static int tuning_applied;
int fpa_kernel_tuning_applied(void) { return tuning_applied; }

__attribute__((constructor)) static void load_default_tuning(void) {
  const char *env = getenv("FPA_TUNING_FILE");
  if (env && strcmp(env, "none") == 0) return;
  if (env && *env) {
    int n = fpa_kernel_load_tuning(env);
    tuning_applied = n > 0 ? n : 0;
    return;
  }
  Dl_info info;
  if (!dladdr((void *)&load_default_tuning, &info) || !info.dli_fname) return;
  char path[4096];
  const char *slash = strrchr(info.dli_fname, '/');
  size_t dir = slash ? (size_t)(slash - info.dli_fname) + 1 : 0;
  if (dir + sizeof "tuning.txt" > sizeof path) return;
  memcpy(path, info.dli_fname, dir);
  memcpy(path + dir, "tuning.txt", sizeof "tuning.txt");
  int n = fpa_kernel_load_tuning(path);
  tuning_applied = n > 0 ? n : 0;
}

void fpa_set_stream(void *stream) { cur_stream = stream; }
void *fpa_get_stream(void) { return cur_stream; }

uint32_t fpa_scalar_round_inc(int mode, uint32_t sign, uint32_t lsb, uint32_t rw, uint32_t sticky) {
  return fpa_round_inc(mode, sign, lsb, rw, sticky);
}
uint32_t fpa_scalar_recip_3by2(uint32_t d1, uint32_t d0) { return fpa_recip_3by2(d1, d0); }
uint32_t fpa_scalar_div3by2(uint32_t u2, uint32_t u1, uint32_t u0, uint32_t d1, uint32_t d0, uint32_t v) {
  return fpa_div3by2(u2, u1, u0, d1, d0, v);
}
uint32_t fpa_scalar_isqrt64(uint64_t u) { return fpa_isqrt64(u); }

/* ------------------------------ fpa.h ------------------------------ */

fpa_result fpa_init(fpa_t fpa, uint32_t count, uint32_t precision) {
  if (!fpa_precision_supported(precision)) return fpa_invalid_precisions;
  fpa->count = count;
  fpa->precision = precision;
  fpa->gpu_exps = NULL;
  fpa->gpu_limbs = NULL;
  size_t words = (size_t)count * (precision / 32);
  FPA_HIP(hipMalloc((void **)&fpa->gpu_exps, (size_t)count * 4 + 4));
  FPA_HIP(hipMalloc((void **)&fpa->gpu_limbs, words * 4 + 16));
  return fpa_success;
}

fpa_result fpa_clear(fpa_t fpa) {
  FPA_HIP(hipFree(fpa->gpu_exps));
  FPA_HIP(hipFree(fpa->gpu_limbs));
  fpa->gpu_exps = NULL;
  fpa->gpu_limbs = NULL;
  return fpa_success;
}

fpa_result fpa_copy_array_to_device(fpa_t fpa, uint32_t *exps, uint32_t *limbs) {
  size_t words = (size_t)fpa->count * (fpa->precision / 32);
  /* work queued on the fpa stream may still read these arrays; a stream created
   * non-blocking does not wait for the null stream's copy by itself */
  if (cur_stream) FPA_HIP(hipStreamSynchronize((hipStream_t)cur_stream));
  FPA_HIP(hipMemcpy(fpa->gpu_exps, exps, (size_t)fpa->count * 4, hipMemcpyHostToDevice));
  FPA_HIP(hipMemcpy(fpa->gpu_limbs, limbs, words * 4, hipMemcpyHostToDevice));
  return fpa_success;
}

fpa_result fpa_copy_array_to_host(fpa_t fpa, uint32_t *exps, uint32_t *limbs) {
  size_t words = (size_t)fpa->count * (fpa->precision / 32);
  if (cur_stream) FPA_HIP(hipStreamSynchronize((hipStream_t)cur_stream));
  FPA_HIP(hipMemcpy(exps, fpa->gpu_exps, (size_t)fpa->count * 4, hipMemcpyDeviceToHost));
  FPA_HIP(hipMemcpy(limbs, fpa->gpu_limbs, words * 4, hipMemcpyDeviceToHost));
  return fpa_success;
}

static fpa_result run(int op, fpa_ptr r, fpa_ptr a, fpa_ptr b, fpa_rounding_mode mode) {
  uint32_t count = r->count, precision = r->precision;
  if (count != a->count || (b && count != b->count)) return fpa_invalid_counts;
  if (precision != a->precision || (b && precision != b->precision)) return fpa_invalid_precisions;
  int p = prec_index(precision);
  if (p < 0) return fpa_invalid_precisions;
  fpa_arguments args;
  memset(&args, 0, sizeof args);
  args.count = count;
  args.precision = precision;
  args.mode = mode;
  args.r_exps = r->gpu_exps;
  args.r_limbs = r->gpu_limbs;
  args.a_exps = a->gpu_exps;
  args.a_limbs = a->gpu_limbs;
  args.b_exps = b ? b->gpu_exps : a->gpu_exps;
  args.b_limbs = b ? b->gpu_limbs : a->gpu_limbs;
  if (count == 0) return fpa_success;
  int rc = fpa_kernel_launch(op, &args, fpa_kernel_get_variant_for_count(op, precision, count), cur_stream);
  if (rc != 0) {
    fprintf(stderr, "\nkernel launch for op %d, precision %u failed: %d (%s)\n", op, precision, rc,
            rc > 0 ? hipGetErrorString((hipError_t)rc) : "no such variant");
    exit(1);
  }
  return fpa_success;
}

fpa_result fpa_neg(fpa_t r, fpa_t a) { return run(FPA_OP_NEG, r, a, NULL, fpa_round_ties_to_even); }
fpa_result fpa_add(fpa_t r, fpa_t a, fpa_t b, fpa_rounding_mode m) { return run(FPA_OP_ADD, r, a, b, m); }
fpa_result fpa_sub(fpa_t r, fpa_t a, fpa_t b, fpa_rounding_mode m) { return run(FPA_OP_SUB, r, a, b, m); }
fpa_result fpa_mul(fpa_t r, fpa_t a, fpa_t b, fpa_rounding_mode m) { return run(FPA_OP_MUL, r, a, b, m); }
fpa_result fpa_div(fpa_t r, fpa_t a, fpa_t b, fpa_rounding_mode m) { return run(FPA_OP_DIV, r, a, b, m); }
fpa_result fpa_sqrt(fpa_t r, fpa_t x, fpa_rounding_mode m) { return run(FPA_OP_SQRT, r, x, NULL, m); }

/* ------------------------- conversions (exact) ------------------------- */

/* fpa_set_ui/si/float/double (fpa_dispatch.h:8-86): the 4- or 8-byte sources go to a
 * staging buffer in device memory and one kernel writes every value's exponent word and
 * limbs there (k_setcmp.hip), as the reference does (cudaMalloc, copy, launch,
 * cudaFree).  The call waits for the kernel before it frees the staging buffer. */
static fpa_result set_scalars(fpa_t r, const void *values, size_t elem, int kind) {
  if (!fpa_precision_supported(r->precision)) return fpa_invalid_precisions;
  if (r->count == 0) return fpa_success;
  void *dv = NULL;
  FPA_HIP(hipMalloc(&dv, elem * (size_t)r->count));
  FPA_HIP(hipMemcpy(dv, values, elem * (size_t)r->count, hipMemcpyHostToDevice));
  fpa_arguments a;
  memset(&a, 0, sizeof a);
  a.count = r->count;
  a.precision = r->precision;
  a.r_exps = r->gpu_exps;
  a.r_limbs = r->gpu_limbs;
  a.values = dv;
  int K, L;
  fpa_kernel_variant_shape(FPA_OP_NEG, r->precision, fpa_kernel_get_variant_for_count(FPA_OP_NEG, r->precision, r->count),
                           &K, &L);
  int rc = fpa_impl_launch_setcmp(2 + kind, &a, K, L, cur_stream);
  if (rc != 0) { fprintf(stderr, "fpa_set: launch failed (%d)\n", rc); exit(1); }
  FPA_HIP(hipStreamSynchronize((hipStream_t)cur_stream));
  FPA_HIP(hipFree(dv));
  return fpa_success;
}

fpa_result fpa_set_ui(fpa_t r, uint64_t *ui) { return set_scalars(r, ui, 8, FPA_SRC_U64); }
fpa_result fpa_set_si(fpa_t r, int64_t *si) { return set_scalars(r, si, 8, FPA_SRC_S64); }
fpa_result fpa_set_float(fpa_t r, float *floats) { return set_scalars(r, floats, 4, FPA_SRC_F32); }
fpa_result fpa_set_double(fpa_t r, double *doubles) { return set_scalars(r, doubles, 8, FPA_SRC_F64); }

uint32_t fpa_scalar_encode(int kind, uint64_t bits, uint64_t *top) { return fpa_encode_scalar(kind, bits, top); }

/* set / sgn / cmp (fpa_dispatch.h:88-153): the kernels in k_setcmp.hip, launched on the
 * fpa stream with the dispatcher's neg shape (a copy-like kernel) for the precision. */
static void copy_shape(uint32_t precision, uint32_t count, int *K, int *L) {
  fpa_kernel_variant_shape(FPA_OP_NEG, precision, fpa_kernel_get_variant_for_count(FPA_OP_NEG, precision, count), K, L);
}

fpa_result fpa_set(fpa_t r, fpa_t x, fpa_rounding_mode mode) {
  if (r->count != x->count) return fpa_invalid_counts;
  if (!fpa_precision_supported(r->precision) || !fpa_precision_supported(x->precision))
    return fpa_invalid_precisions;
  if ((int)mode < 0 || (int)mode > 4) return fpa_math_error;
  fpa_arguments a;
  memset(&a, 0, sizeof a);
  a.count = r->count;
  a.precision = r->precision;
  a.set_source_precision = x->precision;
  a.mode = mode;
  a.r_exps = r->gpu_exps;
  a.r_limbs = r->gpu_limbs;
  a.a_exps = x->gpu_exps;
  a.a_limbs = x->gpu_limbs;
  int K, L;
  copy_shape(r->precision, r->count, &K, &L);
  int rc = fpa_impl_launch_setcmp(1, &a, K, L, cur_stream);
  if (rc != 0) { fprintf(stderr, "fpa_set: launch failed (%d)\n", rc); exit(1); }
  return fpa_success;
}

fpa_result fpa_sgn(int32_t *r, fpa_t x) {
  int rc = fpa_impl_launch_sgn(x->count, x->gpu_exps, r, cur_stream);
  if (rc != 0) { fprintf(stderr, "fpa_sgn: launch failed (%d)\n", rc); exit(1); }
  return fpa_success;
}

fpa_result fpa_cmp(int32_t *r, fpa_t a, fpa_t b) {
  if (a->count != b->count) return fpa_invalid_counts;
  if (a->precision != b->precision) return fpa_invalid_precisions;
  if (!fpa_precision_supported(a->precision)) return fpa_invalid_precisions;
  fpa_arguments g;
  memset(&g, 0, sizeof g);
  g.count = a->count;
  g.precision = a->precision;
  g.r_exps = (uint32_t *)r;
  g.a_exps = a->gpu_exps;
  g.a_limbs = a->gpu_limbs;
  g.b_exps = b->gpu_exps;
  g.b_limbs = b->gpu_limbs;
  int K, L;
  copy_shape(a->precision, a->count, &K, &L);
  int rc = fpa_impl_launch_setcmp(0, &g, K, L, cur_stream);
  if (rc != 0) { fprintf(stderr, "fpa_cmp: launch failed (%d)\n", rc); exit(1); }
  return fpa_success;
}
