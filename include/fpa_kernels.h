/*
 * fpa_kernels.h -- C ABI of the HIP kernel layer under fpa.h.
 *
 * The reference's kernel ABI is one __global__ per op and size, all taking an
 * fpa_arguments struct by value (fpa_kernels.h:1-191, fpa_internal.h:34-46), launched
 * by a precision switch in fpa_dispatch.h.  Here a single launcher takes the same
 * descriptor plus a kernel *variant* -- a (limbs-per-lane K, lanes-per-value L) shape
 * with K*L = precision/32 -- and a HIP stream.  Plain pointers and sizes only.
 */
#ifndef FPA_KERNELS_H
#define FPA_KERNELS_H

#include <stdint.h>
#include "fpa.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Same fields and meaning as the reference's fpa_arguments (fpa_internal.h:34-46). */
typedef struct {
  uint32_t           count;
  uint32_t           precision;
  uint32_t           set_source_precision;
  fpa_rounding_mode  mode;
  uint32_t          *r_exps;
  uint32_t          *r_limbs;
  uint32_t          *a_exps;
  uint32_t          *a_limbs;
  uint32_t          *b_exps;
  uint32_t          *b_limbs;
  void              *values;
} fpa_arguments;

typedef enum {
  FPA_OP_NEG = 0,
  FPA_OP_ADD = 1,
  FPA_OP_SUB = 2,
  FPA_OP_MUL = 3,
  FPA_OP_DIV = 4,
  FPA_OP_SQRT = 5,
  FPA_OP_COUNT = 6
} fpa_op;

/* Variant-table id of the dot-chain kernel (fpa_batch.h): accepted by
 * fpa_kernel_variants / _variant_shape / _get_variant / _set_variant and in tuning
 * files, not by fpa_kernel_launch. */
#define FPA_KERNEL_DOT 6

/* Launch op over args->count elements with kernel variant `variant` (index into
 * fpa_kernel_variants) on `stream` (a hipStream_t, NULL = null stream).
 * Returns 0 on success, otherwise a hipError_t value (or -1 for an unknown
 * op/precision/variant).  count == 0 launches nothing. */
int fpa_kernel_launch(int op, const fpa_arguments *args, int variant, void *stream);

/* Number of compiled variants for (op, precision); 0 if the precision is unsupported.
 * op may also be FPA_KERNEL_DOT (same shapes as mul). */
int fpa_kernel_variants(int op, uint32_t precision);

/* Shape of a variant: limbs per lane, lanes per value. Returns 0, or -1 if invalid. */
int fpa_kernel_variant_shape(int op, uint32_t precision, int variant, int *limbs_per_lane, int *lanes_per_value);

/* Variant the fpa.h dispatcher uses for (op, precision), and an override of it (the
 * autotuner writes these; an override applies to every launch size, so it also removes
 * the small-launch variant below).  Returns -1 for invalid arguments. */
int fpa_kernel_get_variant(int op, uint32_t precision);
int fpa_kernel_set_variant(int op, uint32_t precision, int variant);

/* The variant tables are process-wide: set them before launching from several threads. */

/* Small launches: up to four tiers of (max_count, variant) per (op, precision); a launch of
 * `count` values takes the tier with the smallest max_count >= count, else the variant
 * above.  Setting a tier with an existing max_count replaces it; max_count 0 removes every
 * tier; -1 if the variant is invalid or four other tiers exist.  get_variant_for_count is
 * what the fpa.h calls, batch segments, server groups and dot chains all ask. */
int fpa_kernel_set_small_variant(int op, uint32_t precision, int variant, uint32_t max_count);
int fpa_kernel_get_variant_for_count(int op, uint32_t precision, uint64_t count);

/* Load "op precision K L" lines (as written by the autotuner) into the dispatch table;
 * a fifth field max_count makes the line a small-launch tier.
 * Returns the number of entries applied, or -1 if the file cannot be read.
 * When the library is loaded it applies, if present, the file named by the environment
 * variable FPA_TUNING_FILE, else tuning.txt in the directory holding the library
 * (FPA_TUNING_FILE=none skips both). */
int fpa_kernel_load_tuning(const char *path);

/* Number of entries the load-time tuning file applied (0 if none was read). */
int fpa_kernel_tuning_applied(void);

/* Autotuner timing (no counterpart in the reference): launches variant `variant` of
 * op once untimed, then `reps` times between two HIP events on `stream`, waits, and
 * stores the mean kernel time in ms in *ms.  Returns 0, or the launch / HIP error. */
int fpa_kernel_time(int op, const fpa_arguments *args, int variant, int reps, void *stream, float *ms);

/* Largest count fpa_set_scalars_noalloc accepts. */
#define FPA_NOALLOC_MAX_VALUES (1u << 19)

/* fpa_set_ui / fpa_set_si / fpa_set_float / fpa_set_double (kind 0..3, as
 * fpa_scalar_encode) without allocating: no counterpart in the reference, whose set
 * calls cudaMalloc and cudaFree a staging buffer each time (fpa_dispatch.h:8-86).  For
 * callers that cannot allocate device memory (code under stream capture, tight serving
 * loops).  The 4- or 8-byte sources are staged in a fixed device buffer of
 * FPA_NOALLOC_MAX_VALUES values that each GPU holds from its first call until the
 * process ends; the call copies them there on the fpa stream, launches the same kernel
 * as fpa_set_*, and waits for it.  Results are identical to fpa_set_*.
 * Returns fpa_invalid_counts (with a message) above FPA_NOALLOC_MAX_VALUES values,
 * fpa_invalid_precisions for an unsupported precision, fpa_math_error for a bad kind.
 * All calls on one GPU share the staging buffer: calls that overlap on one GPU (from
 * several host threads) race and must be serialised by the caller. */
fpa_result fpa_set_scalars_noalloc(fpa_t r, int kind, const void *values);

/* Stream used by the fpa.h functions (a hipStream_t; NULL = null stream). */
void fpa_set_stream(void *stream);
void *fpa_get_stream(void);

/* 1 if precision (bits) is supported */
int fpa_precision_supported(uint32_t precision);

/* Diagnostics: throughput probe of the 32x32->64 multiply-add (v_mad_u64_u32) the
 * mul/div/sqrt kernels use.  blocks x 256 lanes each run iters*8 multiply-adds; the
 * kernel time in ms is stored in *ms.  Returns 0 on success. */
int fpa_ubench_mad(uint32_t blocks, uint32_t iters, float *ms);

/* Scalar helpers exported for host-side tests (same code as the kernels use). */
uint32_t fpa_scalar_round_inc(int mode, uint32_t sign, uint32_t lsb, uint32_t round_word, uint32_t sticky);
uint32_t fpa_scalar_recip_3by2(uint32_t d1, uint32_t d0);
uint32_t fpa_scalar_div3by2(uint32_t u2, uint32_t u1, uint32_t u0, uint32_t d1, uint32_t d0, uint32_t v);
uint32_t fpa_scalar_isqrt64(uint64_t u);
/* Exponent word (or class code) and left-aligned significand of one fpa_set_* source:
 * kind 0 uint64, 1 int64 (two's complement bits), 2 binary32 (bits in the low 32),
 * 3 binary64 -- the encoding the set kernels write. */
uint32_t fpa_scalar_encode(int kind, uint64_t bits, uint64_t *top);

#ifdef __cplusplus
}
#endif

#endif
