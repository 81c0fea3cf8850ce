/*
 * fpa_batch.h -- batched mixed-precision execution and fused dot-product chains.
 *
 * No counterpart in the reference, which launches one kernel per op and precision
 * (fpa_dispatch.h:175-421); these are the batched / chained forms of the same ops with
 * the same per-element semantics and rounding.
 *
 * Batch: a list of element-wise ops, each with its own op (add/sub/mul/div/sqrt),
 * precision (any fpa_precision_supported value) and rounding mode.  fpa_batch_create
 * groups the ops into (op, precision, mode) segments, allocates the operand and result
 * pool in device memory once, and captures one hipGraph: a fork, `branches` independent
 * chains of kernel nodes (one node per segment, segments dealt longest-first to the
 * least-loaded branch) and a join.  fpa_batch_run replays that graph with one launch.
 * Host operands and results are packed in op order: op i's exponent word at index i, its
 * limbs at offset sum_{j<i} precision_j/32.
 *
 * Dot chains: r[j] = x[o_j]*y[o_j] + ... + x[o_{j+1}-1]*y[o_{j+1}-1] where every
 * product and every partial sum is rounded in `mode`, accumulated in index order from
 * an exact zero, one launch for all chains.
 */
#ifndef FPA_BATCH_H
#define FPA_BATCH_H

#include <stddef.h>
#include <stdint.h>
#include "fpa_kernels.h"

#ifdef __cplusplus
extern "C" {
#endif

/* parallel branches of a batch graph made by fpa_batch_create (1..8 allowed) */
#define FPA_BATCH_DEFAULT_BRANCHES 4

/* kernel arguments of a persistent batch: the segment table (each segment's arguments and
 * its kind, the (op, precision) shape compiled into the kernel), the cost-sorted tiles
 * (segment, first value) and the work queue (next tile, waves done), in device memory */
#define FPA_PERSIST_MAX_SEGMENTS 40
typedef struct {
  fpa_arguments   seg[FPA_PERSIST_MAX_SEGMENTS];
  uint32_t        kind[FPA_PERSIST_MAX_SEGMENTS];
  const uint32_t *tiles;
  uint32_t       *queue;
  uint32_t        ntiles;
  uint32_t        pad;
} fpa_persist_arguments;

/* device-side arguments of the dot-chain kernel */
typedef struct {
  uint32_t        count;       /* number of chains */
  uint32_t        precision;
  int             mode;
  uint32_t        pad;
  uint32_t       *r_exps, *r_limbs;
  const uint32_t *x_exps, *x_limbs, *y_exps, *y_limbs;
  const uint64_t *offsets;     /* count+1 element offsets, device memory */
  const uint32_t *order;       /* chain processing order, device memory */
} fpa_dot_arguments;

typedef struct fpa_batch_s *fpa_batch_t;

/* 1 if `precision` can appear in a batch (same set as fpa_precision_supported) */
int fpa_batch_precision_supported(uint32_t precision);

/* Plan a batch of nops ops (op codes as fpa_op, modes as fpa_rounding_mode).
 * Returns fpa_invalid_precisions / fpa_invalid_counts / fpa_math_error (bad op). */
fpa_result fpa_batch_create(fpa_batch_t *batch, uint32_t nops, const uint8_t *ops, const uint32_t *precisions,
                            const uint8_t *modes);
/* Same with an explicit branch count (1..8; fpa_invalid_counts otherwise). */
fpa_result fpa_batch_create_branches(fpa_batch_t *batch, uint32_t nops, const uint8_t *ops,
                                     const uint32_t *precisions, const uint8_t *modes, uint32_t branches);
/* The same batch as one persistent kernel launch captured in a hipGraph: waves take
 * one-wave tiles of the segments from a self-resetting work queue, most expensive first,
 * each with its (op, precision) kind's shape (precisions 256, 512, 1024, 2048, 4096 and
 * 8192 only; fpa_invalid_precisions otherwise; at most FPA_PERSIST_MAX_SEGMENTS (op,
 * precision, mode) segments, fpa_invalid_counts otherwise).  Results equal
 * fpa_batch_create's bit for bit.  Runs of one batch must not overlap (they share its
 * pool and queue). */
fpa_result fpa_batch_create_persistent(fpa_batch_t *batch, uint32_t nops, const uint8_t *ops,
                                       const uint32_t *precisions, const uint8_t *modes);
/* 1 if the batch runs as one persistent kernel; its tiles and grid (blocks) */
int fpa_batch_is_persistent(fpa_batch_t batch);
uint32_t fpa_batch_tiles(fpa_batch_t batch);
uint32_t fpa_batch_grid(fpa_batch_t batch);

/* Upload operands packed in op order (b may be NULL if no op is binary). */
fpa_result fpa_batch_set_operands(fpa_batch_t batch, const uint32_t *a_exps, const uint32_t *a_limbs,
                                  const uint32_t *b_exps, const uint32_t *b_limbs);
/* Replay the captured graph (one graph launch) on `stream` (hipStream_t). */
fpa_result fpa_batch_run(fpa_batch_t batch, void *stream);
/* Same kernels launched one after another on `stream`, no graph (comparison path). */
fpa_result fpa_batch_run_plain(fpa_batch_t batch, void *stream);
/* Synchronise `stream` and copy the results back, packed in op order. */
fpa_result fpa_batch_get_results(fpa_batch_t batch, void *stream, uint32_t *r_exps, uint32_t *r_limbs);
fpa_result fpa_batch_destroy(fpa_batch_t batch);

/* Plan facts: segments (= kernel nodes), graph branches, workgroups over all nodes,
 * device bytes held. */
uint32_t fpa_batch_segments(fpa_batch_t batch);
uint32_t fpa_batch_branches(fpa_batch_t batch);
uint64_t fpa_batch_blocks(fpa_batch_t batch);
size_t   fpa_batch_device_bytes(fpa_batch_t batch);
/* total limbs of the packed host arrays (sum of precision/32 over ops) */
uint64_t fpa_batch_total_limbs(fpa_batch_t batch);

/* Dot chains over device arrays: x and y hold all elements, r gets `count` results;
 * lengths[j] = elements in chain j (host array).  Same precision everywhere, any
 * fpa_precision_supported value (fpa_invalid_precisions otherwise); the kernel shape is
 * the dispatcher's FPA_KERNEL_DOT variant for that precision.
 * Returns fpa_invalid_counts if x/y counts differ from sum(lengths) or r->count != count.
 * Synchronous (plans, runs on the fpa stream, waits, frees the plan). */
fpa_result fpa_dot_chains(fpa_t r, fpa_t x, fpa_t y, const uint32_t *lengths, uint32_t count,
                          fpa_rounding_mode mode);

/* Reusable form: the chain offsets and the longest-first order stay in device memory. */
typedef struct fpa_dot_plan_s *fpa_dot_plan_t;
fpa_result fpa_dot_plan_create(fpa_dot_plan_t *plan, const uint32_t *lengths, uint32_t count);
fpa_result fpa_dot_plan_run(fpa_dot_plan_t plan, fpa_t r, fpa_t x, fpa_t y, fpa_rounding_mode mode, void *stream);
fpa_result fpa_dot_plan_destroy(fpa_dot_plan_t plan);
/* Autotuner timing: one untimed run, then `reps` runs between two HIP events on
 * `stream`; mean ms per run in *ms.  Returns 0, or an fpa_result / -1. */
int fpa_dot_plan_time(fpa_dot_plan_t plan, fpa_t r, fpa_t x, fpa_t y, fpa_rounding_mode mode, int reps,
                      void *stream, float *ms);

#ifdef __cplusplus
}
#endif

#endif
