/*
 * Host side of fpa_batch.h: batch planning (segments, one operand/result pool), hipGraph
 * capture of either a fork/join graph of per-segment kernels or one persistent kernel
 * launch over cost-sorted tiles, and dot-chain plans.
 *
 * A batch becomes one graph: a fork node, `branches` independent chains of kernel nodes
 * and a join.  Every kernel node is the element-wise kernel of one (op, precision, mode)
 * segment, launched with the dispatcher's shape for that (op, precision).  Segments are
 * dealt to the branches longest-processing-time first (estimated cost = values x per-value
 * cost of the shape), so the branches end at about the same time and the short segments
 * fill the gaps the long ones leave on the chip.  Nothing in the graph needs resetting
 * between replays.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <hip/hip_runtime_api.h>

#include "../../include/fpa_batch.h"
#include "fpa_scalar.h"

#define FPA_HIP(call)                                                                      \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess) {                                                                \
      fprintf(stderr, "\nCall \"%s\" failed from %s, line %d: %s\n", #call, __FILE__,     \
              __LINE__, hipGetErrorString(e_));                                            \
      exit(1);                                                                             \
    }                                                                                      \
  } while (0)

int fpa_impl_dot_launch(const fpa_dot_arguments *d, int K, int L, void *stream);
int fpa_impl_persist_kind(int op, uint32_t precision, int *K, int *L);
int fpa_impl_persist_blocks_per_cu(void);
int fpa_impl_mul_rounds(int K, int L);
int fpa_impl_persist_launch(const fpa_persist_arguments *P, uint32_t grid, void *stream);
void *fpa_get_stream(void);

#define FPA_BATCH_MAX_BRANCHES 8

typedef struct {
  int op, mode, K, L, variant, branch;
  uint32_t precision, count, blocks;
  double cost;
  size_t off[6];           /* pool offsets: r_exps, r_limbs, a_exps, a_limbs, b_exps, b_limbs */
  fpa_arguments args;      /* device pointers into the pool */
} seg_host;

struct fpa_batch_s {
  uint32_t nops, nseg, branches;
  uint64_t blocks;
  uint32_t *op_seg, *op_slot;
  uint64_t *op_off;        /* packed host limb offset of each op */
  uint64_t total_limbs;
  seg_host *sh;
  uint32_t *order;         /* segments, most expensive first */
  uint8_t *pool;           /* device */
  size_t pool_bytes;
  uint8_t *mirror;         /* pinned host copy of the pool layout */
  hipGraph_t graph;
  hipGraphExec_t exec;
  /* persistent form: kernel arguments (segment table, tiles, queue), grid */
  int persistent;
  uint32_t ntiles, grid;
  fpa_persist_arguments pargs;
  uint32_t *tiles_dev;
  uint32_t *queue_dev;
};

int fpa_batch_precision_supported(uint32_t precision) { return fpa_precision_supported(precision); }

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

/* relative time one wavefront spends on its 64/L values of a segment (mul: product
 * rounds of KD x KD multiply-adds per lane, fewer than L where the shape forms a short
 * product) */
static double wave_cost(int op, int K, int L) {
  int n = K * L, KD = (32 * K + 26) / 27;
  switch (op) {
    case FPA_OP_MUL: {
      const int rounds = fpa_impl_mul_rounds(K, L);
      return (double)(rounds > 0 ? rounds : L) * KD * KD + 40.0 * KD;
    }
    case FPA_OP_DIV: return (double)(n + 1) * (3 * K + 40);
    case FPA_OP_SQRT: return (double)n * (3 * K + 50);
    default: return 8.0 * K + 60.0;
  }
}

typedef struct { double cost; uint32_t seg; } seg_cost;
static int seg_cmp(const void *x, const void *y) {
  const seg_cost *a = (const seg_cost *)x, *b = (const seg_cost *)y;
  if (a->cost != b->cost) return a->cost > b->cost ? -1 : 1;
  return a->seg < b->seg ? -1 : (a->seg > b->seg);
}

static int launch_segment(const seg_host *g, void *stream) {
  return fpa_kernel_launch(g->op, &g->args, g->variant, stream);
}

/* Segments keyed by (op, precision, mode) in first-appearance order, the device pool
 * (per segment: results, a, b) and its pinned host mirror, each segment's kernel
 * arguments.  shape(op, precision, count, &K, &L) gives a segment's shape and returns
 * its variant or kind (-1: unsupported precision). */
typedef int (*shape_fn)(int op, uint32_t precision, uint32_t count, int *K, int *L);

static fpa_result plan_batch(struct fpa_batch_s **out, uint32_t nops, const uint8_t *ops, const uint32_t *precisions,
                             const uint8_t *modes, shape_fn shape) {
  *out = NULL;
  for (uint32_t i = 0; i < nops; i++) {
    if (ops[i] < FPA_OP_ADD || ops[i] > FPA_OP_SQRT || modes[i] > 4) return fpa_math_error;
    if (!fpa_batch_precision_supported(precisions[i])) return fpa_invalid_precisions;
    int K, L;
    if (shape(ops[i], precisions[i], 1, &K, &L) < 0) return fpa_invalid_precisions;
  }
  struct fpa_batch_s *b = (struct fpa_batch_s *)calloc(1, sizeof *b);
  b->nops = nops;
  b->op_seg = (uint32_t *)malloc(sizeof(uint32_t) * (nops + 1));
  b->op_slot = (uint32_t *)malloc(sizeof(uint32_t) * (nops + 1));
  b->op_off = (uint64_t *)malloc(sizeof(uint64_t) * (nops + 1));
  int cap = 16;
  b->sh = (seg_host *)calloc(cap, sizeof(seg_host));
  uint64_t off = 0;
  for (uint32_t i = 0; i < nops; i++) {
    uint32_t s;
    for (s = 0; s < b->nseg; s++)
      if (b->sh[s].op == ops[i] && b->sh[s].precision == precisions[i] && b->sh[s].mode == modes[i]) break;
    if (s == b->nseg) {
      if ((int)b->nseg == cap) { cap *= 2; b->sh = (seg_host *)realloc(b->sh, cap * sizeof(seg_host)); }
      memset(&b->sh[s], 0, sizeof(seg_host));
      b->sh[s].op = ops[i];
      b->sh[s].precision = precisions[i];
      b->sh[s].mode = modes[i];
      b->nseg++;
    }
    b->op_seg[i] = s;
    b->op_slot[i] = b->sh[s].count++;
    b->op_off[i] = off;
    off += precisions[i] / 32;
  }
  b->op_off[nops] = off;
  b->total_limbs = off;
  /* pool layout, then the kernel arguments of every segment */
  size_t pos = 0;
  for (uint32_t s = 0; s < b->nseg; s++) {
    seg_host *g = &b->sh[s];
    size_t ne = (size_t)g->count * 4, nl = (size_t)g->count * (g->precision / 32) * 4;
    int binary = g->op != FPA_OP_SQRT;
    size_t sizes[6] = { ne, nl, ne, nl, binary ? ne : 0, binary ? nl : 0 };
    for (int k = 0; k < 6; k++) { g->off[k] = pos; pos += align256(sizes[k] + 16); }
  }
  b->pool_bytes = pos ? pos : 256;
  FPA_HIP(hipMalloc((void **)&b->pool, b->pool_bytes));
  FPA_HIP(hipHostMalloc((void **)&b->mirror, b->pool_bytes, hipHostMallocDefault));
  for (uint32_t s = 0; s < b->nseg; s++) {
    seg_host *g = &b->sh[s];
    g->variant = shape(g->op, g->precision, g->count, &g->K, &g->L);
    fpa_arguments *a = &g->args;
    a->count = g->count;
    a->precision = g->precision;
    a->mode = (fpa_rounding_mode)g->mode;
    a->r_exps = (uint32_t *)(b->pool + g->off[0]);
    a->r_limbs = (uint32_t *)(b->pool + g->off[1]);
    a->a_exps = (uint32_t *)(b->pool + g->off[2]);
    a->a_limbs = (uint32_t *)(b->pool + g->off[3]);
    a->b_exps = (uint32_t *)(b->pool + (g->op == FPA_OP_SQRT ? g->off[2] : g->off[4]));
    a->b_limbs = (uint32_t *)(b->pool + (g->op == FPA_OP_SQRT ? g->off[3] : g->off[5]));
    uint32_t vpb = 4 * (64 / g->L);
    g->blocks = (g->count + vpb - 1) / vpb;
    b->blocks += g->blocks;
    g->cost = (double)g->blocks * 4 * wave_cost(g->op, g->K, g->L);
  }
  *out = b;
  return fpa_success;
}

/* A segment of the graph form shares the chip with the branches running beside it, so it
 * takes the shape tuned for a launch twice its size: on the 2^18-op mixed batch this beat
 * the shape for its own size (the isolated-latency winner, which takes chip share from the
 * other branches), four times its size and the large-launch shapes (0.445 ms against 0.50,
 * 0.49 and 0.57 ms). */
static int graph_shape(int op, uint32_t precision, uint32_t count, int *K, int *L) {
  int v = fpa_kernel_get_variant_for_count(op, precision, 2 * (uint64_t)count);
  if (v < 0 || fpa_kernel_variant_shape(op, precision, v, K, L) != 0) return -1;
  return v;
}

static int persist_shape(int op, uint32_t precision, uint32_t count, int *K, int *L) {
  (void)count;
  return fpa_impl_persist_kind(op, precision, K, L);
}

fpa_result fpa_batch_create_branches(fpa_batch_t *out, uint32_t nops, const uint8_t *ops,
                                     const uint32_t *precisions, const uint8_t *modes, uint32_t branches) {
  *out = NULL;
  if (branches < 1 || branches > FPA_BATCH_MAX_BRANCHES) return fpa_invalid_counts;
  struct fpa_batch_s *b;
  fpa_result pr = plan_batch(&b, nops, ops, precisions, modes, graph_shape);
  if (pr != fpa_success) return pr;
  /* longest processing time first onto the least-loaded branch */
  b->order = (uint32_t *)malloc(sizeof(uint32_t) * (b->nseg ? b->nseg : 1));
  seg_cost *sc = (seg_cost *)malloc(sizeof(seg_cost) * (b->nseg ? b->nseg : 1));
  for (uint32_t s = 0; s < b->nseg; s++) { sc[s].cost = b->sh[s].cost; sc[s].seg = s; }
  qsort(sc, b->nseg, sizeof(seg_cost), seg_cmp);
  for (uint32_t s = 0; s < b->nseg; s++) b->order[s] = sc[s].seg;
  free(sc);
  b->branches = branches < b->nseg ? branches : (b->nseg ? b->nseg : 1);
  double load[FPA_BATCH_MAX_BRANCHES] = { 0 };
  for (uint32_t k = 0; k < b->nseg; k++) {
    seg_host *g = &b->sh[b->order[k]];
    uint32_t best = 0;
    for (uint32_t j = 1; j < b->branches; j++)
      if (load[j] < load[best]) best = j;
    g->branch = (int)best;
    load[best] += g->cost;
  }
  /* capture: fork on an event, one stream per branch, join back onto the origin stream */
  hipStream_t st[FPA_BATCH_MAX_BRANCHES];
  hipEvent_t ev[FPA_BATCH_MAX_BRANCHES + 1];
  for (uint32_t j = 0; j < b->branches; j++) FPA_HIP(hipStreamCreateWithFlags(&st[j], hipStreamNonBlocking));
  for (uint32_t j = 0; j <= b->branches; j++) FPA_HIP(hipEventCreateWithFlags(&ev[j], hipEventDisableTiming));
  FPA_HIP(hipStreamBeginCapture(st[0], hipStreamCaptureModeThreadLocal));
  FPA_HIP(hipEventRecord(ev[b->branches], st[0]));
  for (uint32_t j = 1; j < b->branches; j++) FPA_HIP(hipStreamWaitEvent(st[j], ev[b->branches], 0));
  int rc = 0;
  for (uint32_t k = 0; k < b->nseg && rc == 0; k++) {
    seg_host *g = &b->sh[b->order[k]];
    rc = launch_segment(g, st[g->branch]);
  }
  for (uint32_t j = 1; j < b->branches; j++) {
    FPA_HIP(hipEventRecord(ev[j], st[j]));
    FPA_HIP(hipStreamWaitEvent(st[0], ev[j], 0));
  }
  FPA_HIP(hipStreamEndCapture(st[0], &b->graph));
  if (rc != 0) { fprintf(stderr, "batch graph capture failed: kernel launch returned %d\n", rc); exit(1); }
  FPA_HIP(hipGraphInstantiate(&b->exec, b->graph, NULL, NULL, 0));
  for (uint32_t j = 0; j <= b->branches; j++) FPA_HIP(hipEventDestroy(ev[j]));
  for (uint32_t j = 0; j < b->branches; j++) FPA_HIP(hipStreamDestroy(st[j]));
  *out = b;
  return fpa_success;
}

fpa_result fpa_batch_create(fpa_batch_t *out, uint32_t nops, const uint8_t *ops, const uint32_t *precisions,
                            const uint8_t *modes) {
  return fpa_batch_create_branches(out, nops, ops, precisions, modes, FPA_BATCH_DEFAULT_BRANCHES);
}

typedef struct { double cost; uint32_t seg, first; } tile_host;
static int tile_cmp(const void *x, const void *y) {
  const tile_host *a = (const tile_host *)x, *b = (const tile_host *)y;
  if (a->cost != b->cost) return a->cost > b->cost ? -1 : 1;
  if (a->seg != b->seg) return a->seg < b->seg ? -1 : 1;
  return a->first < b->first ? -1 : (a->first > b->first);
}

fpa_result fpa_batch_create_persistent(fpa_batch_t *out, uint32_t nops, const uint8_t *ops,
                                       const uint32_t *precisions, const uint8_t *modes) {
  struct fpa_batch_s *b;
  fpa_result pr = plan_batch(&b, nops, ops, precisions, modes, persist_shape);
  if (pr != fpa_success) return pr;
  b->persistent = 1;
  b->branches = 1;
  if (b->nseg > FPA_PERSIST_MAX_SEGMENTS) {
    fpa_batch_destroy(b);
    return fpa_invalid_counts;
  }
  /* the segment table, carried in the kernel arguments */
  memset(&b->pargs, 0, sizeof b->pargs);
  for (uint32_t s = 0; s < b->nseg; s++) {
    b->pargs.seg[s] = b->sh[s].args;
    b->pargs.kind[s] = (uint32_t)b->sh[s].variant;
  }
  /* tiles: one wavefront's values of one segment, most expensive first */
  uint64_t nt = 0;
  for (uint32_t s = 0; s < b->nseg; s++) nt += (b->sh[s].count + (64 / b->sh[s].L) - 1) / (64 / b->sh[s].L);
  tile_host *th = (tile_host *)malloc(sizeof(tile_host) * (nt ? nt : 1));
  uint64_t k = 0;
  for (uint32_t s = 0; s < b->nseg; s++) {
    const seg_host *g = &b->sh[s];
    const uint32_t vpw = 64 / g->L;
    const double c = wave_cost(g->op, g->K, g->L);
    for (uint32_t f = 0; f < g->count; f += vpw) { th[k].cost = c; th[k].seg = s; th[k].first = f; k++; }
  }
  qsort(th, nt, sizeof(tile_host), tile_cmp);
  uint32_t *tt = (uint32_t *)malloc(sizeof(uint32_t) * 2 * (nt ? nt : 1));
  for (uint64_t i = 0; i < nt; i++) { tt[2 * i] = th[i].seg; tt[2 * i + 1] = th[i].first; }
  free(th);
  b->ntiles = (uint32_t)nt;
  FPA_HIP(hipMalloc((void **)&b->tiles_dev, sizeof(uint32_t) * 2 * (nt ? nt : 1)));
  FPA_HIP(hipMemcpy(b->tiles_dev, tt, sizeof(uint32_t) * 2 * nt, hipMemcpyHostToDevice));
  free(tt);
  /* the work queue: next tile, waves done; zero here, and set back to zero by the last
   * wave of every run */
  FPA_HIP(hipMalloc((void **)&b->queue_dev, 256));
  FPA_HIP(hipMemset(b->queue_dev, 0, 256));
  b->pargs.tiles = b->tiles_dev;
  b->pargs.queue = b->queue_dev;
  b->pargs.ntiles = b->ntiles;
  /* grid: every resident block slot of the device, no more blocks than the tiles need */
  int dev = 0, cus = 0;
  FPA_HIP(hipGetDevice(&dev));
  FPA_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  uint64_t grid = (uint64_t)cus * fpa_impl_persist_blocks_per_cu(), need = (nt + 3) / 4;
  b->grid = (uint32_t)(grid < need ? grid : (need ? need : 1));
  FPA_HIP(hipDeviceSynchronize());
  /* capture the one launch */
  hipStream_t cs;
  FPA_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
  FPA_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
  int rc = fpa_impl_persist_launch(&b->pargs, b->grid, cs);
  FPA_HIP(hipStreamEndCapture(cs, &b->graph));
  if (rc != 0) { fprintf(stderr, "persistent batch capture failed: kernel launch returned %d\n", rc); exit(1); }
  FPA_HIP(hipGraphInstantiate(&b->exec, b->graph, NULL, NULL, 0));
  FPA_HIP(hipStreamDestroy(cs));
  *out = b;
  return fpa_success;
}

static void scatter_operand(struct fpa_batch_s *b, int which, const uint32_t *exps, const uint32_t *limbs) {
  for (uint32_t i = 0; i < b->nops; i++) {
    seg_host *g = &b->sh[b->op_seg[i]];
    if (which == 1 && g->op == FPA_OP_SQRT) continue;
    uint32_t n = g->precision / 32, slot = b->op_slot[i];
    uint32_t *e = (uint32_t *)(b->mirror + g->off[2 + 2 * which]);
    uint32_t *l = (uint32_t *)(b->mirror + g->off[3 + 2 * which]);
    e[slot] = exps[i];
    memcpy(l + (size_t)slot * n, limbs + b->op_off[i], (size_t)n * 4);
  }
}

fpa_result fpa_batch_set_operands(fpa_batch_t b, const uint32_t *a_exps, const uint32_t *a_limbs,
                                  const uint32_t *b_exps, const uint32_t *b_limbs) {
  int binary = 0;
  for (uint32_t s = 0; s < b->nseg; s++) binary |= b->sh[s].op != FPA_OP_SQRT;
  if (binary && (!b_exps || !b_limbs)) return fpa_invalid_counts;
  scatter_operand(b, 0, a_exps, a_limbs);
  if (binary) scatter_operand(b, 1, b_exps, b_limbs);
  for (uint32_t s = 0; s < b->nseg; s++) {
    seg_host *g = &b->sh[s];
    size_t ne = (size_t)g->count * 4, nl = ne * (g->precision / 32);
    for (int k = 2; k < (g->op == FPA_OP_SQRT ? 4 : 6); k++) {
      size_t bytes = (k & 1) ? nl : ne;
      FPA_HIP(hipMemcpy(b->pool + g->off[k], b->mirror + g->off[k], bytes, hipMemcpyHostToDevice));
    }
  }
  return fpa_success;
}

fpa_result fpa_batch_run(fpa_batch_t b, void *stream) {
  FPA_HIP(hipGraphLaunch(b->exec, (hipStream_t)stream));
  return fpa_success;
}

fpa_result fpa_batch_run_plain(fpa_batch_t b, void *stream) {
  if (b->persistent) {
    int rc = fpa_impl_persist_launch(&b->pargs, b->grid, stream);
    if (rc != 0) { fprintf(stderr, "persistent batch launch failed: %d\n", rc); exit(1); }
    return fpa_success;
  }
  for (uint32_t k = 0; k < b->nseg; k++) {
    int rc = launch_segment(&b->sh[b->order[k]], stream);
    if (rc != 0) { fprintf(stderr, "plain launch of segment %u failed: %d\n", b->order[k], rc); exit(1); }
  }
  return fpa_success;
}

fpa_result fpa_batch_get_results(fpa_batch_t b, void *stream, uint32_t *r_exps, uint32_t *r_limbs) {
  FPA_HIP(hipStreamSynchronize((hipStream_t)stream));
  for (uint32_t s = 0; s < b->nseg; s++) {
    seg_host *g = &b->sh[s];
    size_t ne = (size_t)g->count * 4, nl = ne * (g->precision / 32);
    FPA_HIP(hipMemcpy(b->mirror + g->off[0], b->pool + g->off[0], ne, hipMemcpyDeviceToHost));
    FPA_HIP(hipMemcpy(b->mirror + g->off[1], b->pool + g->off[1], nl, hipMemcpyDeviceToHost));
  }
  for (uint32_t i = 0; i < b->nops; i++) {
    seg_host *g = &b->sh[b->op_seg[i]];
    uint32_t n = g->precision / 32, slot = b->op_slot[i];
    r_exps[i] = ((uint32_t *)(b->mirror + g->off[0]))[slot];
    memcpy(r_limbs + b->op_off[i], (uint32_t *)(b->mirror + g->off[1]) + (size_t)slot * n, (size_t)n * 4);
  }
  return fpa_success;
}

fpa_result fpa_batch_destroy(fpa_batch_t b) {
  if (!b) return fpa_success;
  if (b->exec) FPA_HIP(hipGraphExecDestroy(b->exec));
  if (b->graph) FPA_HIP(hipGraphDestroy(b->graph));
  FPA_HIP(hipFree(b->pool));
  FPA_HIP(hipHostFree(b->mirror));
  if (b->persistent) {
    if (b->tiles_dev) FPA_HIP(hipFree(b->tiles_dev));
    if (b->queue_dev) FPA_HIP(hipFree(b->queue_dev));
  }
  free(b->op_seg);
  free(b->op_slot);
  free(b->op_off);
  free(b->order);
  free(b->sh);
  free(b);
  return fpa_success;
}

uint32_t fpa_batch_segments(fpa_batch_t b) { return b->nseg; }
uint32_t fpa_batch_branches(fpa_batch_t b) { return b->branches; }
uint64_t fpa_batch_blocks(fpa_batch_t b) { return b->blocks; }
size_t fpa_batch_device_bytes(fpa_batch_t b) { return b->pool_bytes; }
uint64_t fpa_batch_total_limbs(fpa_batch_t b) { return b->total_limbs; }
int fpa_batch_is_persistent(fpa_batch_t b) { return b->persistent; }
uint32_t fpa_batch_tiles(fpa_batch_t b) { return b->ntiles; }
uint32_t fpa_batch_grid(fpa_batch_t b) { return b->grid; }

/* ------------------------------ dot chains ------------------------------ */
struct fpa_dot_plan_s {
  uint32_t count;
  uint64_t total;
  uint64_t *offsets_dev;
  uint32_t *order_dev;
};

typedef struct { uint32_t len, idx; } len_idx;
static int len_cmp(const void *x, const void *y) {
  const len_idx *a = (const len_idx *)x, *b = (const len_idx *)y;
  if (a->len != b->len) return a->len > b->len ? -1 : 1;
  return a->idx < b->idx ? -1 : (a->idx > b->idx);
}

fpa_result fpa_dot_plan_create(fpa_dot_plan_t *out, const uint32_t *lengths, uint32_t count) {
  struct fpa_dot_plan_s *p = (struct fpa_dot_plan_s *)calloc(1, sizeof *p);
  p->count = count;
  uint64_t *off = (uint64_t *)malloc(sizeof(uint64_t) * (count + 1));
  len_idx *li = (len_idx *)malloc(sizeof(len_idx) * (count ? count : 1));
  uint32_t *ord = (uint32_t *)malloc(sizeof(uint32_t) * (count ? count : 1));
  off[0] = 0;
  for (uint32_t j = 0; j < count; j++) { off[j + 1] = off[j] + lengths[j]; li[j].len = lengths[j]; li[j].idx = j; }
  qsort(li, count, sizeof(len_idx), len_cmp);
  for (uint32_t j = 0; j < count; j++) ord[j] = li[j].idx;
  p->total = off[count];
  FPA_HIP(hipMalloc((void **)&p->offsets_dev, sizeof(uint64_t) * (count + 1)));
  FPA_HIP(hipMalloc((void **)&p->order_dev, sizeof(uint32_t) * (count ? count : 1)));
  FPA_HIP(hipMemcpy(p->offsets_dev, off, sizeof(uint64_t) * (count + 1), hipMemcpyHostToDevice));
  FPA_HIP(hipMemcpy(p->order_dev, ord, sizeof(uint32_t) * count, hipMemcpyHostToDevice));
  free(off);
  free(li);
  free(ord);
  *out = p;
  return fpa_success;
}

fpa_result fpa_dot_plan_run(fpa_dot_plan_t p, fpa_t r, fpa_t x, fpa_t y, fpa_rounding_mode mode, void *stream) {
  if (x->precision != y->precision || r->precision != x->precision) return fpa_invalid_precisions;
  if (!fpa_precision_supported(x->precision)) return fpa_invalid_precisions;
  if (x->count != y->count || (uint64_t)x->count != p->total || r->count != p->count) return fpa_invalid_counts;
  fpa_dot_arguments d;
  memset(&d, 0, sizeof d);
  d.count = p->count;
  d.precision = x->precision;
  d.mode = (int)mode;
  d.r_exps = r->gpu_exps;
  d.r_limbs = r->gpu_limbs;
  d.x_exps = x->gpu_exps;
  d.x_limbs = x->gpu_limbs;
  d.y_exps = y->gpu_exps;
  d.y_limbs = y->gpu_limbs;
  d.offsets = p->offsets_dev;
  d.order = p->order_dev;
  int K, L;
  fpa_kernel_variant_shape(FPA_KERNEL_DOT, x->precision,
                           fpa_kernel_get_variant_for_count(FPA_KERNEL_DOT, x->precision, p->count), &K, &L);
  int rc = fpa_impl_dot_launch(&d, K, L, stream);
  if (rc != 0) { fprintf(stderr, "dot kernel launch failed: %d\n", rc); exit(1); }
  return fpa_success;
}

int fpa_dot_plan_time(fpa_dot_plan_t p, fpa_t r, fpa_t x, fpa_t y, fpa_rounding_mode mode, int reps, void *stream,
                      float *ms) {
  if (reps < 1 || !ms) return -1;
  fpa_result rc = fpa_dot_plan_run(p, r, x, y, mode, stream);
  if (rc != fpa_success) return (int)rc;
  hipEvent_t e0, e1;
  FPA_HIP(hipEventCreate(&e0));
  FPA_HIP(hipEventCreate(&e1));
  FPA_HIP(hipEventRecord(e0, (hipStream_t)stream));
  for (int i = 0; i < reps; i++) fpa_dot_plan_run(p, r, x, y, mode, stream);
  FPA_HIP(hipEventRecord(e1, (hipStream_t)stream));
  FPA_HIP(hipEventSynchronize(e1));
  float t = 0.0f;
  FPA_HIP(hipEventElapsedTime(&t, e0, e1));
  FPA_HIP(hipEventDestroy(e0));
  FPA_HIP(hipEventDestroy(e1));
  *ms = t / (float)reps;
  return 0;
}

fpa_result fpa_dot_plan_destroy(fpa_dot_plan_t p) {
  if (!p) return fpa_success;
  FPA_HIP(hipFree(p->offsets_dev));
  FPA_HIP(hipFree(p->order_dev));
  free(p);
  return fpa_success;
}

fpa_result fpa_dot_chains(fpa_t r, fpa_t x, fpa_t y, const uint32_t *lengths, uint32_t count, fpa_rounding_mode mode) {
  uint64_t total = 0;
  for (uint32_t j = 0; j < count; j++) total += lengths[j];
  if (x->precision != y->precision || r->precision != x->precision) return fpa_invalid_precisions;
  if (!fpa_precision_supported(x->precision)) return fpa_invalid_precisions;
  if (x->count != y->count || (uint64_t)x->count != total || r->count != count) return fpa_invalid_counts;
  fpa_dot_plan_t p;
  fpa_dot_plan_create(&p, lengths, count);
  void *st = fpa_get_stream();
  fpa_result rc = fpa_dot_plan_run(p, r, x, y, mode, st);
  FPA_HIP(hipStreamSynchronize((hipStream_t)st));
  fpa_dot_plan_destroy(p);
  return rc;
}
