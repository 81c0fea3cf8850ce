/*
 * Continuous-batching request server (fpa_serve.h) and the plain per-request path.
 *
 * Device memory: one pool of the budget, allocated at creation; requests get regions
 * from a first-fit free list (256-byte granules, neighbours coalesced on free).
 * A request's region holds, per (op, mode) group, its results first and then its
 * operands, so one host->device copy uploads all operands and one device->host copy
 * returns all results.  Each slot has a stream, a pinned staging buffer and a worker
 * thread; one scheduler thread admits queued requests into free slots.  The queue, the
 * pool and the slot states are guarded by one mutex.
 */
#define _GNU_SOURCE
#include <pthread.h>
#include <sched.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>
#include <hip/hip_runtime_api.h>

#include "../../include/fpa_serve.h"

#define FPA_HIP(call)                                                                      \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess) {                                                                \
      fprintf(stderr, "\nCall \"%s\" failed from %s, line %d: %s\n", #call, __FILE__,     \
              __LINE__, hipGetErrorString(e_));                                            \
      exit(1);                                                                             \
    }                                                                                      \
  } while (0)

static double mono_s(void) {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

static size_t a16(size_t x) { return (x + 15) & ~(size_t)15; }

/* ------------------------------ request layout ------------------------------ */
#define MAX_SEG 25

typedef struct {
  int op, mode;
  uint32_t count;
  size_t r_e, r_l, a_e, a_l, b_e, b_l;
} seg_t;

typedef struct {
  int nseg;
  seg_t seg[MAX_SEG];
  size_t r_bytes, in_bytes;
  uint8_t *op_seg;
  uint32_t *op_slot;
  uint32_t cap;
} layout_t;

static void layout_reserve(layout_t *L, uint32_t nops) {
  if (L->cap >= nops) return;
  free(L->op_seg);
  free(L->op_slot);
  L->cap = nops < 64 ? 64 : nops;
  L->op_seg = (uint8_t *)malloc(L->cap);
  L->op_slot = (uint32_t *)malloc(sizeof(uint32_t) * L->cap);
}

static void layout_free(layout_t *L) {
  free(L->op_seg);
  free(L->op_slot);
  L->op_seg = NULL;
  L->op_slot = NULL;
  L->cap = 0;
}

/* groups the request's ops by (op, mode) and lays out results, then operands */
static int build_layout(const fpa_request *q, layout_t *L) {
  if (!fpa_precision_supported(q->precision)) return fpa_invalid_precisions;
  if (q->nops == 0) return fpa_invalid_counts;
  layout_reserve(L, q->nops);
  uint32_t n = q->precision / 32;
  L->nseg = 0;
  int binary = 0;
  for (uint32_t i = 0; i < q->nops; i++) {
    int op = q->ops[i], mode = q->modes[i];
    if (op < FPA_OP_ADD || op > FPA_OP_SQRT || mode < 0 || mode > 4) return fpa_math_error;
    binary |= op != FPA_OP_SQRT;
    int s;
    for (s = 0; s < L->nseg; s++)
      if (L->seg[s].op == op && L->seg[s].mode == mode) break;
    if (s == L->nseg) {
      memset(&L->seg[s], 0, sizeof(seg_t));
      L->seg[s].op = op;
      L->seg[s].mode = mode;
      L->nseg++;
    }
    L->op_seg[i] = (uint8_t)s;
    L->op_slot[i] = L->seg[s].count++;
  }
  if (binary && (!q->b_exps || !q->b_limbs)) return fpa_invalid_counts;
  size_t pos = 0;
  for (int s = 0; s < L->nseg; s++) {
    seg_t *g = &L->seg[s];
    g->r_e = pos; pos += a16((size_t)g->count * 4);
    g->r_l = pos; pos += a16((size_t)g->count * n * 4);
  }
  L->r_bytes = pos;
  for (int s = 0; s < L->nseg; s++) {
    seg_t *g = &L->seg[s];
    g->a_e = pos; pos += a16((size_t)g->count * 4);
    g->a_l = pos; pos += a16((size_t)g->count * n * 4);
    if (g->op != FPA_OP_SQRT) {
      g->b_e = pos; pos += a16((size_t)g->count * 4);
      g->b_l = pos; pos += a16((size_t)g->count * n * 4);
    } else {
      g->b_e = g->a_e;
      g->b_l = g->a_l;
    }
  }
  L->in_bytes = pos - L->r_bytes;
  return 0;
}

size_t fpa_request_bytes(const fpa_request *req) {
  layout_t L;
  memset(&L, 0, sizeof L);
  size_t b = build_layout(req, &L) ? 0 : L.r_bytes + L.in_bytes;
  layout_free(&L);
  return b;
}

/* operands into the staging image of the region */
static void scatter_inputs(const fpa_request *q, const layout_t *L, uint8_t *img) {
  uint32_t n = q->precision / 32;
  for (uint32_t i = 0; i < q->nops; i++) {
    const seg_t *g = &L->seg[L->op_seg[i]];
    uint32_t k = L->op_slot[i];
    ((uint32_t *)(img + g->a_e))[k] = q->a_exps[i];
    memcpy((uint32_t *)(img + g->a_l) + (size_t)k * n, q->a_limbs + (size_t)i * n, (size_t)n * 4);
    if (g->op != FPA_OP_SQRT) {
      ((uint32_t *)(img + g->b_e))[k] = q->b_exps[i];
      memcpy((uint32_t *)(img + g->b_l) + (size_t)k * n, q->b_limbs + (size_t)i * n, (size_t)n * 4);
    }
  }
}

static void gather_results(fpa_request *q, const layout_t *L, const uint8_t *img) {
  uint32_t n = q->precision / 32;
  for (uint32_t i = 0; i < q->nops; i++) {
    const seg_t *g = &L->seg[L->op_seg[i]];
    uint32_t k = L->op_slot[i];
    q->r_exps[i] = ((const uint32_t *)(img + g->r_e))[k];
    memcpy(q->r_limbs + (size_t)i * n, (const uint32_t *)(img + g->r_l) + (size_t)k * n, (size_t)n * 4);
  }
}

static void launch_groups(const fpa_request *q, const layout_t *L, uint8_t *dev, hipStream_t st) {
  for (int s = 0; s < L->nseg; s++) {
    const seg_t *g = &L->seg[s];
    fpa_arguments a;
    memset(&a, 0, sizeof a);
    a.count = g->count;
    a.precision = q->precision;
    a.mode = (fpa_rounding_mode)g->mode;
    a.r_exps = (uint32_t *)(dev + g->r_e);
    a.r_limbs = (uint32_t *)(dev + g->r_l);
    a.a_exps = (uint32_t *)(dev + g->a_e);
    a.a_limbs = (uint32_t *)(dev + g->a_l);
    a.b_exps = (uint32_t *)(dev + g->b_e);
    a.b_limbs = (uint32_t *)(dev + g->b_l);
    int rc = fpa_kernel_launch(g->op, &a, fpa_kernel_get_variant_for_count(g->op, q->precision, g->count), st);
    if (rc != 0) { fprintf(stderr, "server: launch failed (%d)\n", rc); exit(1); }
  }
}

/* --------------------------------- pool --------------------------------- */
typedef struct { size_t off, size; } blk_t;
typedef struct {
  size_t cap, used, peak;
  blk_t *fr;
  int nfr, maxfr;
} pool_t;

static size_t g256(size_t x) { return (x + 255) & ~(size_t)255; }

static void pool_init(pool_t *p, size_t cap) {
  p->cap = cap;
  p->used = p->peak = 0;
  p->maxfr = 64;
  p->fr = (blk_t *)malloc(sizeof(blk_t) * p->maxfr);
  p->nfr = 1;
  p->fr[0].off = 0;
  p->fr[0].size = cap;
}

static int pool_alloc(pool_t *p, size_t size, size_t *off) {
  size = g256(size);
  for (int i = 0; i < p->nfr; i++) {
    if (p->fr[i].size >= size) {
      *off = p->fr[i].off;
      p->fr[i].off += size;
      p->fr[i].size -= size;
      if (p->fr[i].size == 0) {
        memmove(&p->fr[i], &p->fr[i + 1], sizeof(blk_t) * (p->nfr - i - 1));
        p->nfr--;
      }
      p->used += size;
      if (p->used > p->peak) p->peak = p->used;
      return 1;
    }
  }
  return 0;
}

static void pool_free(pool_t *p, size_t off, size_t size) {
  size = g256(size);
  p->used -= size;
  int i = 0;
  while (i < p->nfr && p->fr[i].off < off) i++;
  if (p->nfr == p->maxfr) { p->maxfr *= 2; p->fr = (blk_t *)realloc(p->fr, sizeof(blk_t) * p->maxfr); }
  memmove(&p->fr[i + 1], &p->fr[i], sizeof(blk_t) * (p->nfr - i));
  p->fr[i].off = off;
  p->fr[i].size = size;
  p->nfr++;
  if (i + 1 < p->nfr && p->fr[i].off + p->fr[i].size == p->fr[i + 1].off) {
    p->fr[i].size += p->fr[i + 1].size;
    memmove(&p->fr[i + 1], &p->fr[i + 2], sizeof(blk_t) * (p->nfr - i - 2));
    p->nfr--;
  }
  if (i > 0 && p->fr[i - 1].off + p->fr[i - 1].size == p->fr[i].off) {
    p->fr[i - 1].size += p->fr[i].size;
    memmove(&p->fr[i], &p->fr[i + 1], sizeof(blk_t) * (p->nfr - i - 1));
    p->nfr--;
  }
}

/* -------------------------------- server -------------------------------- */
typedef struct {
  struct fpa_server_s *S;
  hipStream_t st;
  uint8_t *pin;
  int busy;                /* reserved by the scheduler until the worker has finished */
  fpa_request *req;
  layout_t L;
  size_t off, bytes;
  pthread_t thr;
  pthread_cond_t cv;       /* scheduler -> worker: a request was assigned */
} slot_t;

struct fpa_server_s {
  uint8_t *dev;
  size_t budget, max_req;
  pool_t pool;
  uint32_t nslots;
  slot_t *slots;
  pthread_t thr;
  pthread_mutex_t mu;
  pthread_cond_t cv_work, cv_done;
  fpa_request **q;
  uint32_t qhead, qcount, qcap;
  int stop, stop_workers;
  double t0;
  uint64_t budget_waits;
  fpa_request *waited_head;      /* head request already counted as waiting for memory */
};

/* One worker per slot: stages the request's operands, runs it on the slot's stream,
 * waits for it and returns its results, so a large request's host copies never hold
 * up the admission or completion of others. */
static void *worker(void *arg) {
  slot_t *sl = (slot_t *)arg;
  struct fpa_server_s *S = sl->S;
  pthread_mutex_lock(&S->mu);
  for (;;) {
    while (!sl->req && !S->stop_workers) pthread_cond_wait(&sl->cv, &S->mu);
    if (!sl->req) break;
    fpa_request *r = sl->req;
    pthread_mutex_unlock(&S->mu);
    scatter_inputs(r, &sl->L, sl->pin);
    uint8_t *dev = S->dev + sl->off;
    FPA_HIP(hipMemcpyAsync(dev + sl->L.r_bytes, sl->pin + sl->L.r_bytes, sl->L.in_bytes, hipMemcpyHostToDevice,
                           sl->st));
    launch_groups(r, &sl->L, dev, sl->st);
    FPA_HIP(hipMemcpyAsync(sl->pin, dev, sl->L.r_bytes, hipMemcpyDeviceToHost, sl->st));
    FPA_HIP(hipStreamSynchronize(sl->st));
    gather_results(r, &sl->L, sl->pin);
    pthread_mutex_lock(&S->mu);
    r->t_done = mono_s() - S->t0;
    pool_free(&S->pool, sl->off, sl->bytes);
    sl->req = NULL;
    sl->busy = 0;
    r->status = fpa_success;
    pthread_cond_broadcast(&S->cv_done);
    pthread_cond_signal(&S->cv_work);
  }
  pthread_mutex_unlock(&S->mu);
  return NULL;
}

/* Admission in arrival order: the queue head starts as soon as a slot is free and the
 * pool can hold it; a head that does not fit waits (later requests do not pass it). */
static void *scheduler(void *arg) {
  struct fpa_server_s *S = (struct fpa_server_s *)arg;
  pthread_mutex_lock(&S->mu);
  for (;;) {
    while (S->qcount > 0) {
      uint32_t fs;
      for (fs = 0; fs < S->nslots && S->slots[fs].busy; fs++) {}
      if (fs == S->nslots) break;
      fpa_request *r = S->q[S->qhead];
      slot_t *sl = &S->slots[fs];
      int rc = build_layout(r, &sl->L);
      size_t bytes = rc ? 0 : sl->L.r_bytes + sl->L.in_bytes;
      if (rc == 0 && (bytes > S->budget || bytes > S->max_req)) rc = fpa_invalid_counts;
      if (rc != 0) {
        S->qhead = (S->qhead + 1) % S->qcap;
        S->qcount--;
        r->status = rc;
        pthread_cond_broadcast(&S->cv_done);
        continue;
      }
      size_t off;
      if (!pool_alloc(&S->pool, bytes, &off)) {
        if (S->waited_head != r) { S->budget_waits++; S->waited_head = r; }
        break;
      }
      S->qhead = (S->qhead + 1) % S->qcap;
      S->qcount--;
      sl->busy = 1;
      sl->req = r;
      sl->off = off;
      sl->bytes = bytes;
      r->slot = fs;
      r->t_admit = mono_s() - S->t0;
      pthread_cond_signal(&sl->cv);
    }
    if (S->stop && S->qcount == 0) break;
    pthread_cond_wait(&S->cv_work, &S->mu);
  }
  pthread_mutex_unlock(&S->mu);
  return NULL;
}

fpa_result fpa_server_create(fpa_server_t *out, size_t budget_bytes, uint32_t slots, size_t max_request_bytes) {
  if (slots == 0 || budget_bytes == 0) return fpa_invalid_counts;
  struct fpa_server_s *S = (struct fpa_server_s *)calloc(1, sizeof *S);
  S->budget = g256(budget_bytes);
  S->max_req = max_request_bytes ? max_request_bytes : S->budget;
  if (S->max_req > S->budget) S->max_req = S->budget;
  FPA_HIP(hipMalloc((void **)&S->dev, S->budget));
  pool_init(&S->pool, S->budget);
  S->nslots = slots;
  S->slots = (slot_t *)calloc(slots, sizeof(slot_t));
  S->qcap = 1024;
  S->q = (fpa_request **)malloc(sizeof(fpa_request *) * S->qcap);
  pthread_mutex_init(&S->mu, NULL);
  pthread_cond_init(&S->cv_work, NULL);
  pthread_cond_init(&S->cv_done, NULL);
  for (uint32_t i = 0; i < slots; i++) {
    slot_t *sl = &S->slots[i];
    sl->S = S;
    FPA_HIP(hipStreamCreateWithFlags(&sl->st, hipStreamNonBlocking));
    FPA_HIP(hipHostMalloc((void **)&sl->pin, S->max_req, hipHostMallocDefault));
    pthread_cond_init(&sl->cv, NULL);
    pthread_create(&sl->thr, NULL, worker, sl);
  }
  S->t0 = mono_s();
  pthread_create(&S->thr, NULL, scheduler, S);
  *out = S;
  return fpa_success;
}

fpa_result fpa_server_destroy(fpa_server_t S) {
  if (!S) return fpa_success;
  pthread_mutex_lock(&S->mu);
  S->stop = 1;
  pthread_cond_broadcast(&S->cv_work);
  pthread_mutex_unlock(&S->mu);
  pthread_join(S->thr, NULL);
  /* the queue is empty; let the workers finish what they hold, then stop them */
  pthread_mutex_lock(&S->mu);
  S->stop_workers = 1;
  for (uint32_t i = 0; i < S->nslots; i++) pthread_cond_signal(&S->slots[i].cv);
  pthread_mutex_unlock(&S->mu);
  for (uint32_t i = 0; i < S->nslots; i++) pthread_join(S->slots[i].thr, NULL);
  for (uint32_t i = 0; i < S->nslots; i++) {
    FPA_HIP(hipStreamDestroy(S->slots[i].st));
    FPA_HIP(hipHostFree(S->slots[i].pin));
    pthread_cond_destroy(&S->slots[i].cv);
    layout_free(&S->slots[i].L);
  }
  FPA_HIP(hipFree(S->dev));
  free(S->pool.fr);
  free(S->slots);
  free(S->q);
  pthread_mutex_destroy(&S->mu);
  pthread_cond_destroy(&S->cv_work);
  pthread_cond_destroy(&S->cv_done);
  free(S);
  return fpa_success;
}

fpa_result fpa_server_submit(fpa_server_t S, fpa_request *r) {
  pthread_mutex_lock(&S->mu);
  r->status = -1;
  if (S->qcount == S->qcap) {
    fpa_request **nq = (fpa_request **)malloc(sizeof(fpa_request *) * S->qcap * 2);
    for (uint32_t i = 0; i < S->qcount; i++) nq[i] = S->q[(S->qhead + i) % S->qcap];
    free(S->q);
    S->q = nq;
    S->qhead = 0;
    S->qcap *= 2;
  }
  S->q[(S->qhead + S->qcount) % S->qcap] = r;
  S->qcount++;
  pthread_cond_signal(&S->cv_work);
  pthread_mutex_unlock(&S->mu);
  return fpa_success;
}

fpa_result fpa_server_wait(fpa_server_t S, fpa_request *r) {
  pthread_mutex_lock(&S->mu);
  while (r->status == -1) pthread_cond_wait(&S->cv_done, &S->mu);
  int st = r->status;
  pthread_mutex_unlock(&S->mu);
  return (fpa_result)st;
}

typedef struct {
  struct fpa_server_s *S;
  fpa_request *reqs;
  uint32_t n;
  double t0;
} producer_arg;

static void sleep_until(double t_abs) {
  double d = t_abs - mono_s();
  if (d <= 0) return;
  struct timespec ts;
  ts.tv_sec = (time_t)d;
  ts.tv_nsec = (long)((d - (double)ts.tv_sec) * 1e9);
  nanosleep(&ts, NULL);
}

static void *producer(void *arg) {
  producer_arg *p = (producer_arg *)arg;
  for (uint32_t i = 0; i < p->n; i++) {
    sleep_until(p->t0 + p->reqs[i].arrival);
    fpa_server_submit(p->S, &p->reqs[i]);
  }
  return NULL;
}

fpa_result fpa_server_run_trace(fpa_server_t S, fpa_request *reqs, uint32_t nreq) {
  for (uint32_t i = 0; i < nreq; i++) reqs[i].status = -1;
  pthread_mutex_lock(&S->mu);
  S->t0 = mono_s();
  pthread_mutex_unlock(&S->mu);
  producer_arg pa = { S, reqs, nreq, S->t0 };
  pthread_t pt;
  pthread_create(&pt, NULL, producer, &pa);
  pthread_join(pt, NULL);
  fpa_result worst = fpa_success;
  for (uint32_t i = 0; i < nreq; i++) {
    fpa_result r = fpa_server_wait(S, &reqs[i]);
    if (r != fpa_success) worst = r;
  }
  return worst;
}

size_t fpa_server_peak_bytes(fpa_server_t S) { return S->pool.peak; }
uint64_t fpa_server_budget_waits(fpa_server_t S) { return S->budget_waits; }

/* ------------------------------ plain path ------------------------------ */
fpa_result fpa_serve_plain_trace(fpa_request *reqs, uint32_t nreq) {
  layout_t L;
  memset(&L, 0, sizeof L);
  double t0 = mono_s();
  fpa_result worst = fpa_success;
  for (uint32_t i = 0; i < nreq; i++) {
    fpa_request *r = &reqs[i];
    sleep_until(t0 + r->arrival);
    r->t_admit = mono_s() - t0;
    int rc = build_layout(r, &L);
    if (rc) { r->status = rc; worst = (fpa_result)rc; r->t_done = mono_s() - t0; continue; }
    size_t bytes = L.r_bytes + L.in_bytes;
    uint8_t *img = (uint8_t *)malloc(bytes);
    uint8_t *dev;
    FPA_HIP(hipMalloc((void **)&dev, bytes));
    scatter_inputs(r, &L, img);
    FPA_HIP(hipMemcpy(dev + L.r_bytes, img + L.r_bytes, L.in_bytes, hipMemcpyHostToDevice));
    launch_groups(r, &L, dev, NULL);
    FPA_HIP(hipMemcpy(img, dev, L.r_bytes, hipMemcpyDeviceToHost));
    gather_results(r, &L, img);
    FPA_HIP(hipFree(dev));
    free(img);
    r->status = fpa_success;
    r->t_done = mono_s() - t0;
  }
  layout_free(&L);
  return worst;
}
