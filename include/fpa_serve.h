/*
 * fpa_serve.h -- continuous-batching request server for element-wise fpa ops.
 *
 * No counterpart in the reference (its API is one synchronous-looking call per op);
 * this is for callers that receive many independent requests.  A request is a batch of
 * ops (mixed add/sub/mul/div/sqrt) on values of one precision, with host operands and
 * host result buffers packed in op order (value i at limbs [i*p/32, (i+1)*p/32)).
 *
 * The server owns one device pool of `budget_bytes` (allocated once, kept resident) and
 * `slots` HIP streams.  A scheduler thread admits queued requests in arrival order as
 * soon as a slot is free and the pool can hold the request's operands and results;
 * each slot's worker thread then stages, uploads, runs (one kernel per (op, mode)
 * group, via the tuned dispatcher) and downloads its request on the slot's own stream,
 * so requests overlap on the device and on the host, and new ones start while others
 * are still running.  Results are identical to running each
 * request alone (every element is computed independently).
 */
#ifndef FPA_SERVE_H
#define FPA_SERVE_H

#include <stddef.h>
#include <stdint.h>
#include "fpa_kernels.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  uint32_t        nops;
  uint32_t        precision;
  const uint8_t  *ops;                  /* fpa_op per op (ADD..SQRT) */
  const uint8_t  *modes;                /* fpa_rounding_mode per op */
  const uint32_t *a_exps, *a_limbs;     /* host operands, packed in op order */
  const uint32_t *b_exps, *b_limbs;     /* may be NULL if no op is binary */
  uint32_t       *r_exps, *r_limbs;     /* host results */
  double          arrival;              /* trace time of arrival (s) */
  /* written by the server */
  double          t_admit;              /* trace time the request started on the device */
  double          t_done;               /* trace time its results were in r_* */
  int32_t         status;               /* fpa_result; -1 while pending */
  uint32_t        slot;
} fpa_request;

typedef struct fpa_server_s *fpa_server_t;

/* Pool of budget_bytes device memory, `slots` streams, pinned staging per slot sized for
 * requests of up to max_request_bytes of operands+results. */
fpa_result fpa_server_create(fpa_server_t *server, size_t budget_bytes, uint32_t slots, size_t max_request_bytes);
fpa_result fpa_server_destroy(fpa_server_t server);

/* Asynchronous use: submit (thread-safe) and wait for one request. */
fpa_result fpa_server_submit(fpa_server_t server, fpa_request *req);
fpa_result fpa_server_wait(fpa_server_t server, fpa_request *req);

/* Replay a trace: a producer thread submits reqs[i] at reqs[i].arrival seconds after the
 * start (arrivals non-decreasing); returns when every request is done.  Trace times in
 * t_admit/t_done are on the same clock. */
fpa_result fpa_server_run_trace(fpa_server_t server, fpa_request *reqs, uint32_t nreq);

/* The same trace served the plain way: one request at a time on the null stream, each
 * allocating its device arrays, copying, launching, copying back and freeing. */
fpa_result fpa_serve_plain_trace(fpa_request *reqs, uint32_t nreq);

/* Device bytes a request needs (operands + results), 0 if it is invalid. */
size_t fpa_request_bytes(const fpa_request *req);

/* Peak pool bytes in use and admissions delayed by the budget since creation. */
size_t fpa_server_peak_bytes(fpa_server_t server);
uint64_t fpa_server_budget_waits(fpa_server_t server);

#ifdef __cplusplus
}
#endif

#endif
