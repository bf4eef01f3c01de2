/*
 * psba_server.h -- a request loop over batches (C ABI, part of libpsba.so).
 *
 * No counterpart in the reference, which solves one problem per process
 * (PSBA/main.cpp:193-208).  Requests (independent problems, the same data
 * contract as psba_batch_add) are queued on submit.  Every psba_server_step
 * call, without blocking:
 *   - polls the batches on the device and retires each request as soon as its
 *     own solve has stopped (its parameters are copied back while the rest of
 *     its batch keeps running);
 *   - admits queued requests, oldest first, into a new batch on every idle
 *     lane (one HIP stream each), grouping requests of one size class, as long
 *     as the device memory all lanes hold stays within the HBM budget;
 *   - launches that batch, so its rounds overlap on the device with the
 *     batches already running.
 * Each request is solved by the LM rules of psba_batch (psba_levmar with
 * PSBA_FLAG_NO_HANDOFF plus the relative-decrease stop); a result does not
 * depend on which other requests shared its batch.
 */
#ifndef PSBA_SERVER_H
#define PSBA_SERVER_H

#include "psba_batch.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct psba_server psba_server;

typedef struct {
  int cnp;              /* camera parameters, the same for every request */
  int precision;        /* PSBA_PREC_FP64 or PSBA_PREC_MIXED */
  int device;           /* HIP device ordinal */
  int lanes;            /* batches in flight at once, one stream each (0 = 4) */
  int max_batch;        /* requests per batch (0 = 512) */
  long long hbm_budget; /* device bytes all lanes may hold (0 = 200e9) */
  int max_iter;         /* LM iterations per request (0 = 100) */
  double rel_tol;       /* relative-decrease stop (< 0 = 1e-8, 0 = off) */
} psba_server_desc;

typedef struct {
  long long submitted, completed, rejected, batches;
  long long bytes_held;  /* device memory the lanes hold now */
  long long bytes_peak;  /* most ever held */
  int queued, running;   /* requests waiting / admitted and not yet complete */
} psba_server_stats;

int psba_server_create(const psba_server_desc* d, psba_server** out);
int psba_server_destroy(psba_server* s);

/* Seconds on the clock the server stamps requests with (steady, arbitrary origin). */
double psba_server_clock(void);

/* Queue one request (host arrays are copied).  *id identifies it. */
int psba_server_submit(psba_server* s, int n_cams, int n_pts, int n_obs, const double* K, const double* impts,
                       const double* initrot, const double* cams, const double* pts, const int* iidx,
                       const int* jidx, const unsigned char* fixmask, long long* id);

/* One non-blocking pass (retire, admit, launch); returns the number of
 * requests completed by this call (>= 0) or an error. */
int psba_server_step(psba_server* s);

/* Step until every submitted request has completed. */
int psba_server_drain(psba_server* s);

/* Take a completed request's result: returns 1 and releases it, 0 while it is
 * still queued or running, PSBA_ERR_ARG for an unknown id.  *code is the
 * ITER_* stop code, or PSBA_ERR_NOMEM for a request larger than the budget.
 * cams / pts (may be NULL) receive the final parameters; t_submit / t_done
 * (may be NULL) the psba_server_clock times of submission and completion. */
int psba_server_result(psba_server* s, long long id, int* code, int* itno, double* init_err, double* final_err,
                       double* cams, double* pts, double* t_submit, double* t_done);

/* Ids of requests completed since the last call, in completion order (up to
 * cap; the rest stay listed).  Returns how many were written. */
int psba_server_completed(psba_server* s, long long* ids, int cap);

int psba_server_get_stats(psba_server* s, psba_server_stats* out);

#ifdef __cplusplus
}
#endif

#endif
