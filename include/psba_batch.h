/*
 * psba_batch.h -- many independent bundle-adjustment problems solved together
 * (C ABI, part of libpsba.so).
 *
 * No counterpart in the reference, which solves one problem per process
 * (PSBA/main.cpp:193-208).  Each problem follows the reference's data
 * contract (psba_capi.h: point-major, camera-ascending observations; pose-last
 * camera vectors; K per camera) and is solved by the same Levenberg-Marquardt
 * rules as psba_levmar with PSBA_FLAG_NO_HANDOFF (PSBA/levmar.cpp:45-280),
 * stopping on a relative cost decrease below rel_tol, on max_iter
 * iterations, or on the reference's own tests.
 *
 * All problems of a batch advance in lock-step LM rounds: one launch per stage
 * covers every problem (observations, points and cameras are concatenated),
 * the accept / reject decision, damping and iteration counters live on the
 * device, and a round is a captured HIP graph replayed until every problem
 * has stopped.  Problems that stopped are skipped by every stage.
 *
 * A batch runs on its own HIP stream; several batches can be in flight at
 * once (the serving loop admits new requests into a new batch while earlier
 * ones are still running).
 */
#ifndef PSBA_BATCH_H
#define PSBA_BATCH_H

#include "psba_capi.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct psba_batch psba_batch;

typedef struct {
  int cnp;          /* camera parameters (6 | 9 | 11 | 16), the same for all problems */
  int precision;    /* PSBA_PREC_FP64 or PSBA_PREC_MIXED */
  int device;       /* HIP device ordinal */
  int max_problems; /* problems per batch (0 = no limit) */
} psba_batch_desc;

/* Device memory one problem of the given size takes inside a batch (bytes, an
 * upper bound); pair_items = sum over its points of t(t+1)/2, t the track
 * length.  For admission control by HBM budget. */
long long psba_problem_bytes(int cnp, int precision, int n_cams, int n_pts, int n_obs, long long pair_items);

int psba_batch_create(const psba_batch_desc* d, psba_batch** out);
int psba_batch_destroy(psba_batch* b);

/* Device memory the batch holds now (bytes). */
long long psba_batch_bytes(const psba_batch* b);

/* Drop every problem (device memory is kept for the next commit). */
int psba_batch_reset(psba_batch* b);

/* Add one problem (host arrays, copied).  iidx / jidx are the problem's own
 * point / camera indices, point-major and camera-ascending; fixmask
 * (cnp*n_cams, may be NULL) holds parameters fixed.  *slot receives the
 * problem's index in the batch. */
int psba_batch_add(psba_batch* b, int n_cams, int n_pts, int n_obs, const double* K, const double* impts,
                   const double* initrot, const double* cams, const double* pts, const int* iidx, const int* jidx,
                   const unsigned char* fixmask, int* slot);

/* Build the concatenated index of every problem added since the last reset
 * and upload it (all problems start from their initial parameters). */
int psba_batch_commit(psba_batch* b);

/* Start an asynchronous solve of the committed problems on the batch's
 * stream and return.  o may be NULL: max_iter 100, rel_tol 1e-8; otherwise
 * max_iter (<= 0: 100) and rel_tol (0: no relative test) are used. */
int psba_batch_launch(psba_batch* b, const psba_opts* o);

/* 1 once every problem of the launched solve has stopped, 0 while running
 * (non-blocking; keeps the next rounds queued). */
int psba_batch_poll(psba_batch* b);

/* Launch and wait. *rounds (may be NULL) receives the LM rounds executed. */
int psba_batch_solve(psba_batch* b, const psba_opts* o, int* rounds);

/* Per-problem outcome: the ITER_* code, iterations, initial and final sum of
 * squared residuals, and (if non-NULL) the final parameters. */
int psba_batch_result(psba_batch* b, int slot, int* code, int* itno, double* init_err, double* final_err,
                      double* cams, double* pts);

/* Device time of the last solve (milliseconds, from launch to the round that
 * found every problem stopped) and its LM rounds. */
int psba_batch_elapsed(psba_batch* b, double* ms, int* rounds);

/* Auto-tuning: time the candidate launch shapes of the per-observation and
 * per-point stage kernels on the committed problems and keep the fastest.
 * Shapes that would change a reduction order are not tuned, so results do
 * not depend on the choice.  *obs_block / *pt_block (may be NULL) receive it. */
int psba_batch_autotune(psba_batch* b, int* obs_block, int* pt_block);

#ifdef __cplusplus
}
#endif

#endif
