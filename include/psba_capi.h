/*
 * psba_capi.h -- C-ABI of the MI355X bundle adjuster (libpsba.so).
 *
 * Plain pointers and sizes, int status codes (0 = OK, negative = error), no
 * exit().  One context owns the device state of one problem on one HIP device
 * and stream; contexts are independent.  The caller owns all host memory; the
 * context copies what it needs.  The C++ drop-in layer (cl_psba.h, levmar.h,
 * trust_region.h, misc.h) is implemented on top of these entry points.
 *
 * Layout conventions (identical to PSBA):
 *   observations point-major, camera-ascending within a point
 *     (PSBA/misc.cpp:189-216);
 *   cams: cnp doubles per camera, pose last [... v1 v2 v3 t1 t2 t3]
 *     cnp=6  [v t]                         (PSBA/CL_files/PSBA.cl:5 cnp=6)
 *     cnp=9  [f k1 k2 v t]
 *     cnp=11 [fx cx cy ar skew v t]
 *     cnp=16 [fx cx cy ar skew kc1..kc5 v t]
 *   K: 5 doubles per camera [fx cx cy ar skew] (PSBA/psba.h:3 K_DIM);
 *   initrot: unit quaternion (s,x,y,z) per camera (PSBA/cl_psba.h:20);
 *   impts: 2 doubles per observation (PSBA/cl_psba.h:19).
 */
#ifndef PSBA_CAPI_H
#define PSBA_CAPI_H

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define PSBA_OK 0
#define PSBA_ERR_ARG (-1)
#define PSBA_ERR_HIP (-2)
#define PSBA_ERR_NOMEM (-3)
#define PSBA_ERR_STATE (-4)
#define PSBA_ERR_UNSUPPORTED (-5)
#define PSBA_ERR_NODEVICE (-6)

/* driver return codes, PSBA/psba.h:12-18 */
#define PSBA_ITER_TURN_TO_LM 1
#define PSBA_ITER_TURN_TO_TR 2
#define PSBA_ITER_CONTINUE 3
#define PSBA_ITER_ERR 4
#define PSBA_ITER_DP_NO_CHANGE 5
#define PSBA_ITER_ERR_SMALL_ENOUGH 6
#define PSBA_ITER_PASS 7

/* arithmetic of the Jacobian / normal-equation / Schur stages */
#define PSBA_PREC_FP64 0  /* plain path: everything in FP64 */
#define PSBA_PREC_MIXED 1 /* FP32 Jacobian/W/V/Schur, FP64 factor, cost and refinement */

typedef struct psba_ctx psba_ctx;

typedef struct {
  int cnp, pnp, mnp;        /* camera params (6|9|11|16), 3, 2 */
  int n_cams, n_pts, n_obs; /* nCams, n3Dpts, n2Dprojs */
  int precision;            /* PSBA_PREC_* */
  int device;               /* HIP device ordinal */
} psba_desc;

/* psba_opts.flags */
#define PSBA_FLAG_NO_HANDOFF 1 /* LM never returns ITER_TURN_TO_TR (pure LM solve) */
#define PSBA_FLAG_HOST_CONTROL 2 /* psba_levmar decides each trial on the host (one
                                    read-back per trial) instead of on the device */

typedef struct {
  int max_iter;      /* shared iteration cap (reference: 50) */
  int max_pb_tries;  /* guard on the dogleg's factor-retry loop (0 = 64) */
  int refine_steps;  /* mixed precision: FP64 refinement sweeps of every step (0 = none) */
  int flags;         /* PSBA_FLAG_* (0 = reference behaviour) */
  double rel_tol;    /* > 0: LM also stops (ITER_DP_NO_CHANGE) once an accepted
                        step lowers the cost by less than rel_tol relative */
} psba_opts;

/* one record per LM trial or dogleg step (same fields as the oracle's) */
typedef struct {
  int itno, kind, factor_ok, accepted;
  double cost, damp, rho, lambda, dp_sq;
} psba_trace_rec;

/* ---- lifecycle ------------------------------------------------------------ */
/* replaces setup_cl, PSBA/cl_psba.h:93 (PSBA/cl_psba.cpp:16-133) */
int psba_create(const psba_desc* desc, psba_ctx** out);
/* replaces release_buffer, PSBA/cl_psba.h:94 */
int psba_destroy(psba_ctx* ctx);
/* replaces fill_initBuffer2, PSBA/cl_psba.h:114-121 */
int psba_upload(psba_ctx* ctx, const double* K, const double* impts, const double* initrot,
                const double* cams, const double* pts);
/* replaces fill_idxBuffer, PSBA/cl_psba.h:123-126: the device index (point CSR,
 * camera CSR, camera-pair lists) is built from the observation order */
int psba_set_index(psba_ctx* ctx, const int* iidx, const int* jidx);
/* the same index built on the device (stable radix sorts and scans of the
 * observation order); every index array is bit-identical to psba_set_index's */
int psba_set_index_device(psba_ctx* ctx, const int* iidx, const int* jidx);
/* addition (no reference counterpart): hold parameters fixed; mask has
 * cnp*n_cams bytes, 1 = fixed.  NULL clears the mask. */
int psba_set_fixed(psba_ctx* ctx, const unsigned char* mask);
/* addition (no reference counterpart; the reference minimises the plain sum
 * of squares): a robust loss on each observation's squared residual norm s,
 *   PSBA_LOSS_HUBER   rho(s) = s (s <= d^2), 2 d sqrt(s) - d^2 beyond
 *   PSBA_LOSS_CAUCHY  rho(s) = d^2 log(1 + s / d^2)
 * with scale d (pixels).  Steps solve the reweighted problem (residual and
 * Jacobian rows scaled by sqrt(rho'(s))); costs, the gain ratio and finalErr
 * are sum rho(s).  PSBA_LOSS_NONE (the default) is the reference's cost. */
#define PSBA_LOSS_NONE 0
#define PSBA_LOSS_HUBER 1
#define PSBA_LOSS_CAUCHY 2
int psba_set_loss(psba_ctx* ctx, int kind, double scale);
/* accessor for the optimised parameters (the reference leaves them in
 * cams_buffer / pts3D_buffer, PSBA/cl_psba.h:21-22) */
int psba_read_params(psba_ctx* ctx, double* cams, double* pts);
int psba_write_params(psba_ctx* ctx, const double* cams, const double* pts);
/* device-side copy of the current parameters and its restore (no host traffic) */
int psba_snapshot_params(psba_ctx* ctx);
int psba_restore_params(psba_ctx* ctx);
const char* psba_strerror(int code);
const char* psba_last_error(void);
int psba_device_count(void);
/* name (with the architecture), compute units, peak clock and memory of a
 * device; any output may be NULL */
int psba_device_info(int device, char* name, int name_cap, int* cus, int* clock_mhz, long long* mem_bytes);

/* ---- host integer indexing ------------------------------------------------- */
/* replaces generate_idxs, PSBA/misc.h:22-23 (PSBA/misc.cpp:178-218); bit-exact
 * restatement.  comm3DIdx may be NULL (it needs nCams*nCams*n3Dpts ints). */
int psba_generate_idxs(int nCams, int n3Dpts, int n2Dprojs, const char* vmask, int* comm3DIdx,
                       int* comm3DIdxCnt, int* iidx, int* jidx, int* blk_idx);
/* generate_idxs (PSBA/misc.cpp:178-218) on a device: compaction of the
 * visibility mask by a scan; iidx, jidx, blk_idx and comm3DIdxCnt (may be
 * NULL) bit-exact.  The dense comm3DIdx lists are not produced (the device
 * index replaces them). */
int psba_generate_idxs_device(int device, int nCams, int n3Dpts, int n2Dprojs, const char* vmask,
                              int* comm3DIdxCnt, int* iidx, int* jidx, int* blk_idx);

/* generate_idxs with no allocation (no counterpart in the reference: for
 * callers that cannot allocate, such as code under HIP stream capture or
 * serving loops that forbid hipMalloc).  Device pointers in and out: d_vmask
 * (nCams*n3Dpts bytes), d_iidx and d_jidx (n2Dprojs ints), d_blk_idx
 * (nCams*n3Dpts ints), d_count (1 int, may be NULL: the number of visible
 * entries; iidx/jidx entries beyond n2Dprojs are not written and their
 * blk_idx is -1).  Runs on `stream` (a hipStream_t, NULL for the null
 * stream) without synchronising; the results equal psba_generate_idxs_device's
 * (comm3DIdxCnt is not produced).  Its only intermediate lives in a fixed
 * device array of 65,537 ints per device, so nCams*n3Dpts is limited to
 * 67,108,864 (PSBA_ERR_ARG beyond); calls that overlap on one device race. */
int psba_generate_idxs_device_noalloc(int device, void* stream, int nCams, int n3Dpts, int n2Dprojs,
                                      const char* d_vmask, int* d_iidx, int* d_jidx, int* d_blk_idx,
                                      int* d_count);

/* ---- drivers ---------------------------------------------------------------- */
/* replaces levmar, PSBA/levmar.h:6-11.  *itno is the shared iteration counter
 * (the reference's global itno, PSBA/main.cpp:27); *finalErr receives the
 * final sum of squared residuals.  Returns a PSBA_ITER_* code (> 0) or an error
 * (< 0).  trace may be NULL. */
int psba_levmar(psba_ctx* ctx, const psba_opts* opts, int* itno, double* finalErr,
                psba_trace_rec* trace, int trace_cap, int* trace_len);
/* the sum of squared residuals at the entry of the last psba_levmar call on
 * ctx: the value PSBA/levmar.cpp:95 stores in the global initErr */
int psba_levmar_entry_cost(psba_ctx* ctx, double* sse);
/* replaces trust_region, PSBA/trust_region.h:5-11 (its blk_idx argument is
 * unused by the reference and absent here) */
int psba_trust_region(psba_ctx* ctx, const psba_opts* opts, int* itno, double* finalErr,
                      psba_trace_rec* trace, int trace_cap, int* trace_len);
/* the LM <-> dogleg hand-off loop of PSBA/main.cpp:193-208 */
int psba_hybrid(psba_ctx* ctx, const psba_opts* opts, int* itno, double* initErr,
                double* finalErr, psba_trace_rec* trace, int trace_cap, int* trace_len);

/* The hybrid loop of PSBA/main.cpp:193-208 with the report that
 * PSBA/main.cpp:209-225 prints: stop code, wall time, the reference's error
 * measure sqrt(SSE) / n_obs (initial and final), total iterations and the time
 * per stage (here device time, from HIP events on the context's stream).
 * Added: the RMS reprojection error per observation sqrt(SSE / n_obs), the
 * split of iterations between LM and the dogleg, and the hand-offs. */
typedef struct {
  double jacobian_normal, schur, factor, solve, backsub, residual, other;
  int trials, jacobians;
} psba_report_stages;
typedef struct {
  int code;            /* PSBA_ITER_* */
  int itno;            /* the shared iteration counter after the loop */
  int lm_iterations, tr_iterations, handoffs; /* iterations entered by each driver */
  int lm_trials, tr_steps;
  double init_sse, final_sse;
  double init_error_ref, final_error_ref; /* sqrt(SSE) / n_obs (PSBA/main.cpp:214-215) */
  double init_rms_px, final_rms_px;       /* sqrt(SSE / n_obs) */
  double wall_ms;
  psba_report_stages stages;               /* milliseconds per stage, counts */
} psba_report;
int psba_hybrid_report(psba_ctx* ctx, const psba_opts* opts, int* itno, psba_report* out);

/* ---- stage entry points (PSBA/sba_func.h:10-138) ---------------------------- */
/* residual and its squared norm at the current parameters (use_trial = 0) or
 * at p + dp (use_trial = 1); PSBA/sba_func.cpp:81-145 */
int psba_stage_residual(psba_ctx* ctx, int use_trial, double* cost);
/* Jacobian + U/V/W/g with the reference's coefficients (LM: 1,1; dogleg: 2,-2);
 * PSBA/sba_func.cpp:153-617 */
int psba_stage_jacobian_normal(psba_ctx* ctx, double coeff_uvw, double coeff_g);
/* damped V^-1, S = U + mu I - sum W V^-1 W^T and ea; PSBA/sba_func.cpp:624-995 */
int psba_stage_schur(psba_ctx* ctx, double mu);
/* Cholesky of S and the two triangular solves for dpa; *ok = 1 on success;
 * replaces SPDinv + matVec_mul, PSBA/cl_spdinv.h + PSBA/cl_linearalg.h */
int psba_stage_factor_solve(psba_ctx* ctx, int* ok);
/* eb = gb - W^T dpa and dpb = V^-1 eb; PSBA/sba_func.cpp:1001-1117 */
int psba_stage_backsub(psba_ctx* ctx);
/* modified Cholesky of the current S, returns lambda = |sum E| / n;
 * PSBA/cl_cholmod.h (cholmod_blk, compute_cholmod_E) */
int psba_stage_modchol(psba_ctx* ctx, double* lambda);
/* mixed-precision contexts: FP64 iterative refinement of the current step dp
 * at damping mu (after the backsub stage): per sweep, r = J^T (e - J dp) -
 * mu dp with the Jacobian and residual recomputed in FP64, reduced and solved
 * with the existing blocks and factor, dp += correction; g is replaced by the
 * FP64 J^T e.  No reference counterpart (the reference is FP64 throughout). */
int psba_stage_refine(psba_ctx* ctx, double mu, int sweeps);
/* test hook: the damped 3x3 inverses of compute_Vinv (PSBA/CL_files/
 * compute_Vinv.cl:6-90, including its small-determinant branch :31-73) on
 * n_pts caller-given blocks V9 (9 per point, row-major, symmetric part used);
 * replaces the context's V and writes the inverses to out9 (9 per point) */
int psba_stage_vinv(psba_ctx* ctx, const double* V9, double mu, double* out9);
/* test hook: load an n x n matrix (n = cnp*n_cams, row-major, lower triangle
 * used) as S and, if b != NULL, b as e_a, so the dense stages can be checked
 * on arbitrary matrices */
int psba_stage_load_S(psba_ctx* ctx, const double* S, const double* b);

/* device buffer read-back for stage parity tests */
#define PSBA_BUF_EX 0      /* 2*n_obs residual */
#define PSBA_BUF_JA 1      /* 2*cnp*n_obs */
#define PSBA_BUF_JB 2      /* 6*n_obs */
#define PSBA_BUF_U 3       /* cnp*cnp*n_cams */
#define PSBA_BUF_V 4       /* 9*n_pts */
#define PSBA_BUF_W 5       /* 3*cnp*n_obs */
#define PSBA_BUF_G 6       /* n_tot */
#define PSBA_BUF_VINV 7    /* 9*n_pts */
#define PSBA_BUF_S 9       /* n*n, lower triangle valid (n = cnp*n_cams) */
#define PSBA_BUF_EA 10     /* n */
#define PSBA_BUF_DP 11     /* n_tot */
#define PSBA_BUF_EB 12     /* 3*n_pts */
#define PSBA_BUF_L 13      /* n*n lower factor */
long psba_buffer_size(const psba_ctx* ctx, int which);
int psba_read_buffer(psba_ctx* ctx, int which, double* out, long n);

/* device index read-back (ints) */
#define PSBA_IDX_OBS_PT 0         /* n_obs */
#define PSBA_IDX_OBS_CAM 1        /* n_obs */
#define PSBA_IDX_PT_PTR 2         /* n_pts + 1 */
#define PSBA_IDX_CAM_PTR 3        /* n_cams + 1 */
#define PSBA_IDX_CAM_OBS 4        /* n_obs: observations of each camera, ascending */
#define PSBA_IDX_CCHUNK_BEGIN 5   /* camera chunks */
#define PSBA_IDX_CCHUNK_END 6
#define PSBA_IDX_CAM_CCHUNK_PTR 7 /* n_cams + 1 */
#define PSBA_IDX_PAIR_ITEMS 8     /* 2 per item: (obs in camera k, obs in camera l) */
#define PSBA_IDX_PAIR_K 9
#define PSBA_IDX_PAIR_L 10
#define PSBA_IDX_SCHUNK_BEGIN 11  /* pair chunks */
#define PSBA_IDX_SCHUNK_END 12
#define PSBA_IDX_PAIR_SCHUNK_PTR 13
#define PSBA_IDX_LBLOCK_PAIR 14   /* n_cams(n_cams+1)/2 */
#define PSBA_IDX_SCHUNK_ORDER 15  /* launch order of the pair chunks (n_schunk) */
#define PSBA_IDX_PT_ORDER 16      /* points by descending track length, ties ascending id (n_pts) */
long psba_index_size(const psba_ctx* ctx, int which);
int psba_read_index(psba_ctx* ctx, int which, int* out, long n);

/* problem/index statistics */
typedef struct {
  int cnp, n_cams, n_pts, n_obs;
  int n;              /* cnp * n_cams, order of S */
  int n_pad;          /* S storage order (multiple of the tile size) */
  int n_pairs;        /* camera pairs k >= l sharing a point */
  long n_pair_items;  /* sum over points of t(t+1)/2 */
  double sum_t2;      /* sum over points of t^2 */
} psba_info;
int psba_get_info(psba_ctx* ctx, psba_info* out);

/* Kernel probe: HIP events recorded on the context's stream around every
 * launch of one kernel (or kernel group); psba_probe_read returns the summed
 * device time and the number of launches since the last read. */
#define PSBA_PROBE_OFF 0
#define PSBA_PROBE_SCHUR 1    /* Schur-complement assembly (k_schur_chunks) */
#define PSBA_PROBE_FACTOR 2   /* whole blocked Cholesky + forward substitution */
#define PSBA_PROBE_JACOBIAN 3 /* Jacobian / W per observation (k_jacobian) */
#define PSBA_PROBE_BSOLVE 4   /* back substitution */
int psba_probe(psba_ctx* ctx, int which);
int psba_probe_read(psba_ctx* ctx, double* total_ms, int* count);

/* per-stage device timings of the last driver call (milliseconds) */
typedef struct {
  double jacobian_normal, schur, factor, solve, backsub, residual, other;
  int trials, jacobians;
} psba_timing;
int psba_get_timing(psba_ctx* ctx, psba_timing* out);
int psba_enable_timing(psba_ctx* ctx, int on);

#ifdef __cplusplus
}
#endif
#endif
