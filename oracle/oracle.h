/*
 * oracle.h -- CPU FP64 restatement of PSBA's Schur-complement LM / dogleg path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * path in psba_amd/.  Only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py may load it; the product never links, calls or falls back to
 * it.
 *
 * Every routine restates the behaviour of a reference file (cited per function
 * in oracle.c).  Conventions shared with the product:
 *   - observations are point-major and camera-ascending within a point
 *     (PSBA/misc.cpp:189-216);
 *   - camera parameter vector of cnp doubles per camera, pose last:
 *       cnp=6 : [v1 v2 v3 t1 t2 t3]                 intrinsics fixed in K
 *       cnp=9 : [f k1 k2 v1 v2 v3 t1 t2 t3]         cx,cy,ar,skew fixed in K
 *       cnp=11: [fx cx cy ar s v1 v2 v3 t1 t2 t3]
 *       cnp=16: [fx cx cy ar s kc1..kc5 v1 v2 v3 t1 t2 t3]
 *     (intrinsic order fu,u0,v0,ar,s,kc(1:5) as data/54camsvarKD.txt:1);
 *   - K is 5 doubles per camera [fx cx cy ar skew] (PSBA/psba.h:3 K_DIM=5);
 *   - q0 is the fixed initial unit quaternion per camera (s,x,y,z);
 *   - residual e = measured - projected; J is the Jacobian of the projection.
 */
#ifndef PSBA_ORACLE_H
#define PSBA_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* return codes of the drivers, PSBA/psba.h:12-18 */
#define ORC_ITER_TURN_TO_LM 1
#define ORC_ITER_TURN_TO_TR 2
#define ORC_ITER_CONTINUE 3
#define ORC_ITER_ERR 4
#define ORC_ITER_DP_NO_CHANGE 5
#define ORC_ITER_ERR_SMALL_ENOUGH 6
#define ORC_ITER_PASS 7

/* how the reduced camera system S dpa = ea is solved */
#define ORC_SOLVE_CHOLESKY 0 /* L L^T factor + forward/back substitution */
#define ORC_SOLVE_INVERSE 1  /* reference-faithful explicit S^-1 then mat-vec */

typedef struct {
  int cnp;
  int n_cams, n_pts, n_obs;
  const double *K;              /* 5*n_cams */
  const double *q0;             /* 4*n_cams */
  const double *meas;           /* 2*n_obs  */
  const int *iidx, *jidx;       /* n_obs    */
  const unsigned char *fixmask; /* cnp*n_cams, 1 = parameter held fixed; may be NULL */
  const double *cams;           /* cnp*n_cams initial parameters */
  const double *pts;            /* 3*n_pts initial parameters */
} orc_problem;

#define ORC_FLAG_NO_HANDOFF 1 /* LM never returns ORC_ITER_TURN_TO_TR */

typedef struct {
  int solve_mode;   /* ORC_SOLVE_* */
  int max_iter;     /* shared iteration cap, 50 in the reference */
  int max_pb_tries; /* guard on the dogleg's factor-retry loop (reference: unbounded) */
  int flags;        /* ORC_FLAG_* */
  double rel_tol;   /* > 0: stop after an accepted step with relative decrease < rel_tol */
} orc_opts;

/* one record per LM trial or dogleg step */
typedef struct {
  int itno;      /* value of the shared iteration counter */
  int kind;      /* 0 = LM trial, 1 = dogleg step */
  int factor_ok; /* 1 if the reduced system was factored */
  int accepted;  /* 1 if the step was committed */
  double cost;   /* trial cost (sum of squared residuals), NaN if not evaluated */
  double damp;   /* LM: mu of the trial; dogleg: radius Delta used for the step */
  double rho;    /* gain ratio, NaN if not evaluated */
  double lambda; /* dogleg: Levenberg lambda in the reduced system; LM: 0 */
  double dp_sq;  /* squared norm of the step */
} orc_trace_rec;

typedef struct orc_ctx orc_ctx;

/* robust loss on each observation's squared residual norm (extension; the
 * reference has none): ORC_LOSS_NONE reproduces the reference */
#define ORC_LOSS_NONE 0
#define ORC_LOSS_HUBER 1
#define ORC_LOSS_CAUCHY 2
int orc_set_loss(orc_ctx *c, int kind, double scale);
void orc_loss_eval(int kind, double scale, double s, double *rho, double *drho);

/* ---- model --------------------------------------------------------------- */
int orc_cnp_supported(int cnp);
/* projection of point M by camera (K5, q0, cam); optional Jacobians
 * JA (2 x cnp row-major) and JB (2 x 3 row-major) of the projection. */
int orc_project(int cnp, const double *K5, const double *q0, const double *cam,
                const double *M, double *uv, double *JA, double *JB);

/* ---- dense kernels on caller matrices (n x n row-major, lower used) ------- */
/* factor + solve S x = b; returns 1 on success.  L (n*n) may be NULL. */
int orc_dense_factor_solve(int n, int solve_mode, const double *S, const double *b, double *x, double *L);
/* modified Cholesky: returns lambda = |sum E| / n; L and E may be NULL */
double orc_dense_modchol(int n, const double *S, double *L, double *E);

/* ---- integer indexing, PSBA/misc.cpp:178-218 ------------------------------ */
/* comm3DIdx may be NULL (it is n_cams*n_cams*n_pts ints). */
void orc_generate_idxs(int nCams, int n3Dpts, const char *vmask, int *comm3DIdx,
                       int *comm3DIdxCnt, int *iidx, int *jidx, int *blk_idx);

/* ---- context ------------------------------------------------------------- */
orc_ctx *orc_create(const orc_problem *p);
void orc_destroy(orc_ctx *c);
void orc_set_threads(int n);
/* order of every reduction: 0 ascending index (default, the reference's),
 * 1 descending (a second valid FP64 evaluation order, for experiments) */
void orc_set_sum_order(int descending);
int orc_get_threads(void);
/* current parameters */
void orc_get_params(const orc_ctx *c, double *cams, double *pts);
void orc_set_params(orc_ctx *c, const double *cams, const double *pts);

/* ---- stages (each mirrors a reference kernel; see oracle.c) --------------- */
double orc_stage_residual(orc_ctx *c, int use_new, double *ex);
void orc_stage_jacobian(orc_ctx *c);
void orc_stage_normal(orc_ctx *c, double coeff_uvw, double coeff_g);
double orc_max_uvdiag(const orc_ctx *c);
int orc_stage_vinv(orc_ctx *c, double mu);
/* one point's damped inverse of a 3x3 block V9 (row-major, symmetric part
 * used) as orc_stage_vinv forms it; returns 1 for the small-determinant branch */
int orc_vinv3(const double *V9, double mu, double *out9);
/* the dogleg / 2-D subspace step of orc_trust_region (PSBA/trust_region.cpp:
 * 520-595): P_U the Cauchy step, P_B the (damped) Newton step, the three
 * curvature products x^T B y, radius delta; p receives the step, the return
 * value is the "norm" the reference reports (with its :564-569 bookkeeping) */
double orc_dogleg_step(int N, double pUtBpU, double pUtBpB, double pBtBpB, double delta,
                       const double *P_U, const double *P_B, const double *g, double *p);
void orc_stage_schur(orc_ctx *c, double mu);
void orc_stage_ea(orc_ctx *c);
int orc_stage_factor_solve(orc_ctx *c, int solve_mode);
void orc_stage_backsub(orc_ctx *c);
void orc_stage_newp(orc_ctx *c, const double *dp);
void orc_stage_commit(orc_ctx *c);
void orc_jmul(const orc_ctx *c, const double *x, double *out); /* 2*n_obs */
double orc_modchol_lambda(orc_ctx *c, double *E);

/* buffer read-back for stage parity tests */
enum {
  ORC_BUF_EX = 0,  /* 2*n_obs */
  ORC_BUF_JA,      /* 2*cnp*n_obs */
  ORC_BUF_JB,      /* 6*n_obs */
  ORC_BUF_U,       /* cnp*cnp*n_cams (undamped) */
  ORC_BUF_V,       /* 9*n_pts (undamped) */
  ORC_BUF_W,       /* 3*cnp*n_obs */
  ORC_BUF_G,       /* n_tot */
  ORC_BUF_VINV,    /* 9*n_pts (damped inverse, full) */
  ORC_BUF_Y,       /* 3*cnp*n_obs */
  ORC_BUF_S,       /* n*n, n = cnp*n_cams (full symmetric) */
  ORC_BUF_EA,      /* n */
  ORC_BUF_DP,      /* n_tot */
  ORC_BUF_EB,      /* 3*n_pts */
  ORC_BUF_L,       /* n*n lower factor (after factor_solve) */
  ORC_BUF_NEWCAMS, /* cnp*n_cams */
  ORC_BUF_NEWPTS,  /* 3*n_pts */
  ORC_BUF_COUNT
};
long orc_buffer_size(const orc_ctx *c, int which);
int orc_read_buffer(const orc_ctx *c, int which, double *out);

/* ---- drivers -------------------------------------------------------------- */
/* PSBA/levmar.cpp:45-256 */
int orc_levmar(orc_ctx *c, const orc_opts *o, int *itno, double *finalErr,
               orc_trace_rec *trace, int trace_cap, int *trace_len);
/* PSBA/trust_region.cpp:49-288 */
int orc_trust_region(orc_ctx *c, const orc_opts *o, int *itno, double *finalErr,
                     orc_trace_rec *trace, int trace_cap, int *trace_len);
/* PSBA/main.cpp:193-208: LM <-> dogleg hand-off loop; returns last code */
int orc_hybrid(orc_ctx *c, const orc_opts *o, int *itno, double *initErr,
               double *finalErr, orc_trace_rec *trace, int trace_cap, int *trace_len);

#ifdef __cplusplus
}
#endif
#endif
