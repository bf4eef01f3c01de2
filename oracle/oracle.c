/*
 * oracle.c -- CPU FP64 restatement of PSBA's Schur-complement LM / dogleg path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  The product path in psba_amd/ never
 * links or calls this file.
 *
 * Each function names the reference file:line whose behaviour it restates.
 * Summation orders follow the reference kernels (observations of a camera in
 * ascending point order, of a point in ascending camera order, shared points
 * of a camera pair in ascending point order).
 *
 * Pinning status (what checks each function in tests/):
 *   orc_project / Jacobians      pinned: quaternion rotation vs an independent
 *                                rotation library, central differences, and
 *                                the pinhole special case of every model.
 *   orc_generate_idxs            pinned: brute force over vmask on the
 *                                reference's data files (bit-exact).
 *   normal equations, Vinv, Y, S pinned: dense J^T J and its Schur complement
 *                                built with numpy on small scenes; inverse vs
 *                                numpy.linalg.inv.
 *   factor/solve (both modes)    pinned: numpy Cholesky / solve, L L^T = S.
 *   orc_modchol_lambda           pinned: E == 0 on SPD input (closed form) and
 *                                L L^T = S + diag(E) on indefinite input.
 *   LM / dogleg drivers          pinned: noise-free scenes converge to cost ~0;
 *                                gain-ratio and damping identities restated
 *                                from levmar.cpp / trust_region.cpp (the
 *                                decision sequence itself has no reference
 *                                fixture: the reference publishes no runs).
 */
#include "oracle.h"

#include <float.h>
#include <limits.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* PSBA/psba.h:6-10 */
#define PSBA_INIT_MU 1e-3
#define PSBA_STOP_THRESH 1e-12
#define PSBA_EPSILON 1e-12
#define PSBA_EPSILON2 1e-12
#define PSBA_EPSILON_SQ (PSBA_EPSILON * PSBA_EPSILON)
/* PSBA/trust_region.cpp:18 */
#define MAX_DELTA 10000.0

struct orc_ctx {
  int cnp, nc, np, no, n, ntot;
  double *K, *q0, *meas;
  int *iidx, *jidx;
  unsigned char *fix; /* cnp*nc */
  double *cams, *pts, *newcams, *newpts;
  /* index structures */
  int *pt_ptr;  /* np+1: observations are point-major */
  int *cam_ptr; /* nc+1 */
  int *cam_obs; /* no: observations of each camera, ascending point */
  int npairs;
  int *pair_k, *pair_l; /* k >= l */
  long *pair_ptr;       /* npairs+1 */
  int *pair_ok, *pair_ol;
  /* work buffers */
  double *ex, *JA, *JB, *U, *V, *W, *g, *uvdiag, *Vinv, *Y, *S, *Saux, *L, *ea, *eb,
      *dp, *xtmp, *Sinv;
  int L_valid;
  /* robust loss on each observation's squared residual norm (an extension:
   * the reference minimises the plain sum of squares) */
  int loss;
  double loss_scale;
};

static int g_threads = 1;

void orc_set_threads(int n) {
  g_threads = n < 1 ? 1 : n;
#ifdef _OPENMP
  omp_set_num_threads(g_threads);
#endif
}
int orc_get_threads(void) { return g_threads; }

int orc_cnp_supported(int cnp) { return cnp == 6 || cnp == 9 || cnp == 11 || cnp == 16; }

/* ------------------------------------------------------------------------ */
/* model                                                                    */
/* ------------------------------------------------------------------------ */

/* Hamilton product a (x) b, scalar first; the composition q = q_v (x) q0 of
 * PSBA/CL_files/compute_exQT.cl:36-49. */
static void quat_mul(const double *a, const double *b, double *o) {
  o[0] = a[0] * b[0] - (a[1] * b[1] + a[2] * b[2] + a[3] * b[3]);
  o[1] = a[0] * b[1] + b[0] * a[1] + (a[2] * b[3] - a[3] * b[2]);
  o[2] = a[0] * b[2] + b[0] * a[2] + (a[3] * b[1] - a[1] * b[3]);
  o[3] = a[0] * b[3] + b[0] * a[3] + (a[1] * b[2] - a[2] * b[1]);
}

/* rotation matrix of a unit quaternion, PSBA/misc.cpp:93-113 */
static void quat_R(const double *q, double *R) {
  double s = q[0], x = q[1], y = q[2], z = q[3];
  R[0] = 1 - 2 * (y * y + z * z);
  R[1] = 2 * (x * y - z * s);
  R[2] = 2 * (x * z + y * s);
  R[3] = 2 * (x * y + z * s);
  R[4] = 1 - 2 * (x * x + z * z);
  R[5] = 2 * (y * z - x * s);
  R[6] = 2 * (x * z - y * s);
  R[7] = 2 * (y * z + x * s);
  R[8] = 1 - 2 * (x * x + y * y);
}

static void cross3(const double *a, const double *b, double *o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}

/*
 * Projection of PSBA/CL_files/compute_exQT.cl:18-71 (pose part: q = q_v (x) q0,
 * X' = R(q) M + t, u = fx x + skew y + cx, v = fx ar y + cy with x = X'/Z',
 * y = Y'/Z'), generalised to the variable-intrinsic / distortion models of the
 * camera files (data/54camsvarKD.txt:1).  The Jacobian restates the analytic
 * derivative of PSBA/CL_files/compute_jacobiQT.cl:7-141 by the chain rule.
 */
int orc_project(int cnp, const double *K5, const double *q0, const double *cam,
                const double *M, double *uv, double *JA, double *JB) {
  if (!orc_cnp_supported(cnp)) return -1;
  const double *v = cam + cnp - 6, *t = cam + cnp - 3;
  double s = sqrt(1.0 - v[0] * v[0] - v[1] * v[1] - v[2] * v[2]);
  double qv[4] = {s, v[0], v[1], v[2]}, q[4], R[9];
  quat_mul(qv, q0, q);
  quat_R(q, R);
  double X[3];
  for (int r = 0; r < 3; r++) X[r] = R[3 * r] * M[0] + R[3 * r + 1] * M[1] + R[3 * r + 2] * M[2] + t[r];
  double iz = 1.0 / X[2];
  double x = X[0] * iz, y = X[1] * iz;
  double fx, cx, cy, ar, sk;
  if (cnp == 6) {
    fx = K5[0]; cx = K5[1]; cy = K5[2]; ar = K5[3]; sk = K5[4];
  } else if (cnp == 9) {
    fx = cam[0]; cx = K5[1]; cy = K5[2]; ar = K5[3]; sk = K5[4];
  } else {
    fx = cam[0]; cx = cam[1]; cy = cam[2]; ar = cam[3]; sk = cam[4];
  }
  double r2 = x * x + y * y;
  double xd = x, yd = y;
  double dxx = 1, dxy = 0, dyx = 0, dyy = 1; /* d(xd,yd)/d(x,y) */
  if (cnp == 9) {
    double k1 = cam[1], k2 = cam[2];
    double d = 1.0 + k1 * r2 + k2 * r2 * r2;
    double dd = k1 + 2.0 * k2 * r2; /* d d / d r2 */
    xd = d * x;
    yd = d * y;
    dxx = d + 2.0 * x * x * dd;
    dxy = 2.0 * x * y * dd;
    dyx = dxy;
    dyy = d + 2.0 * y * y * dd;
  } else if (cnp == 16) {
    const double *kc = cam + 5;
    double rad = 1.0 + kc[0] * r2 + kc[1] * r2 * r2 + kc[4] * r2 * r2 * r2;
    double drad = kc[0] + 2.0 * kc[1] * r2 + 3.0 * kc[4] * r2 * r2;
    xd = x * rad + 2.0 * kc[2] * x * y + kc[3] * (r2 + 2.0 * x * x);
    yd = y * rad + kc[2] * (r2 + 2.0 * y * y) + 2.0 * kc[3] * x * y;
    dxx = rad + 2.0 * x * x * drad + 2.0 * kc[2] * y + 6.0 * kc[3] * x;
    dxy = 2.0 * x * y * drad + 2.0 * kc[2] * x + 2.0 * kc[3] * y;
    dyx = dxy;
    dyy = rad + 2.0 * y * y * drad + 6.0 * kc[2] * y + 2.0 * kc[3] * x;
  }
  uv[0] = fx * xd + sk * yd + cx;
  uv[1] = fx * ar * yd + cy;
  if (!JA && !JB) return 0;

  /* d(u,v)/d(xd,yd) = [[fx, sk],[0, fx ar]] ; then distortion ; then pinhole */
  double P00 = fx, P01 = sk, P11 = fx * ar;
  double A00 = P00 * dxx + P01 * dyx, A01 = P00 * dxy + P01 * dyy;
  double A10 = P11 * dyx, A11 = P11 * dyy;
  double G[6]; /* d(u,v)/dX' */
  G[0] = A00 * iz; G[1] = A01 * iz; G[2] = -(A00 * x + A01 * y) * iz;
  G[3] = A10 * iz; G[4] = A11 * iz; G[5] = -(A10 * x + A11 * y) * iz;
  if (JB) {
    for (int r = 0; r < 2; r++)
      for (int c = 0; c < 3; c++)
        JB[3 * r + c] = G[3 * r] * R[c] + G[3 * r + 1] * R[3 + c] + G[3 * r + 2] * R[6 + c];
  }
  if (JA) {
    /* pose columns: X' = R(q_v) M0 + t with M0 = R(q0) M */
    double R0[9], M0[3];
    quat_R(q0, R0);
    for (int r = 0; r < 3; r++) M0[r] = R0[3 * r] * M[0] + R0[3 * r + 1] * M[1] + R0[3 * r + 2] * M[2];
    double a[3];
    cross3(v, M0, a); /* v x M0 */
    int cv = cnp - 6;
    for (int k = 0; k < 3; k++) {
      double ek[3] = {0, 0, 0};
      ek[k] = 1.0;
      double ds = -v[k] / s;
      double t1[3], t2[3], t3[3];
      cross3(ek, M0, t1);  /* e_k x M0 */
      cross3(ek, a, t2);   /* e_k x (v x M0) */
      cross3(v, t1, t3);   /* v x (e_k x M0) */
      double D[3];
      for (int r = 0; r < 3; r++) D[r] = 2.0 * ds * a[r] + 2.0 * s * t1[r] + 2.0 * (t2[r] + t3[r]);
      for (int r = 0; r < 2; r++) JA[r * cnp + cv + k] = G[3 * r] * D[0] + G[3 * r + 1] * D[1] + G[3 * r + 2] * D[2];
    }
    for (int k = 0; k < 3; k++)
      for (int r = 0; r < 2; r++) JA[r * cnp + cv + 3 + k] = G[3 * r + k];
    /* intrinsic / distortion columns */
    if (cnp == 9) {
      JA[0] = xd; JA[cnp + 0] = ar * yd;                     /* f  */
      JA[1] = P00 * x * r2 + P01 * y * r2; JA[cnp + 1] = P11 * y * r2;           /* k1 */
      JA[2] = P00 * x * r2 * r2 + P01 * y * r2 * r2; JA[cnp + 2] = P11 * y * r2 * r2; /* k2 */
    } else if (cnp == 11 || cnp == 16) {
      JA[0] = xd; JA[cnp + 0] = ar * yd; /* fx */
      JA[1] = 1.0; JA[cnp + 1] = 0.0;    /* cx */
      JA[2] = 0.0; JA[cnp + 2] = 1.0;    /* cy */
      JA[3] = 0.0; JA[cnp + 3] = fx * yd; /* ar */
      JA[4] = yd; JA[cnp + 4] = 0.0;     /* skew */
      if (cnp == 16) {
        double r4 = r2 * r2, r6 = r4 * r2;
        /* d(xd,yd)/dkc for kc1..kc5 */
        double dx[5] = {x * r2, x * r4, 2.0 * x * y, r2 + 2.0 * x * x, x * r6};
        double dy[5] = {y * r2, y * r4, r2 + 2.0 * y * y, 2.0 * x * y, y * r6};
        for (int k = 0; k < 5; k++) {
          JA[5 + k] = P00 * dx[k] + P01 * dy[k];
          JA[cnp + 5 + k] = P11 * dy[k];
        }
      }
    }
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* integer indexing                                                         */
/* ------------------------------------------------------------------------ */

/* PSBA/misc.cpp:178-218: observation order point-major, camera-ascending;
 * blk_idx[i*nCams+j] = observation or -1; camera-pair shared-point lists,
 * ascending point id, stored for both (j,k) and (k,j). */
void orc_generate_idxs(int nCams, int n3Dpts, const char *vmask, int *comm3DIdx,
                       int *comm3DIdxCnt, int *iidx, int *jidx, int *blk_idx) {
  long rowsize = (long)nCams * n3Dpts;
  memset(comm3DIdxCnt, 0, sizeof(int) * (size_t)nCams * nCams);
  int k = 0;
  for (int i = 0; i < n3Dpts; i++) {
    for (int j = 0; j < nCams; j++) {
      if (vmask[(long)i * nCams + j]) {
        iidx[k] = i;
        jidx[k] = j;
        blk_idx[(long)i * nCams + j] = k;
        for (int l = 0; l <= j; l++) {
          if (vmask[(long)i * nCams + l]) {
            int t = comm3DIdxCnt[j * nCams + l]++;
            if (comm3DIdx) comm3DIdx[j * rowsize + (long)l * n3Dpts + t] = i;
            if (l != j) {
              comm3DIdxCnt[l * nCams + j]++;
              if (comm3DIdx) comm3DIdx[l * rowsize + (long)j * n3Dpts + t] = i;
            }
          }
        }
        k++;
      } else {
        blk_idx[(long)i * nCams + j] = -1;
      }
    }
  }
}

/* ------------------------------------------------------------------------ */
/* context                                                                  */
/* ------------------------------------------------------------------------ */

static void *xcalloc(size_t n, size_t sz) {
  void *p = calloc(n ? n : 1, sz);
  return p;
}

orc_ctx *orc_create(const orc_problem *p) {
  if (!p || !orc_cnp_supported(p->cnp) || p->n_cams <= 0 || p->n_pts <= 0 || p->n_obs <= 0) return NULL;
  orc_ctx *c = (orc_ctx *)xcalloc(1, sizeof(orc_ctx));
  int C = p->cnp, nc = p->n_cams, np = p->n_pts, no = p->n_obs;
  c->cnp = C; c->nc = nc; c->np = np; c->no = no;
  c->n = C * nc; c->ntot = C * nc + 3 * np;
  c->K = (double *)xcalloc((size_t)5 * nc, sizeof(double));
  c->q0 = (double *)xcalloc((size_t)4 * nc, sizeof(double));
  c->meas = (double *)xcalloc((size_t)2 * no, sizeof(double));
  c->iidx = (int *)xcalloc(no, sizeof(int));
  c->jidx = (int *)xcalloc(no, sizeof(int));
  c->fix = (unsigned char *)xcalloc((size_t)C * nc, 1);
  c->cams = (double *)xcalloc((size_t)C * nc, sizeof(double));
  c->pts = (double *)xcalloc((size_t)3 * np, sizeof(double));
  c->newcams = (double *)xcalloc((size_t)C * nc, sizeof(double));
  c->newpts = (double *)xcalloc((size_t)3 * np, sizeof(double));
  memcpy(c->K, p->K, sizeof(double) * 5 * nc);
  memcpy(c->q0, p->q0, sizeof(double) * 4 * nc);
  memcpy(c->meas, p->meas, sizeof(double) * 2 * no);
  memcpy(c->iidx, p->iidx, sizeof(int) * no);
  memcpy(c->jidx, p->jidx, sizeof(int) * no);
  if (p->fixmask) memcpy(c->fix, p->fixmask, (size_t)C * nc);
  memcpy(c->cams, p->cams, sizeof(double) * C * nc);
  memcpy(c->pts, p->pts, sizeof(double) * 3 * np);

  /* point CSR (observations must be point-major) */
  c->pt_ptr = (int *)xcalloc(np + 1, sizeof(int));
  for (int k = 0; k < no; k++) c->pt_ptr[c->iidx[k] + 1]++;
  for (int i = 0; i < np; i++) c->pt_ptr[i + 1] += c->pt_ptr[i];
  /* camera CSR, ascending point within a camera */
  c->cam_ptr = (int *)xcalloc(nc + 1, sizeof(int));
  c->cam_obs = (int *)xcalloc(no, sizeof(int));
  for (int k = 0; k < no; k++) c->cam_ptr[c->jidx[k] + 1]++;
  for (int j = 0; j < nc; j++) c->cam_ptr[j + 1] += c->cam_ptr[j];
  int *fill = (int *)xcalloc(nc, sizeof(int));
  for (int k = 0; k < no; k++) {
    int j = c->jidx[k];
    c->cam_obs[c->cam_ptr[j] + fill[j]++] = k;
  }
  free(fill);
  /* camera pairs k >= l with their shared points, ascending point */
  long *cnt = (long *)xcalloc((size_t)nc * nc, sizeof(long));
  for (int i = 0; i < np; i++)
    for (int a = c->pt_ptr[i]; a < c->pt_ptr[i + 1]; a++)
      for (int b = c->pt_ptr[i]; b <= a; b++) cnt[(long)c->jidx[a] * nc + c->jidx[b]]++;
  int npairs = 0;
  for (long q = 0; q < (long)nc * nc; q++) npairs += cnt[q] > 0;
  c->npairs = npairs;
  c->pair_k = (int *)xcalloc(npairs, sizeof(int));
  c->pair_l = (int *)xcalloc(npairs, sizeof(int));
  c->pair_ptr = (long *)xcalloc(npairs + 1, sizeof(long));
  long *pid = (long *)xcalloc((size_t)nc * nc, sizeof(long));
  int pp = 0;
  for (int k = 0; k < nc; k++)
    for (int l = 0; l <= k; l++)
      if (cnt[(long)k * nc + l]) {
        c->pair_k[pp] = k;
        c->pair_l[pp] = l;
        c->pair_ptr[pp + 1] = c->pair_ptr[pp] + cnt[(long)k * nc + l];
        pid[(long)k * nc + l] = pp;
        pp++;
      }
  long tot = c->pair_ptr[npairs];
  c->pair_ok = (int *)xcalloc(tot, sizeof(int));
  c->pair_ol = (int *)xcalloc(tot, sizeof(int));
  long *pf = (long *)xcalloc(npairs, sizeof(long));
  for (int i = 0; i < np; i++)
    for (int a = c->pt_ptr[i]; a < c->pt_ptr[i + 1]; a++)
      for (int b = c->pt_ptr[i]; b <= a; b++) {
        long q = pid[(long)c->jidx[a] * nc + c->jidx[b]];
        long at = c->pair_ptr[q] + pf[q]++;
        c->pair_ok[at] = a;
        c->pair_ol[at] = b;
      }
  free(pf);
  free(pid);
  free(cnt);

  c->ex = (double *)xcalloc((size_t)2 * no, sizeof(double));
  c->JA = (double *)xcalloc((size_t)2 * C * no, sizeof(double));
  c->JB = (double *)xcalloc((size_t)6 * no, sizeof(double));
  c->U = (double *)xcalloc((size_t)C * C * nc, sizeof(double));
  c->V = (double *)xcalloc((size_t)9 * np, sizeof(double));
  c->W = (double *)xcalloc((size_t)3 * C * no, sizeof(double));
  c->Y = (double *)xcalloc((size_t)3 * C * no, sizeof(double));
  c->g = (double *)xcalloc(c->ntot, sizeof(double));
  c->uvdiag = (double *)xcalloc(c->ntot, sizeof(double));
  c->Vinv = (double *)xcalloc((size_t)9 * np, sizeof(double));
  c->S = (double *)xcalloc((size_t)c->n * c->n, sizeof(double));
  c->Saux = (double *)xcalloc((size_t)c->n * c->n, sizeof(double));
  c->L = (double *)xcalloc((size_t)c->n * c->n, sizeof(double));
  c->ea = (double *)xcalloc(c->n, sizeof(double));
  c->eb = (double *)xcalloc((size_t)3 * np, sizeof(double));
  c->dp = (double *)xcalloc(c->ntot, sizeof(double));
  c->xtmp = (double *)xcalloc(c->ntot, sizeof(double));
  c->Sinv = (double *)xcalloc((size_t)c->n * c->n, sizeof(double));
  return c;
}

void orc_destroy(orc_ctx *c) {
  if (!c) return;
  void *ptrs[] = {c->K, c->q0, c->meas, c->iidx, c->jidx, c->fix, c->cams, c->pts, c->newcams,
                  c->newpts, c->pt_ptr, c->cam_ptr, c->cam_obs, c->pair_k, c->pair_l,
                  c->pair_ptr, c->pair_ok, c->pair_ol, c->ex, c->JA, c->JB, c->U, c->V, c->W,
                  c->Y, c->g, c->uvdiag, c->Vinv, c->S, c->Saux, c->L, c->ea, c->eb, c->dp,
                  c->xtmp, c->Sinv};
  for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); i++) free(ptrs[i]);
  free(c);
}

void orc_get_params(const orc_ctx *c, double *cams, double *pts) {
  if (cams) memcpy(cams, c->cams, sizeof(double) * c->n);
  if (pts) memcpy(pts, c->pts, sizeof(double) * 3 * c->np);
}

void orc_set_params(orc_ctx *c, const double *cams, const double *pts) {
  if (cams) memcpy(c->cams, cams, sizeof(double) * c->n);
  if (pts) memcpy(c->pts, pts, sizeof(double) * 3 * c->np);
}

/* ------------------------------------------------------------------------ */
/* reductions, PSBA/misc.cpp:151-169 (serial FP64, index order)              */
/* ------------------------------------------------------------------------ */
/* Summation order of every reduction below: ascending index (the
 * reference's order, default) or, for experiments on how far two valid FP64
 * evaluation orders stay together, descending.  RIDX maps the k-th step of a
 * loop over [b, e) to the index it adds. */
static int g_reverse = 0;
#define RIDX(b, e, k) (g_reverse ? (e) - 1 - ((k) - (b)) : (k))

void orc_set_sum_order(int descending) { g_reverse = descending ? 1 : 0; }

static double l2sq(long n, const double *x) {
  double s = 0;
  for (long i0 = 0; i0 < n; i0++) {
    const long i = RIDX(0, n, i0);
    s += x[i] * x[i];
  }
  return s;
}
static double dotp(const double *a, const double *b, long n) {
  double s = 0.0;
  for (long i0 = 0; i0 < n; i0++) {
    const long i = RIDX(0, n, i0);
    s += a[i] * b[i];
  }
  return s;
}

/* ------------------------------------------------------------------------ */
/* stages                                                                   */
/* ------------------------------------------------------------------------ */

/* Robust loss of one observation's squared residual norm s, with scale d:
 *   Huber   rho(s) = s for s <= d^2, 2 d sqrt(s) - d^2 beyond; rho' = 1 or d / sqrt(s)
 *   Cauchy  rho(s) = d^2 log(1 + s / d^2);                       rho' = 1 / (1 + s / d^2)
 * Steps use the reweighted problem: residual and Jacobian rows of the
 * observation scaled by sqrt(rho'(s)) (iteratively reweighted least squares);
 * the cost that accepts or rejects a step is sum rho(s). */
void orc_loss_eval(int kind, double d, double s, double *rho, double *drho) {
  const double d2 = d * d;
  if (kind == ORC_LOSS_HUBER) {
    if (s <= d2) {
      *rho = s;
      *drho = 1.0;
    } else {
      const double r = sqrt(s);
      *rho = 2.0 * d * r - d2;
      *drho = d / r;
    }
  } else if (kind == ORC_LOSS_CAUCHY) {
    *rho = d2 * log1p(s / d2);
    *drho = 1.0 / (1.0 + s / d2);
  } else {
    *rho = s;
    *drho = 1.0;
  }
}

int orc_set_loss(orc_ctx *c, int kind, double scale) {
  if (kind != ORC_LOSS_NONE && kind != ORC_LOSS_HUBER && kind != ORC_LOSS_CAUCHY) return -1;
  if (kind != ORC_LOSS_NONE && !(scale > 0)) return -1;
  c->loss = kind;
  c->loss_scale = scale;
  return 0;
}

/* residual e = m - pi(p) per observation (PSBA/CL_files/compute_exQT.cl:18-71)
 * and its squared norm (PSBA/misc.cpp:151-157), or the robust cost */
double orc_stage_residual(orc_ctx *c, int use_new, double *ex) {
  const double *cams = use_new ? c->newcams : c->cams;
  const double *pts = use_new ? c->newpts : c->pts;
  int C = c->cnp;
#pragma omp parallel for schedule(static)
  for (int k = 0; k < c->no; k++) {
    int i = c->iidx[k], j = c->jidx[k];
    double uv[2];
    orc_project(C, c->K + 5 * j, c->q0 + 4 * j, cams + (long)C * j, pts + 3 * i, uv, NULL, NULL);
    c->ex[2 * k] = c->meas[2 * k] - uv[0];
    c->ex[2 * k + 1] = c->meas[2 * k + 1] - uv[1];
  }
  if (ex) memcpy(ex, c->ex, sizeof(double) * 2 * c->no);
  if (c->loss == ORC_LOSS_NONE) return l2sq(2L * c->no, c->ex);
  double sum = 0;
  for (int k0 = 0; k0 < c->no; k0++) {
    const int k = RIDX(0, c->no, k0);
    double rho, drho;
    orc_loss_eval(c->loss, c->loss_scale, c->ex[2 * k] * c->ex[2 * k] + c->ex[2 * k + 1] * c->ex[2 * k + 1], &rho,
                  &drho);
    sum += rho;
  }
  return sum;
}

/* JA (2 x cnp) and JB (2 x 3) per observation at the current parameters,
 * PSBA/CL_files/compute_jacobiQT.cl:7-141; columns of held-fixed parameters
 * are zero.  Also refreshes the residual at the current parameters. */
void orc_stage_jacobian(orc_ctx *c) {
  int C = c->cnp;
#pragma omp parallel for schedule(static)
  for (int k = 0; k < c->no; k++) {
    int i = c->iidx[k], j = c->jidx[k];
    double uv[2];
    double *ja = c->JA + (long)2 * C * k;
    orc_project(C, c->K + 5 * j, c->q0 + 4 * j, c->cams + (long)C * j, c->pts + 3 * i, uv, ja,
                c->JB + 6L * k);
    for (int col = 0; col < C; col++)
      if (c->fix[(long)C * j + col]) ja[col] = ja[C + col] = 0.0;
    c->ex[2 * k] = c->meas[2 * k] - uv[0];
    c->ex[2 * k + 1] = c->meas[2 * k + 1] - uv[1];
    if (c->loss != ORC_LOSS_NONE) {
      double rho, drho;
      orc_loss_eval(c->loss, c->loss_scale, c->ex[2 * k] * c->ex[2 * k] + c->ex[2 * k + 1] * c->ex[2 * k + 1], &rho,
                    &drho);
      const double w = sqrt(drho);
      for (int q = 0; q < 2 * C; q++) ja[q] *= w;
      for (int q = 0; q < 6; q++) c->JB[6L * k + q] *= w;
      c->ex[2 * k] *= w;
      c->ex[2 * k + 1] *= w;
    }
  }
}

/* U_j  = coeff * sum_i A_ij^T A_ij   PSBA/CL_files/compute_U.cl:5-35
 * V_i  = coeff * sum_j B_ij^T B_ij   PSBA/CL_files/compute_V.cl:6-38
 * W_ij = coeff * A_ij^T B_ij         PSBA/CL_files/compute_Wblks.cl:7-34
 * g    = coeff_g * J^T e             PSBA/CL_files/compute_g.cl:6-60
 * UVdiag = undamped diagonals        PSBA/CL_files/compute_U.cl:33-34 */
void orc_stage_normal(orc_ctx *c, double coeff, double coeff_g) {
  int C = c->cnp;
#pragma omp parallel for schedule(dynamic, 1)
  for (int j = 0; j < c->nc; j++) {
    for (int r = 0; r < C; r++) {
      for (int cc = 0; cc < C; cc++) {
        double sum = 0;
        for (int a0 = c->cam_ptr[j]; a0 < c->cam_ptr[j + 1]; a0++) {
          const int a = RIDX(c->cam_ptr[j], c->cam_ptr[j + 1], a0);
          const double *ja = c->JA + (long)2 * C * c->cam_obs[a];
          sum = sum + ja[r] * ja[cc] + ja[C + r] * ja[C + cc];
        }
        sum = coeff * sum;
        c->U[(long)C * C * j + C * r + cc] = sum;
        if (r == cc) c->uvdiag[C * j + r] = sum;
      }
      double sg = 0;
      for (int a0 = c->cam_ptr[j]; a0 < c->cam_ptr[j + 1]; a0++) {
        const int a = RIDX(c->cam_ptr[j], c->cam_ptr[j + 1], a0);
        int o = c->cam_obs[a];
        const double *ja = c->JA + (long)2 * C * o;
        sg = sg + ja[r] * c->ex[2 * o] + ja[C + r] * c->ex[2 * o + 1];
      }
      c->g[C * j + r] = coeff_g * sg;
    }
  }
#pragma omp parallel for schedule(static)
  for (int i = 0; i < c->np; i++) {
    for (int r = 0; r < 3; r++) {
      for (int cc = 0; cc < 3; cc++) {
        double sum = 0;
        for (int o0 = c->pt_ptr[i]; o0 < c->pt_ptr[i + 1]; o0++) {
          const int o = RIDX(c->pt_ptr[i], c->pt_ptr[i + 1], o0);
          const double *jb = c->JB + 6L * o;
          sum = sum + jb[r] * jb[cc] + jb[3 + r] * jb[3 + cc];
        }
        sum = coeff * sum;
        c->V[9L * i + 3 * r + cc] = sum;
        if (r == cc) c->uvdiag[c->n + 3 * i + r] = sum;
      }
      double sg = 0;
      for (int o0 = c->pt_ptr[i]; o0 < c->pt_ptr[i + 1]; o0++) {
        const int o = RIDX(c->pt_ptr[i], c->pt_ptr[i + 1], o0);
        const double *jb = c->JB + 6L * o;
        sg = sg + jb[r] * c->ex[2 * o] + jb[3 + r] * c->ex[2 * o + 1];
      }
      c->g[c->n + 3 * i + r] = coeff_g * sg;
    }
  }
#pragma omp parallel for schedule(static)
  for (int o = 0; o < c->no; o++) {
    const double *ja = c->JA + (long)2 * C * o;
    const double *jb = c->JB + 6L * o;
    for (int r = 0; r < C; r++)
      for (int cc = 0; cc < 3; cc++) {
        double sum = ja[r] * jb[cc] + ja[C + r] * jb[3 + cc];
        c->W[(long)3 * C * o + 3 * r + cc] = coeff * sum;
      }
  }
}

/* mu0 = tau * max(UVdiag), PSBA/CL_files/../PSBA/sba_func.cpp:422-444 */
double orc_max_uvdiag(const orc_ctx *c) {
  double m = 0.0;
  for (int i = 0; i < c->ntot; i++)
    if (c->uvdiag[i] > m) m = c->uvdiag[i];
  return m;
}

/*
 * Batched 3x3 inverse of the damped point blocks V_i + mu I,
 * PSBA/CL_files/update_UV.cl:5-31 + PSBA/CL_files/compute_Vinv.cl:6-90.
 * Regular branch: adjugate over T = -det.  When |T| < 1e-16 the reference
 * takes a pivoted-elimination fallback whose divisor is the product of the
 * eliminated diagonal (signed pivot tests; the `a[0,0] != 0` test at
 * compute_Vinv.cl:41 is always true).  Returns 1 if any point took the
 * fallback (the reference's racy `ret`, here OR-reduced).
 */
/* one point's damped 3x3 inverse (PSBA/CL_files/compute_Vinv.cl:6-90) into
 * inv[9]; returns 1 when the small-determinant branch (:31-73) was taken */
static int vinv3(const double *Vi, double mu, double *inv_out) {
  int any = 0;
  {
    double a11 = Vi[0] + mu, a12 = Vi[1], a13 = Vi[2];
    double a22 = Vi[4] + mu, a23 = Vi[5], a33 = Vi[8] + mu;
    double a21 = a12, a31 = a13, a32 = a23;
    double T = a33 * a12 * a12 - 2 * a12 * a13 * a23 + a22 * a13 * a13 + a11 * a23 * a23 - a11 * a22 * a33;
    double inv[9];
    if (fabs(T) < 1e-16) {
      any = 1;
      double m[3][3] = {{a11, a12, a13}, {a21, a22, a23}, {a31, a32, a33}}, tmp;
      int p = 0;
      if (m[0][0] < m[1][0]) p = 1;
      if (m[p][0] < m[2][0]) p = 2;
      if (p != 0)
        for (int q = 0; q < 3; q++) { tmp = m[0][q]; m[0][q] = m[p][q]; m[p][q] = tmp; }
      m[1][0] = m[1][0] / m[0][0];
      m[2][0] = m[2][0] / m[0][0];
      m[1][1] = m[1][1] - m[1][0] * m[0][1];
      m[1][2] = m[1][2] - m[1][0] * m[0][2];
      m[2][1] = m[2][1] - m[2][0] * m[0][1];
      m[2][2] = m[2][2] - m[2][0] * m[0][2];
      if (m[1][1] < m[2][1])
        for (int q = 0; q < 3; q++) { tmp = m[1][q]; m[1][q] = m[2][q]; m[2][q] = tmp; }
      if (m[1][1] != 0.0) {
        m[2][1] = m[2][1] / m[1][1];
        m[2][2] = m[2][2] - m[2][1] * m[1][2];
      }
      double D = m[0][0] * m[1][1] * m[2][2];
      inv[0] = (a22 * a33 - a23 * a32) / D;
      inv[3] = -(a21 * a33 - a23 * a31) / D;
      inv[4] = (a11 * a33 - a13 * a31) / D;
      inv[6] = (a21 * a32 - a22 * a31) / D;
      inv[7] = -(a11 * a32 - a12 * a31) / D;
      inv[8] = (a11 * a22 - a12 * a21) / D;
    } else {
      inv[0] = -(-a23 * a23 + a22 * a33) / T;
      inv[3] = -(a13 * a23 - a12 * a33) / T;
      inv[4] = -(-a13 * a13 + a11 * a33) / T;
      inv[6] = -(a12 * a23 - a13 * a22) / T;
      inv[7] = -(a12 * a13 - a11 * a23) / T;
      inv[8] = -(-a12 * a12 + a11 * a22) / T;
    }
    /* the reference keeps only the lower triangle; read it symmetrically */
    inv[1] = inv[3]; inv[2] = inv[6]; inv[5] = inv[7];
    memcpy(inv_out, inv, sizeof(inv));
  }
  return any;
}

int orc_vinv3(const double *V9, double mu, double *out9) { return vinv3(V9, mu, out9); }

int orc_stage_vinv(orc_ctx *c, double mu) {
  int any = 0;
#pragma omp parallel for schedule(static) reduction(| : any)
  for (int i = 0; i < c->np; i++) any |= vinv3(c->V + 9L * i, mu, c->Vinv + 9L * i);
  return any;
}

/*
 * Y_ij = W_ij V_i^-1 (PSBA/CL_files/compute_Yblks.cl:6-39) and the reduced
 * camera matrix S_kl = delta_kl (U_k + mu I) - sum_{i in P(k,l)} Y_ik W_il^T
 * (PSBA/CL_files/update_UV.cl + PSBA/CL_files/compute_S.cl:6-78).  Blocks with
 * k >= l are summed over the shared points in ascending point order and the
 * strictly lower blocks are mirrored.  Diagonal entries of held-fixed
 * parameters are set to 1 (their rows and columns are otherwise zero).
 */
void orc_stage_schur(orc_ctx *c, double mu) {
  int C = c->cnp, n = c->n;
#pragma omp parallel for schedule(static)
  for (int o = 0; o < c->no; o++) {
    const double *w = c->W + (long)3 * C * o;
    const double *vi = c->Vinv + 9L * c->iidx[o];
    double *y = c->Y + (long)3 * C * o;
    for (int r = 0; r < C; r++)
      for (int cc = 0; cc < 3; cc++)
        y[3 * r + cc] = w[3 * r] * vi[cc] + w[3 * r + 1] * vi[3 + cc] + w[3 * r + 2] * vi[6 + cc];
  }
  memset(c->S, 0, sizeof(double) * (size_t)n * n);
#pragma omp parallel for schedule(dynamic, 4)
  for (int q = 0; q < c->npairs; q++) {
    int k = c->pair_k[q], l = c->pair_l[q];
    for (int r = 0; r < C; r++)
      for (int cc = 0; cc < C; cc++) {
        double sum = 0;
        for (long it0 = c->pair_ptr[q]; it0 < c->pair_ptr[q + 1]; it0++) {
          const long it = RIDX(c->pair_ptr[q], c->pair_ptr[q + 1], it0);
          const double *y = c->Y + (long)3 * C * c->pair_ok[it] + 3 * r;
          const double *w = c->W + (long)3 * C * c->pair_ol[it] + 3 * cc;
          sum += y[0] * w[0] + y[1] * w[1] + y[2] * w[2];
        }
        if (k == l) {
          double u = c->U[(long)C * C * k + C * r + cc];
          if (r == cc) u = u + mu;
          sum = u - sum;
        } else {
          sum = -sum;
        }
        c->S[(long)(k * C + r) * n + l * C + cc] = sum;
      }
  }
  /* diagonal blocks of cameras without shared points (no observations) */
  for (int k = 0; k < c->nc; k++)
    if (c->cam_ptr[k + 1] == c->cam_ptr[k])
      for (int r = 0; r < C; r++) {
        for (int cc = 0; cc < C; cc++) {
          double u = c->U[(long)C * C * k + C * r + cc];
          if (r == cc) u = u + mu;
          c->S[(long)(k * C + r) * n + k * C + cc] = u;
        }
      }
  for (int q = 0; q < c->npairs; q++) {
    int k = c->pair_k[q], l = c->pair_l[q];
    if (k == l) continue;
    for (int r = 0; r < C; r++)
      for (int cc = 0; cc < C; cc++)
        c->S[(long)(l * C + cc) * n + k * C + r] = c->S[(long)(k * C + r) * n + l * C + cc];
  }
  for (int p = 0; p < n; p++)
    if (c->fix[p]) c->S[(long)p * n + p] = 1.0;
}

/* ea_j = g_a,j - sum_i Y_ij g_b,i, PSBA/CL_files/compute_ea.cl:6-37 */
void orc_stage_ea(orc_ctx *c) {
  int C = c->cnp;
#pragma omp parallel for schedule(static)
  for (int j = 0; j < c->nc; j++)
    for (int r = 0; r < C; r++) {
      double sum = 0;
      for (int a0 = c->cam_ptr[j]; a0 < c->cam_ptr[j + 1]; a0++) {
        const int a = RIDX(c->cam_ptr[j], c->cam_ptr[j + 1], a0);
        int o = c->cam_obs[a];
        const double *y = c->Y + (long)3 * C * o + 3 * r;
        const double *gb = c->g + c->n + 3 * c->iidx[o];
        sum = sum + y[0] * gb[0] + y[1] * gb[1] + y[2] * gb[2];
      }
      c->ea[C * j + r] = c->g[C * j + r] - sum;
    }
}

/* plain column Cholesky S = L L^T on the lower triangle; 1 = success */
static int chol_plain(int n, const double *S, double *L) {
  memset(L, 0, sizeof(double) * (size_t)n * n);
  for (int j = 0; j < n; j++) {
    const double *Lj = L + (long)j * n;
    double d = S[(long)j * n + j];
    for (int k = 0; k < j; k++) d -= Lj[k] * Lj[k];
    if (!(d > 0.0) || !isfinite(d)) return 0;
    double ljj = sqrt(d);
    L[(long)j * n + j] = ljj;
    int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
    for (int i = j + 1; i < n; i++) {
      double *Li = L + (long)i * n;
      double s = S[(long)i * n + j];
      for (int k = 0; k < j; k++) s -= Li[k] * Lj[k];
      Li[j] = s / ljj;
      if (!isfinite(Li[j])) bad = 1;
    }
    if (bad) return 0;
  }
  return 1;
}

/*
 * Reference-faithful explicit inverse, PSBA/PSBA/cl_spdinv.cpp:18-204 with
 * PSBA/CL_files/SPD_inv.cl:20-411: 3x3-block left-looking Cholesky with
 * closed-form diagonal blocks and their inverses (fails on any non-finite
 * value), block-diagonal-by-diagonal L^-1, then S^-1 = L^-T L^-1.
 * Returns 1 on success; Sinv is n x n.
 */
static int spd_inverse_ref(int n, const double *S, double *L, double *Sinv) {
  int nb = n / 3;
  double *dinv = (double *)xcalloc((size_t)9 * nb, sizeof(double));
  double *X = (double *)xcalloc((size_t)n * n, sizeof(double));
  memset(L, 0, sizeof(double) * (size_t)n * n);
  int ok = 1;
  for (int j = 0; j < nb && ok; j++) {
    double T[9];
    for (int u = 0; u < 3; u++)
      for (int v = 0; v < 3; v++) {
        double s = S[(long)(3 * j + u) * n + 3 * j + v];
        for (int k = 0; k < 3 * j; k++) s -= L[(long)(3 * j + u) * n + k] * L[(long)(3 * j + v) * n + k];
        T[3 * u + v] = s;
      }
    double l0 = sqrt(T[0]);
    double l1 = T[3] / sqrt(T[0]);
    double l2 = sqrt(T[4] - T[3] * T[3] / T[0]);
    double l3 = T[6] / sqrt(T[0]);
    double l4 = sqrt(T[0] / (T[0] * T[4] - T[3] * T[3])) * (T[7] - T[3] * T[6] / T[0]);
    double t1 = -T[8] * T[3] * T[3], t2 = 2 * T[7] * T[3] * T[6], t3 = -T[4] * T[6] * T[6];
    double t4 = T[0] * (T[4] * T[8] - T[7] * T[7]), t5 = -T[3] * T[3] + T[0] * T[4];
    double l5 = sqrt((t1 + t2 + t3 + t4) / t5);
    double lv[6] = {l0, l1, l2, l3, l4, l5};
    for (int q = 0; q < 6; q++)
      if (!isfinite(lv[q])) ok = 0;
    double *di = dinv + 9 * j;
    di[0] = 1 / l0; di[1] = 0; di[2] = 0;
    di[3] = -l1 / (l0 * l2); di[4] = 1 / l2; di[5] = 0;
    di[6] = (l1 * l4 - l2 * l3) / (l0 * l2 * l5); di[7] = -l4 / (l2 * l5); di[8] = 1 / l5;
    for (int q = 0; q < 9; q++)
      if (!isfinite(di[q])) ok = 0;
    if (!ok) break;
    long b = (long)(3 * j) * n + 3 * j;
    L[b] = l0; L[b + n] = l1; L[b + n + 1] = l2; L[b + 2 * n] = l3; L[b + 2 * n + 1] = l4; L[b + 2 * n + 2] = l5;
#pragma omp parallel for schedule(static)
    for (int i = j + 1; i < nb; i++) {
      double Tij[9];
      for (int u = 0; u < 3; u++)
        for (int v = 0; v < 3; v++) {
          double s = S[(long)(3 * i + u) * n + 3 * j + v];
          for (int k = 0; k < 3 * j; k++) s -= L[(long)(3 * i + u) * n + k] * L[(long)(3 * j + v) * n + k];
          Tij[3 * u + v] = s;
        }
      for (int u = 0; u < 3; u++)
        for (int v = 0; v < 3; v++)
          L[(long)(3 * i + u) * n + 3 * j + v] =
              Tij[3 * u] * di[3 * v] + Tij[3 * u + 1] * di[3 * v + 1] + Tij[3 * u + 2] * di[3 * v + 2];
    }
  }
  if (ok) {
    /* X = L^-1, diagonal blocks from the closed-form inverses, then by block
     * diagonals: X_ij = -L_ii^-1 sum_{k=j}^{i-1} L_ik X_kj (SPD_inv.cl:248-328) */
    for (int i = 0; i < nb; i++)
      for (int u = 0; u < 3; u++)
        for (int v = 0; v < 3; v++) X[(long)(3 * i + u) * n + 3 * i + v] = dinv[9 * i + 3 * u + v];
    for (int dd = 1; dd < nb; dd++) {
#pragma omp parallel for schedule(static)
      for (int j = 0; j < nb - dd; j++) {
        int i = j + dd;
        double T[9];
        for (int u = 0; u < 3; u++)
          for (int v = 0; v < 3; v++) {
            double s = 0.0;
            for (int k = 3 * j; k < 3 * i; k++) s += L[(long)(3 * i + u) * n + k] * X[(long)k * n + 3 * j + v];
            T[3 * u + v] = s;
          }
        for (int u = 0; u < 3; u++)
          for (int v = 0; v < 3; v++) {
            double s = 0.0;
            for (int w = 0; w < 3; w++) s += dinv[9 * i + 3 * u + w] * T[3 * w + v];
            X[(long)(3 * i + u) * n + 3 * j + v] = -s;
          }
      }
    }
    /* S^-1 = X^T X (SPD_inv.cl:343-411) */
#pragma omp parallel for schedule(dynamic, 8)
    for (int u = 0; u < n; u++)
      for (int v = 0; v <= u; v++) {
        double s = 0;
        for (int k = u; k < n; k++) s += X[(long)k * n + u] * X[(long)k * n + v];
        Sinv[(long)u * n + v] = s;
        Sinv[(long)v * n + u] = s;
      }
  }
  free(dinv);
  free(X);
  return ok;
}

/* factor the reduced camera system and solve S dpa = ea into dp[0..n);
 * PSBA/PSBA/levmar.cpp:134-140 (SPDinv + matVec_mul) in the inverse mode,
 * Cholesky + two triangular solves otherwise.  Returns 1 on success. */
static int factor_solve(int n, int solve_mode, const double *S, const double *b, double *x, double *L,
                        double *Sinv, double *y) {
  /* the reference's explicit inverse works on 3x3 blocks (cnp = 6); other
   * orders use the Cholesky solve */
  if (solve_mode == ORC_SOLVE_INVERSE && n % 3 == 0) {
    int ok = spd_inverse_ref(n, S, L, Sinv);
    if (!ok) return 0;
    for (int r = 0; r < n; r++) {
      double s = 0;
      for (int k = 0; k < n; k++) s = s + Sinv[(long)r * n + k] * b[k];
      x[r] = s;
    }
    return 1;
  }
  if (!chol_plain(n, S, L)) return 0;
  for (int i = 0; i < n; i++) {
    double s = b[i];
    const double *Li = L + (long)i * n;
    for (int k = 0; k < i; k++) s -= Li[k] * y[k];
    y[i] = s / Li[i];
  }
  for (int i = n - 1; i >= 0; i--) {
    double s = y[i];
    for (int k = i + 1; k < n; k++) s -= L[(long)k * n + i] * x[k];
    x[i] = s / L[(long)i * n + i];
  }
  return 1;
}

int orc_stage_factor_solve(orc_ctx *c, int solve_mode) {
  c->L_valid = 0;
  int ok = factor_solve(c->n, solve_mode, c->S, c->ea, c->dp, c->L, c->Sinv, c->xtmp);
  c->L_valid = ok;
  return ok;
}

int orc_dense_factor_solve(int n, int solve_mode, const double *S, const double *b, double *x, double *L) {
  double *Lw = L ? L : (double *)xcalloc((size_t)n * n, sizeof(double));
  double *Sinv = (double *)xcalloc((size_t)n * n, sizeof(double));
  double *y = (double *)xcalloc(n, sizeof(double));
  int ok = factor_solve(n, solve_mode, S, b, x, Lw, Sinv, y);
  if (!L) free(Lw);
  free(Sinv);
  free(y);
  return ok;
}

/* eb_i = g_b,i - sum_j W_ij^T dpa_j (PSBA/CL_files/compute_eb.cl:6-41) and
 * dpb_i = V_i^-1 eb_i (PSBA/CL_files/compute_dpb.cl:6-35) */
void orc_stage_backsub(orc_ctx *c) {
  int C = c->cnp;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < c->np; i++) {
    double eb[3];
    for (int cc = 0; cc < 3; cc++) {
      double sum = 0;
      for (int o0 = c->pt_ptr[i]; o0 < c->pt_ptr[i + 1]; o0++) {
        const int o = RIDX(c->pt_ptr[i], c->pt_ptr[i + 1], o0);
        const double *w = c->W + (long)3 * C * o;
        const double *da = c->dp + C * c->jidx[o];
        for (int k = 0; k < C; k++) sum = sum + w[3 * k + cc] * da[k];
      }
      eb[cc] = c->g[c->n + 3 * i + cc] - sum;
      c->eb[3 * i + cc] = eb[cc];
    }
    const double *vi = c->Vinv + 9L * i;
    for (int r = 0; r < 3; r++) c->dp[c->n + 3 * i + r] = vi[3 * r] * eb[0] + vi[3 * r + 1] * eb[1] + vi[3 * r + 2] * eb[2];
  }
}

/* new_p = p + dp, PSBA/CL_files/compute_newp.cl:6-26 */
void orc_stage_newp(orc_ctx *c, const double *dp) {
  for (int q = 0; q < c->n; q++) c->newcams[q] = c->cams[q] + dp[q];
  for (int q = 0; q < 3 * c->np; q++) c->newpts[q] = c->pts[q] + dp[c->n + q];
}

/* p = new_p, PSBA/CL_files/update_p.cl:6-25 */
void orc_stage_commit(orc_ctx *c) {
  memcpy(c->cams, c->newcams, sizeof(double) * c->n);
  memcpy(c->pts, c->newpts, sizeof(double) * 3 * c->np);
}

/* J x over the observations (the non-zero rows of the dense grid of
 * PSBA/CL_files/compute_Jmultiply.cl:6-52, in the same order) */
void orc_jmul(const orc_ctx *c, const double *x, double *out) {
  int C = c->cnp;
#pragma omp parallel for schedule(static)
  for (int o = 0; o < c->no; o++) {
    const double *ja = c->JA + (long)2 * C * o, *jb = c->JB + 6L * o;
    const double *xa = x + C * c->jidx[o], *xb = x + c->n + 3 * c->iidx[o];
    for (int r = 0; r < 2; r++) {
      double s = 0;
      for (int k = 0; k < C; k++) s += ja[r * C + k] * xa[k];
      for (int k = 0; k < 3; k++) s += jb[r * 3 + k] * xb[k];
      out[2 * o + r] = s;
    }
  }
}

/*
 * Modified Cholesky of S and lambda = |sum E| / n, the dogleg's recovery when S
 * is not positive definite (PSBA/PSBA/trust_region.cpp:343-366,
 * PSBA/PSBA/cl_cholmod.cpp:25-202, PSBA/CL_files/cholmod_blk.cl:87-846).
 * delta/beta as cl_cholmod.cpp:161-164.  Per 3-column block: try a plain
 * block step (positive closed-form pivots and every off-diagonal-block entry
 * of L <= beta, signed test as cholmod_blk.cl:340); otherwise restore the
 * block column and process its three columns one by one (Gill-Murray-Wright
 * style: d_j = max(|c_jj|, delta); if any L_ij > beta then sqrt(d_j) :=
 * theta_j / beta, cholmod_blk.cl:446-697).  E_i = (L L^T)_ii - S_ii.
 * S (lower triangle) is read from c->S; the factor goes to c->L.
 */
/* one column of the column-wise modified Cholesky, cholmod_blk.cl:446-697 */
static void modchol_column(int n, const double *A, double *L, double *Cc, int xj, double delta, double beta) {
  double sum = A[(long)xj * n + xj];
  for (int k = 0; k < xj; k++) sum -= L[(long)xj * n + k] * L[(long)xj * n + k];
  double dj = fmax(fabs(sum), delta);
  double sd = sqrt(dj);
  L[(long)xj * n + xj] = sd;
  int over = 0;
  for (int i = xj + 1; i < n; i++) {
    double s = A[(long)i * n + xj];
    for (int k = 0; k < xj; k++) s = s - L[(long)i * n + k] * L[(long)xj * n + k];
    Cc[i] = s;
    L[(long)i * n + xj] = s / sd;
    if (L[(long)i * n + xj] > beta) over = 1;
  }
  if (over) {
    double theta = 0;
    for (int i = xj + 1; i < n; i++) theta = fmax(theta, fabs(Cc[i]));
    double nd = theta / beta;
    L[(long)xj * n + xj] = nd;
    for (int i = xj + 1; i < n; i++) L[(long)i * n + xj] = Cc[i] / nd;
  }
}

static double modchol_dense(int n, const double *A, double *L, double *Eout, double *Cc) {
  int nb = n / 3;
  double xi = 0, gamma = 0;
  for (int r = 0; r < n; r++) {
    for (int k = 0; k < n; k++) {
      if (k == r) continue;
      double t = fabs(A[(long)r * n + k]);
      if (t > xi) xi = t;
    }
    double t = fabs(A[(long)r * n + r]);
    if (t > gamma) gamma = t;
  }
  double delta = 1e-15 * fmax(xi + gamma, 1);
  double beta = fmax(gamma, 1e-15);
  beta = fmax(beta, xi / sqrt((double)(n * n - 1)));
  beta = sqrt(beta);

  memset(L, 0, sizeof(double) * (size_t)n * n);
  for (int j = 0; j < nb; j++) {
    int blk_ok = 1;
    double T[9];
    for (int u = 0; u < 3; u++)
      for (int v = 0; v < 3; v++) {
        double s = A[(long)(3 * j + u) * n + 3 * j + v];
        for (int k = 0; k < 3 * j; k++) s -= L[(long)(3 * j + u) * n + k] * L[(long)(3 * j + v) * n + k];
        T[3 * u + v] = s;
      }
    double l0 = T[0];
    if (!isfinite(l0) || l0 <= 0) blk_ok = 0;
    l0 = sqrt(l0);
    double l1 = T[3] / sqrt(T[0]);
    if (!isfinite(l1)) blk_ok = 0;
    double l2 = T[4] - T[3] * T[3] / T[0];
    if (!isfinite(l2) || l2 <= 0) blk_ok = 0;
    l2 = sqrt(l2);
    double l3 = T[6] / sqrt(T[0]);
    if (!isfinite(l3)) blk_ok = 0;
    double l4 = sqrt(T[0] / (T[0] * T[4] - T[3] * T[3])) * (T[7] - T[3] * T[6] / T[0]);
    if (!isfinite(l4)) blk_ok = 0;
    double t1 = -T[8] * T[3] * T[3], t2 = 2 * T[7] * T[3] * T[6], t3 = -T[4] * T[6] * T[6];
    double t4 = T[0] * (T[4] * T[8] - T[7] * T[7]), t5 = -T[3] * T[3] + T[0] * T[4];
    double l5 = (t1 + t2 + t3 + t4) / t5;
    if (!isfinite(l5) || l5 <= 0) blk_ok = 0;
    l5 = sqrt(l5);
    if (blk_ok) {
      double di[9] = {1 / l0, 0, 0, -l1 / (l0 * l2), 1 / l2, 0,
                      (l1 * l4 - l2 * l3) / (l0 * l2 * l5), -l4 / (l2 * l5), 1 / l5};
      long b = (long)(3 * j) * n + 3 * j;
      L[b] = l0; L[b + n] = l1; L[b + n + 1] = l2; L[b + 2 * n] = l3; L[b + 2 * n + 1] = l4; L[b + 2 * n + 2] = l5;
      int over = 0;
      /* row triplets below the diagonal block; the last one is partial when
       * n is not a multiple of 3 */
      for (int i = j + 1; 3 * i < n; i++) {
        double Tij[9];
        int nu = n - 3 * i < 3 ? n - 3 * i : 3;
        for (int u = 0; u < nu; u++)
          for (int v = 0; v < 3; v++) {
            double s = A[(long)(3 * i + u) * n + 3 * j + v];
            for (int k = 0; k < 3 * j; k++) s -= L[(long)(3 * i + u) * n + k] * L[(long)(3 * j + v) * n + k];
            Tij[3 * u + v] = s;
          }
        for (int u = 0; u < nu; u++)
          for (int v = 0; v < 3; v++) {
            double s = Tij[3 * u] * di[3 * v] + Tij[3 * u + 1] * di[3 * v + 1] + Tij[3 * u + 2] * di[3 * v + 2];
            L[(long)(3 * i + u) * n + 3 * j + v] = s;
            if (s > beta) over = 1;
          }
      }
      if (!over) continue;
      /* restore block column j and fall back to the column-wise path */
      for (int i = 3 * j; i < n; i++)
        for (int v = 0; v < 3; v++) L[(long)i * n + 3 * j + v] = 0.0;
    } else {
      for (int i = 3 * j; i < 3 * j + 3; i++)
        for (int v = 0; v < 3; v++) L[(long)i * n + 3 * j + v] = 0.0;
    }
    for (int col = 0; col < 3; col++) modchol_column(n, A, L, Cc, 3 * j + col, delta, beta);
  }
  /* columns past the last full 3-block (n not a multiple of 3, possible with
   * cnp 11 or 16; the reference's cnp = 6 never has them): column-wise path */
  for (int xj = 3 * nb; xj < n; xj++) modchol_column(n, A, L, Cc, xj, delta, beta);
  double sumE = 0;
  for (int i = 0; i < n; i++) {
    double s = 0;
    for (int k = 0; k <= i; k++) s += L[(long)i * n + k] * L[(long)i * n + k];
    double e = s - A[(long)i * n + i];
    if (Eout) Eout[i] = e;
    sumE += e;
  }
  return fabs(sumE) / n;
}

double orc_modchol_lambda(orc_ctx *c, double *Eout) {
  c->L_valid = 1;
  return modchol_dense(c->n, c->S, c->L, Eout, c->xtmp);
}

double orc_dense_modchol(int n, const double *S, double *L, double *E) {
  double *Lw = L ? L : (double *)xcalloc((size_t)n * n, sizeof(double));
  double *Cc = (double *)xcalloc(n, sizeof(double));
  double lam = modchol_dense(n, S, Lw, E, Cc);
  if (!L) free(Lw);
  free(Cc);
  return lam;
}

long orc_buffer_size(const orc_ctx *c, int which) {
  int C = c->cnp;
  switch (which) {
    case ORC_BUF_EX: return 2L * c->no;
    case ORC_BUF_JA: return 2L * C * c->no;
    case ORC_BUF_JB: return 6L * c->no;
    case ORC_BUF_U: return (long)C * C * c->nc;
    case ORC_BUF_V: return 9L * c->np;
    case ORC_BUF_W: return 3L * C * c->no;
    case ORC_BUF_G: return c->ntot;
    case ORC_BUF_VINV: return 9L * c->np;
    case ORC_BUF_Y: return 3L * C * c->no;
    case ORC_BUF_S: return (long)c->n * c->n;
    case ORC_BUF_EA: return c->n;
    case ORC_BUF_DP: return c->ntot;
    case ORC_BUF_EB: return 3L * c->np;
    case ORC_BUF_L: return (long)c->n * c->n;
    case ORC_BUF_NEWCAMS: return c->n;
    case ORC_BUF_NEWPTS: return 3L * c->np;
  }
  return -1;
}

int orc_read_buffer(const orc_ctx *c, int which, double *out) {
  const double *src = NULL;
  switch (which) {
    case ORC_BUF_EX: src = c->ex; break;
    case ORC_BUF_JA: src = c->JA; break;
    case ORC_BUF_JB: src = c->JB; break;
    case ORC_BUF_U: src = c->U; break;
    case ORC_BUF_V: src = c->V; break;
    case ORC_BUF_W: src = c->W; break;
    case ORC_BUF_G: src = c->g; break;
    case ORC_BUF_VINV: src = c->Vinv; break;
    case ORC_BUF_Y: src = c->Y; break;
    case ORC_BUF_S: src = c->S; break;
    case ORC_BUF_EA: src = c->ea; break;
    case ORC_BUF_DP: src = c->dp; break;
    case ORC_BUF_EB: src = c->eb; break;
    case ORC_BUF_L: src = c->L; break;
    case ORC_BUF_NEWCAMS: src = c->newcams; break;
    case ORC_BUF_NEWPTS: src = c->newpts; break;
    default: return -1;
  }
  memcpy(out, src, sizeof(double) * (size_t)orc_buffer_size(c, which));
  return 0;
}

/* ------------------------------------------------------------------------ */
/* drivers                                                                  */
/* ------------------------------------------------------------------------ */

static void trace_push(orc_trace_rec *tr, int cap, int *len, const orc_trace_rec *r) {
  if (!tr || !len) return;
  if (*len < cap) tr[*len] = *r;
  (*len)++;
}

static void opts_default(const orc_opts *o, orc_opts *d) {
  d->solve_mode = ORC_SOLVE_CHOLESKY;
  d->max_iter = 50;
  d->max_pb_tries = 64;
  d->flags = 0;
  d->rel_tol = 0.0;
  if (o) {
    d->solve_mode = o->solve_mode;
    if (o->max_iter > 0) d->max_iter = o->max_iter;
    if (o->max_pb_tries > 0) d->max_pb_tries = o->max_pb_tries;
    d->flags = o->flags;
    d->rel_tol = o->rel_tol;
  }
}

static double params_l2_full(const orc_ctx *c) {
  /* PSBA/PSBA/levmar.cpp:211-212: compute_L2_sq over [cams ; pts] in order */
  double s = 0;
  for (int q = 0; q < c->n; q++) s += c->cams[q] * c->cams[q];
  for (int q = 0; q < 3 * c->np; q++) s += c->pts[q] * c->pts[q];
  return s;
}

/* PSBA/PSBA/levmar.cpp:271-280 */
static double compute_rho(const orc_ctx *c, double ex_L2, double new_ex_L2, double mu) {
  double sum = 0;
  for (int q0 = 0; q0 < c->ntot; q0++) {
    const int q = RIDX(0, c->ntot, q0);
    sum += c->dp[q] * (mu * c->dp[q] + c->g[q]);
  }
  return (ex_L2 - new_ex_L2) / sum;
}

/* PSBA/PSBA/levmar.cpp:45-256 */
int orc_levmar(orc_ctx *c, const orc_opts *oin, int *itno, double *finalErr,
               orc_trace_rec *trace, int trace_cap, int *trace_len) {
  orc_opts o;
  opts_default(oin, &o);
  double tau = PSBA_INIT_MU, mu = 0, p_L2 = 0, rho;
  int gooditer_cnt = 0, nu = 2, first = 1;
  double ex_L2 = orc_stage_residual(c, 0, NULL);
  int iter_flag = ORC_ITER_CONTINUE;
  for (; *itno < o.max_iter && iter_flag == ORC_ITER_CONTINUE; (*itno)++) {
    orc_stage_jacobian(c);
    orc_stage_normal(c, 1.0, 1.0);
    if (first) {
      mu = tau * orc_max_uvdiag(c);
      first = 0;
      p_L2 = 1e+3;
      nu = 2;
    }
    while (1) {
      orc_trace_rec rec = {*itno, 0, 0, 0, NAN, mu, NAN, 0.0, NAN};
      orc_stage_vinv(c, mu);
      orc_stage_schur(c, mu);
      orc_stage_ea(c);
      int ok = orc_stage_factor_solve(c, o.solve_mode);
      rec.factor_ok = ok;
      if (ok) {
        orc_stage_backsub(c);
        double dp_L2 = l2sq(c->ntot, c->dp);
        rec.dp_sq = dp_L2;
        if (dp_L2 < p_L2 * PSBA_STOP_THRESH * PSBA_STOP_THRESH) {
          iter_flag = ORC_ITER_DP_NO_CHANGE;
          trace_push(trace, trace_cap, trace_len, &rec);
          break;
        }
        if (dp_L2 >= (p_L2 + PSBA_STOP_THRESH) / PSBA_EPSILON_SQ) {
          iter_flag = ORC_ITER_ERR;
          trace_push(trace, trace_cap, trace_len, &rec);
          break;
        }
        orc_stage_newp(c, c->dp);
        double new_ex_L2 = orc_stage_residual(c, 1, NULL);
        rho = compute_rho(c, ex_L2, new_ex_L2, mu);
        rec.cost = new_ex_L2;
        rec.rho = rho;
        if (rho > 0) {
          double tmp = 2 * rho - 1;
          tmp = 1.0 - tmp * tmp * tmp;
          mu = mu * ((tmp >= (1.0 / 3.0)) ? tmp : (1.0 / 3.0));
          nu = 2;
          orc_stage_commit(c);
          p_L2 = params_l2_full(c);
          double rel_dec = (ex_L2 - new_ex_L2) / ex_L2;
          ex_L2 = new_ex_L2;
          rec.accepted = 1;
          trace_push(trace, trace_cap, trace_len, &rec);
          if (o.rel_tol > 0 && rel_dec < o.rel_tol) {
            iter_flag = ORC_ITER_DP_NO_CHANGE;
            break;
          }
          if (o.flags & ORC_FLAG_NO_HANDOFF) {
            gooditer_cnt = 0;
          } else if (fabs(rho - 1) < (1.0 / 5.0)) {
            gooditer_cnt++;
            if (gooditer_cnt >= 5) {
              iter_flag = ORC_ITER_TURN_TO_TR;
              break;
            }
          } else {
            gooditer_cnt = 0;
          }
          break;
        }
      } else {
        gooditer_cnt = 0;
      }
      trace_push(trace, trace_cap, trace_len, &rec);
      mu *= nu;
      /* levmar.cpp:238-243: int nu doubles until 2*nu wraps (nu = 2^30) */
      if (nu >= (1 << 30)) {
        iter_flag = ORC_ITER_ERR;
        break;
      }
      nu = 2 * nu;
    }
    if (ex_L2 <= PSBA_STOP_THRESH) iter_flag = ORC_ITER_ERR_SMALL_ENOUGH;
  }
  *finalErr = ex_L2;
  return iter_flag;
}

/* PSBA/PSBA/trust_region.cpp:292-405; returns 1 when P_B was produced */
static int compute_PB(orc_ctx *c, const orc_opts *o, double *lambda, double *P_B) {
  orc_stage_vinv(c, *lambda);
  orc_stage_schur(c, *lambda);
  memcpy(c->Saux, c->S, sizeof(double) * (size_t)c->n * c->n);
  /* ea is formed before the factorisation here; the reference forms it after
   * (trust_region.cpp:378) but it does not depend on the factor */
  orc_stage_ea(c);
  int ok = orc_stage_factor_solve(c, o->solve_mode);
  if (!ok) {
    if (*lambda == 0.0) {
      memcpy(c->S, c->Saux, sizeof(double) * (size_t)c->n * c->n);
      *lambda = orc_modchol_lambda(c, NULL);
      return 0;
    }
    *lambda = 2 * (*lambda);
    return 0;
  }
  orc_stage_backsub(c);
  for (int q = 0; q < c->ntot; q++) P_B[q] = -c->dp[q];
  return 1;
}

/* PSBA/PSBA/trust_region.cpp:520-595 (p_norm bookkeeping of :564-569 kept) */
static double compute_p_2(int N, double pUtBpU, double pUtBpB, double pBtBpB,
                          double delta, const double *P_U, const double *P_B, double *p,
                          const double *g) {
  double pUg = dotp(P_U, g, N), pBg = dotp(P_B, g, N);
  double pUpU = dotp(P_U, P_U, N), pUpB = dotp(P_U, P_B, N), pBpB = dotp(P_B, P_B, N);
  (void)pUpU; (void)pUpB; (void)pBpB;
  double den = -pUtBpB * pUtBpB + pBtBpB * pUtBpU;
  double eta1 = (pBg * pUtBpB) / den - (pBtBpB * pUg) / den;
  double eta2 = (pUg * pUtBpB) / den - (pBg * pUtBpU) / den;
  double p_norm = 0.0;
  for (int q = 0; q < N; q++) {
    p[q] = eta1 * P_U[q] + eta2 * P_B[q];
    p_norm += p[q] * p[q];
  }
  p_norm = sqrt(p_norm);
  if (p_norm > delta) {
    double pU_norm = 0.0, pB_norm = 0.0;
    for (int q = 0; q < N; q++) {
      pU_norm += P_U[q] * P_U[q];
      pB_norm += P_B[q] * P_B[q];
    }
    pU_norm = sqrt(pU_norm);
    pB_norm = sqrt(pB_norm);
    if (pU_norm > delta) {
      for (int q = 0; q < N; q++) p[q] = delta * P_U[q] / pU_norm;
      return delta;
    } else if (pB_norm <= delta) {
      for (int q = 0; q < N; q++) {
        p[q] = P_B[q];
        p_norm += p[q] * p[q];
      }
      return sqrt(p_norm);
    } else {
      double a = 0.0, b = 0.0, cc = 0.0;
      for (int q = 0; q < N; q++) {
        double Ai = P_B[q] - P_U[q], Bi = 2 * P_U[q] - P_B[q];
        a += Ai * Ai;
        b += Ai * Bi;
        cc += Bi * Bi;
      }
      b = 2 * b;
      cc = cc - delta * delta;
      double disc = b * b - 4 * a * cc;
      if (fabs(disc) < 1e-12) disc = 0;
      double tau = (-b + sqrt(disc)) / (2 * a);
      for (int q = 0; q < N; q++) p[q] = P_U[q] + (tau - 1) * (P_B[q] - P_U[q]);
      return delta;
    }
  }
  return p_norm;
}

double orc_dogleg_step(int N, double pUtBpU, double pUtBpB, double pBtBpB, double delta,
                       const double *P_U, const double *P_B, const double *g, double *p) {
  return compute_p_2(N, pUtBpU, pUtBpB, pBtBpB, delta, P_U, P_B, p, g);
}

/* PSBA/PSBA/trust_region.cpp:49-288 */
int orc_trust_region(orc_ctx *c, const orc_opts *oin, int *itno, double *finalErr,
                     orc_trace_rec *trace, int trace_cap, int *trace_len) {
  orc_opts o;
  opts_default(oin, &o);
  int N = c->ntot;
  int notgood_cnt = 0, good_iters = 0, nu = 2;
  double dk = 1, lambda = 0, origin_lambda = 0.0;
  double *Jx1 = (double *)xcalloc((size_t)2 * c->no, sizeof(double));
  double *Jx2 = (double *)xcalloc((size_t)2 * c->no, sizeof(double));
  double *P_U = (double *)xcalloc(N, sizeof(double));
  double *P_B = (double *)xcalloc(N, sizeof(double));
  double *P = (double *)xcalloc(N, sizeof(double));
  double *g = (double *)xcalloc(N, sizeof(double));
  double ex_L2 = orc_stage_residual(c, 0, NULL);
  int iter_flag = ORC_ITER_CONTINUE;
  int ret_early = 0;
  for (; *itno < o.max_iter; (*itno)++) {
    orc_stage_jacobian(c);
    orc_stage_normal(c, 2.0, -2.0);
    memcpy(g, c->g, sizeof(double) * N);
    orc_jmul(c, g, Jx1);
    double gtBg = 2 * l2sq(2L * c->no, Jx1);
    double gtg = dotp(g, g, N);
    for (int q = 0; q < N; q++) P_U[q] = -(g[q] * gtg) / gtBg;
    int solved = 0, tries = 0;
    while (!solved) {
      solved = compute_PB(c, &o, &lambda, P_B);
      if (!solved) {
        if (origin_lambda != 0.0) {
          if (nu > 4) {
            iter_flag = ORC_ITER_TURN_TO_LM;
            *finalErr = ex_L2;
            ret_early = 1;
            break;
          }
          lambda = lambda * nu;
          nu = nu * 2;
        }
        if (++tries >= o.max_pb_tries) {
          iter_flag = ORC_ITER_TURN_TO_LM;
          *finalErr = ex_L2;
          ret_early = 1;
          break;
        }
      } else {
        nu = 2;
        origin_lambda = lambda;
      }
    }
    if (ret_early) break;
    orc_jmul(c, P_U, Jx1);
    orc_jmul(c, P_B, Jx2);
    double pUtBpU = 2 * dotp(Jx1, Jx1, 2L * c->no);
    double pUtBpB = 2 * dotp(Jx1, Jx2, 2L * c->no);
    double pBtBpB = 2 * dotp(Jx2, Jx2, 2L * c->no);
    iter_flag = ORC_ITER_CONTINUE;
    while (iter_flag == ORC_ITER_CONTINUE) {
      orc_trace_rec rec = {*itno, 1, 1, 0, NAN, dk, NAN, lambda, NAN};
      double p_norm = compute_p_2(c->ntot, pUtBpU, pUtBpB, pBtBpB, dk, P_U, P_B, P, g);
      rec.dp_sq = p_norm * p_norm;
      orc_stage_newp(c, P);
      double act_ex_L2 = orc_stage_residual(c, 1, NULL);
      rec.cost = act_ex_L2;
      if (fabs((ex_L2 - act_ex_L2) / ex_L2) < PSBA_EPSILON2) {
        iter_flag = ORC_ITER_DP_NO_CHANGE;
        trace_push(trace, trace_cap, trace_len, &rec);
        break;
      }
      orc_jmul(c, P, Jx1);
      double Jx_norm = 2 * l2sq(2L * c->no, Jx1);
      double pred_ex_L2 = dotp(g, P, N);
      pred_ex_L2 += ex_L2 + Jx_norm / 2;
      double rho = (ex_L2 - act_ex_L2) / (ex_L2 - pred_ex_L2);
      rec.rho = rho;
      if (rho < (1.0 / 4.0) || act_ex_L2 > ex_L2) {
        dk = dk / 4;
      } else if (rho >= (3.0 / 4.0) && act_ex_L2 < ex_L2) {
        iter_flag = ORC_ITER_PASS;
        orc_stage_commit(c);
        *finalErr = act_ex_L2;
        dk = fmin(2 * dk, MAX_DELTA);
        rec.accepted = 1;
      } else if (rho >= (1.0 / 4.0) && rho < (3.0 / 4.0) && act_ex_L2 < ex_L2) {
        iter_flag = ORC_ITER_PASS;
        orc_stage_commit(c);
        *finalErr = act_ex_L2;
        rec.accepted = 1;
      } else if (isnan(rho)) {
        iter_flag = ORC_ITER_TURN_TO_LM;
        *finalErr = ex_L2;
        trace_push(trace, trace_cap, trace_len, &rec);
        ret_early = 1;
        break;
      }
      trace_push(trace, trace_cap, trace_len, &rec);
      if (fabs((act_ex_L2 - ex_L2) / ex_L2) <= PSBA_EPSILON2) {
        iter_flag = ORC_ITER_ERR_SMALL_ENOUGH;
        break;
      }
      if (rho < 1.0 / 4) {
        notgood_cnt++;
        if (notgood_cnt >= 5) {
          iter_flag = ORC_ITER_TURN_TO_LM;
          break;
        }
      } else {
        notgood_cnt = 0;
      }
      if (rho > 3.0 / 4 && act_ex_L2 < ex_L2) {
        good_iters++;
        if (good_iters >= 10) {
          lambda = 0.0;
          origin_lambda = 0.0;
          good_iters = 0;
        }
      } else {
        good_iters = 0;
      }
      if (rho > (1.0 / 4) && act_ex_L2 < ex_L2) ex_L2 = act_ex_L2;
    }
    if (ret_early) break;
    if (iter_flag != ORC_ITER_PASS) break;
  }
  free(Jx1); free(Jx2); free(P_U); free(P_B); free(P); free(g);
  return iter_flag;
}

/* PSBA/PSBA/main.cpp:193-208 */
int orc_hybrid(orc_ctx *c, const orc_opts *o, int *itno, double *initErr, double *finalErr,
               orc_trace_rec *trace, int trace_cap, int *trace_len) {
  if (initErr) *initErr = orc_stage_residual(c, 0, NULL);
  int flag;
  while (1) {
    flag = orc_levmar(c, o, itno, finalErr, trace, trace_cap, trace_len);
    if (flag != ORC_ITER_TURN_TO_TR) break;
    flag = orc_trust_region(c, o, itno, finalErr, trace, trace_cap, trace_len);
    if (flag != ORC_ITER_TURN_TO_LM) break;
  }
  return flag;
}
