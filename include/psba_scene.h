/*
 * psba_scene.h -- seeded synthetic bundle-adjustment scenes (C ABI).
 *
 * Deterministic across machines: xoshiro256** seeded by splitmix64, Box-Muller
 * normals.  Observations are point-major and camera-ascending, i.e. already in
 * the order generate_idxs produces.  Camera 0 is held fixed (fixmask).
 *
 * config 0: 7 cameras on a circle (r=10) looking at the origin, pose-only
 *           (cnp 6), fx=fy=800, cx=320, cy=240; 1000 points in [-2,2]^3;
 *           observations inside 640x480; sigma 0.5 px; init: 0.02 rad
 *           rotation, 0.1 translation, N(0, 0.05^2) points.
 * config 1: 54 cameras on an arc, cnp 11 (or 16 with distortion); 5200 points;
 *           geometric track lengths (p=0.26) truncated to [2,20]; sigma 1 px;
 *           config-0 perturbations plus 2% on the intrinsics.
 * config 2: 88 cameras along a polyline, cnp 9 (f, k1, k2 + pose); 64298
 *           points in a 100x100x20 slab; Zipf track lengths on [2,88] with the
 *           exponent fitted to 384000 observations (zipf_s overrides);
 *           sigma 1 px and 1% uniform outliers over 1024x768.
 * config 3: one window of the batched workload: cameras U[5,30], points
 *           U[200,4000], tracks U[2,min(8,cams)], sigma 0.5 px, cnp 6; the
 *           problem seed is splitmix64(seed ^ problem).
 * config 4: one request of the serving stream, size_class 0 (10 cams, 1000
 *           pts, ~6k obs), 1 (50 cams, 20000 pts, ~100k obs), 2 (138 cams,
 *           19878 pts, Zipf tracks, ~100k obs); cnp 6.
 */
#ifndef PSBA_SCENE_H
#define PSBA_SCENE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int cnp, n_cams, n_pts, n_obs;
  double* K;      /* 5*n_cams */
  double* q0;     /* 4*n_cams initial rotation (perturbed) */
  double* cams;   /* cnp*n_cams initial parameters (v = 0) */
  double* pts;    /* 3*n_pts initial points */
  double* meas;   /* 2*n_obs */
  int* iidx;      /* n_obs point of each observation */
  int* jidx;      /* n_obs camera of each observation */
  unsigned char* fixmask; /* cnp*n_cams */
  double* gt_q0;  /* 4*n_cams true rotations */
  double* gt_cams;/* cnp*n_cams true parameters (v = 0) */
  double* gt_pts; /* 3*n_pts */
  double sum_t2;  /* sum of squared track lengths */
  int n_outliers;
  double zipf_s;
} psba_scene;

typedef struct {
  int cnp;             /* 0 = default of the config (config 1: 11 or 16) */
  int n_cams, n_pts;   /* 0 = default */
  double zipf_s;       /* 0 = default */
  double noise;        /* < 0 = default; 0 = exact measurements */
  double outlier_frac; /* < 0 = default */
  double perturb;      /* scale of the initial perturbation; < 0 = 1 */
  int size_class;      /* config 4 */
  unsigned long long problem; /* config 3/4: index of the problem in the stream */
} psba_scene_opts;

/* opts may be NULL (all defaults) */
psba_scene* psba_scene_generate(int config, unsigned long long seed, const psba_scene_opts* opts);
void psba_scene_free(psba_scene* s);
/* splitmix64(seed ^ k): the per-problem seed of configs 3 and 4 */
unsigned long long psba_problem_seed(unsigned long long seed, unsigned long long k);

#ifdef __cplusplus
}
#endif
#endif
