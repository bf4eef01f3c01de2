// Device state of one bundle-adjustment problem (the HIP-backed PSBA_struct).
//
// HBM layout (all FP64 in the plain path):
//   parameters   cams[cnp*nc], pts[3*np]                 (AoS per camera / point)
//   observations meas[2*no], obs_cam[no], obs_pt[no]     (point-major order)
//   per-obs      e[2*no], JA[2*cnp*no], JB[6*no], W[3*cnp*no]  (AoS per obs)
//   per-camera   U[cnp*cnp*nc] (full, symmetric), g[0..n)
//   per-point    V[6*np] (xx xy xz yy yz zz), Vinv[6*np], z[3*np] = V^-1 g_b
//   reduced      S[n_pad*n_pad] row-major, lower triangle; padded rows/cols
//                beyond n = cnp*nc hold the identity so every tile is full.
//   index        point CSR (pt_ptr), camera CSR (cam_ptr, cam_obs ascending
//                point), camera-pair item lists (obs of camera k, obs of camera
//                l) for k >= l in ascending point order, split into fixed
//                chunks for the deterministic two-stage reductions.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "common.hpp"

namespace psba {

constexpr int kCamChunk = 256;   // observations per camera chunk (U/g_a/e_a)
constexpr int kPairChunk = 64;   // shared points per camera-pair chunk (S)
constexpr int kNB = 64;          // dense tile size of the Cholesky / solves
constexpr int kMaxSolveN = 8192; // largest reduced system the single-block solve holds in LDS

inline int cam_chunk_size(int cnp) { return cnp > 11 ? 128 : kCamChunk; }

// scalar slots in ctx->d_scal (device) / ctx->h_scal (pinned host)
enum Scalar {
  SC_COST = 0,     // sum of squared residuals at the evaluated parameters
  SC_DP2,          // |dp|^2
  SC_RHODEN,       // sum dp (mu dp + g)
  SC_P2,           // |p|^2 after commit
  SC_MAXDIAG,      // max(UVdiag)
  SC_GG,           // g.g
  SC_D0, SC_D1, SC_D2, SC_D3, SC_D4, SC_D5,  // generic dot products
  SC_LAMBDA,       // modified-Cholesky lambda
  SC_FACTOR_OK,    // 1.0 when the Cholesky succeeded
  SC_COUNT = 32
};


// Device-held LM state of one problem (k_lm.hip) and the gates it sets for
// the kernels of the next trial.
struct LmDev {
  double mu, ex_L2, p_L2, rel_tol, entry_cost;
  int nu, gooditer, flag, itno, max_iter, first, p_pending, no_handoff;
  int ntrace, trace_cap, trials, pad_;
};
enum { kGateActive = 0, kGateJac = 1, kGateAccept = 2, kGateCount = 4 };

struct Timing {
  double jac = 0, schur = 0, factor = 0, solve = 0, backsub = 0, resid = 0, other = 0;
  int trials = 0, jacobians = 0;
};

}  // namespace psba

struct psba_ctx {
  int cnp = 0, nc = 0, np = 0, no = 0;
  int n = 0, ntot = 0, n_pad = 0;
  int precision = 0, device = 0;
  int loss = PSBA_LOSS_NONE;  // robust loss (PSBA_LOSS_*) and its scale
  double loss_scale = 1.0;
  bool uploaded = false, indexed = false;
  bool s_loaded = false;  // a reduced system was loaded directly (psba_stage_load_S)
  hipStream_t stream = nullptr;

  // problem data
  double *K = nullptr, *q0 = nullptr, *meas = nullptr, *cams = nullptr, *pts = nullptr;
  int *obs_cam = nullptr, *obs_pt = nullptr;
  unsigned char* fix = nullptr;
  std::vector<unsigned char> h_fix;

  // index
  int* pt_ptr = nullptr;
  int *cam_ptr = nullptr, *cam_obs = nullptr;
  int n_cchunk = 0;
  int *cchunk_begin = nullptr, *cchunk_end = nullptr, *cam_cchunk_ptr = nullptr;
  int npairs = 0;
  long n_items = 0;
  int2* pair_items = nullptr;  // (obs in camera k, obs in camera l)
  int *pair_k = nullptr, *pair_l = nullptr;
  int n_schunk = 0;
  int *schunk_begin = nullptr, *schunk_end = nullptr, *pair_schunk_ptr = nullptr;
  int n_lblocks = 0;            // nc*(nc+1)/2 lower block positions
  int* lblock_pair = nullptr;   // pair index of each lower block position or -1
  int* schunk_order = nullptr;  // launch order of the pair chunks (schunk_launch_key)
  // the pair chunks' items in launch order, kPairChunk records per launch
  // slot: (obs in camera k, obs in camera l, point, chunk), then (-1, -1, -1,
  // chunk) past the chunk's items; what k_schur_chunks reads first, at an
  // address its workgroup index gives (built by build_schur_slots)
  int4* slot_items = nullptr;
  // the camera chunks' observations and their points, cam_chunk_size(cnp)
  // records per chunk ((-1, -1) past the chunk's end), read at an address the
  // workgroup index gives (built by build_schur_slots)
  int2* cchunk_rec = nullptr;
  int* pt_order = nullptr;      // points by descending track length (ties: ascending id)
  double sum_t2 = 0;            // sum over points of track length squared
  double entry_cost = 0;        // SSE at the entry of the last psba_levmar call

  bool mixed() const { return precision == PSBA_PREC_MIXED; }

  // work buffers (plain path: FP64; the mixed path uses the *32 FP32 copies of
  // the per-observation arrays and chunk partials instead)
  double *e = nullptr, *JA = nullptr, *JB = nullptr, *W = nullptr;
  float *e32 = nullptr, *JA32 = nullptr, *JB32 = nullptr, *W32 = nullptr;
  float *ucpart32 = nullptr, *schpart32 = nullptr, *eapart32 = nullptr;
  double *U = nullptr, *V = nullptr, *g = nullptr, *uvdiag = nullptr;
  double *Vinv = nullptr, *z = nullptr;
  double *ucpart = nullptr, *schpart = nullptr, *eapart = nullptr;
  double *S = nullptr, *ea = nullptr, *xsol = nullptr, *dp = nullptr, *eb = nullptr;
  double *vecA = nullptr, *vecB = nullptr, *vecC = nullptr, *vecD = nullptr;  // n_tot scratch
  double* mc_work = nullptr;   // modified Cholesky scratch (3*n_pad)
  // mixed path's FP64 refinement: per-observation products, camera-chunk partials
  double* ref_obs = nullptr;   // 2 (cnp + 3) n_obs
  double* ref_part = nullptr;  // 2 cnp per camera chunk
  long ref_part_cap = 0;
  int refine_steps = 0;        // sweeps per trial (psba_opts.refine_steps)
  // fused factor + solve (k_chol_solve): factor, right-hand side copy, y = L^-1 b,
  // bar = [grid-barrier counter (only grows), x flag, T row flags, epoch
  // (the counter's value before the next launch)]
  double *Lf = nullptr, *bw = nullptr, *yv = nullptr;
  unsigned long long* bar = nullptr;
  int* chol_fail = nullptr;
  int chol_grid = 0, chol_resident = 0;
  const double* L_cur = nullptr;  // where the last factor lives (Lf, or S after the modified Cholesky)
  double* Saux = nullptr;      // copy of S for the dogleg's recovery path (allocated on first use)
  double* snap = nullptr;      // parameter snapshot (psba_snapshot_params)
  double* red_part = nullptr;  // reduction partials
  long red_part_cap = 0;
  double* d_scal = nullptr;
  double* h_scal = nullptr;  // pinned

  // device-side LM control (psba_levmar without a host round trip per
  // trial): state, gates, trace records, a zero-filled owner array (the
  // batch kernels' per-element problem index, all 0 for one problem), pinned
  // status copies and their events.  While `devctl` is set the launchers
  // read the damping from lm->mu and pass the gates to their kernels.
  bool devctl = false;
  psba::LmDev* lm = nullptr;
  unsigned char* gate = nullptr;
  psba_trace_rec* lm_trace = nullptr;
  int lm_trace_cap = 0;
  psba::LmDev* h_lm = nullptr;  // pinned, two slots
  hipEvent_t ev_lm[2] = {nullptr, nullptr};

  // kernel probe (events around one kernel's launches)
  int probe_id = 0;
  std::vector<hipEvent_t> probe_pool;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> probe_pending;

  // timing
  bool timing_on = false;
  psba::Timing timing;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

namespace psba {

// probe helpers: record an event before/after a launch when probing `id`
void probe_begin(psba_ctx* c, int id);
void probe_end(psba_ctx* c, int id);

// ---- launchers (k_obs.hip) ----
int launch_residual(psba_ctx* c, bool trial);            // -> d_scal[SC_COST], e (current only)
int launch_jacobian_normal(psba_ctx* c, double coeff, double coeff_g);  // e, JA, JB, W, U, V, g, uvdiag
int launch_maxdiag(psba_ctx* c);                          // -> d_scal[SC_MAXDIAG]
int launch_vinv(psba_ctx* c, double mu);                  // Vinv, z
int launch_backsub(psba_ctx* c);                          // eb, dp[n..)
int launch_step_stats(psba_ctx* c, double mu);            // SC_DP2, SC_RHODEN
int launch_jacobian_normal_vinv(psba_ctx* c);  // device control, after the first trial: the stage + V^-1
int launch_trial_eval(psba_ctx* c);  // device control: step statistics + trial cost partials, one launch
int launch_commit(psba_ctx* c);                           // p += dp, SC_P2
// out[SC_D0..] = dots of (J x1).(J x1), (J x1).(J x2), (J x2).(J x2); x2 may be null
int launch_jmul_dots(psba_ctx* c, const double* x1, const double* x2);
// generic n_tot dot products: SC_D0 = a.b (+ more)
int launch_dot(psba_ctx* c, const double* a, const double* b, int slot);
int launch_axpby(psba_ctx* c, double* out, double alpha, const double* a, double beta,
                 const double* b, long n);

// ---- launchers (k_schur.hip) ----
int launch_schur(psba_ctx* c, double mu);                 // S (lower), ea
int build_point_order(psba_ctx* c);                        // pt_order (index builds)
int build_schur_slots(psba_ctx* c);                        // slot_items, cchunk_rec (index builds)
int launch_ea_rhs(psba_ctx* c, const double* ga, const double* z, double* out);  // ga - sum W z

// ---- launchers (k_dense.hip) ----
int launch_factor_solve(psba_ctx* c);                     // S -> Lf, SC_FACTOR_OK; dp[0..n) = S^-1 ea
int launch_modchol(psba_ctx* c, double* S_src);           // SC_LAMBDA (factor in S)
int launch_copy_S(psba_ctx* c, double* dst, const double* src);
int launch_tri_solve(psba_ctx* c, const double* b, double* out);  // S x = b with the factor in Lf

// ---- mixed-path refinement (k_obs.hip) ----
int launch_refine(psba_ctx* c, double mu, int sweeps);           // correction of dp -> vecB, g <- FP64 J^T e
int refine_step(psba_ctx* c, double mu, int sweeps);             // guarded sweeps (drivers.cpp)
int launch_backsub_rhs(psba_ctx* c, const double* gb, double* dp_out);

// ---- index (ctx.cpp, k_index.hip) ----
int check_obs_order(int nc, int np, int no, const int* iidx, const int* jidx);  // host validation
// frees the previous index and allocates every index buffer and the chunk partials
int alloc_index(psba_ctx* c, long nitems, int npairs, int n_cchunk, int n_schunk);

// ---- device-side LM control (k_lm.hip) ----
int launch_lm_init(psba_ctx* c, int itno, int max_iter, double rel_tol, int no_handoff);
int launch_lm_mu0(psba_ctx* c);
int launch_lm_control(psba_ctx* c, LmDev* out);  // out: the state after the decision (pinned host slot)
// gates of the launchers in device-control mode (nullptr otherwise)
inline const unsigned char* gate_of(const psba_ctx* c, int which) { return c->devctl ? c->gate + which : nullptr; }
// Under device control a trial's cost partials (one per 256 observations) sit
// in red_part after the step statistics' (two per 256 parameters), so that the
// LM control kernel can sum both: red_part holds 8 per block of the larger.
inline long trial_cost_part_off(const psba_ctx* c) { return 2L * ((c->ntot + 255) / 256); }
// and the commit's |p|^2 partials (one per 256 parameters) after those, summed
// by the next trial's LM control kernel
inline long commit_part_off(const psba_ctx* c) { return trial_cost_part_off(c) + (c->no + 255) / 256; }
inline const double* mu_of(const psba_ctx* c) { return c->devctl ? &c->lm->mu : nullptr; }

// ---- launchers (k_tr.hip) ----
int launch_cauchy(psba_ctx* c, double* out, const double* g, double gtg, double gtBg);
int launch_tr_dots(psba_ctx* c, const double* PU, const double* PB, const double* g, int mode);  // SC_D0..D4
int launch_comb(psba_ctx* c, double* out, const double* a, const double* b, double s1, double s2, int op);

}  // namespace psba
