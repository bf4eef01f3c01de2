// Device state of a batch of independent problems (psba_batch.h).
//
// Layout: the problems' cameras, points and observations are concatenated
// (global indices); per-problem offsets delimit each problem's ranges.  The
// reduced system of problem p is a dense n_pad(p) x n_pad(p) block at S_off(p)
// (n_pad a multiple of 64, padding rows/cols the identity); right-hand side,
// forward and back substitution vectors are padded the same way at vec_off(p).
// Every device buffer is a view into one arena allocation that only grows.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <vector>

#include "common.hpp"
#include "psba_batch.h"

namespace psba {
// index of one problem in its own numbering: point CSR, camera-ordered
// observations, camera chunks, camera-pair items and chunks, lower-block pairs
struct LocalIndex {
  std::vector<int> pt_ptr, cam_obs, cb, ce, ccp, sb, se, psp, lbpair;
  std::vector<int2> items;
  int npairs = 0;
};
}  // namespace psba

// Per-problem outcome, copied to pinned host memory at the end of every round
// (entry 0 is a header whose `active` field counts the problems still running).
struct BatchStat {
  double cost, init_err;
  int itno, code, active, pad;
};

struct psba_batch {
  int cnp = 0, precision = 0, device = 0;
  int cap_p = 0;
  hipStream_t stream = nullptr;

  // host copies of the problems added since the last reset
  struct Staged {
    int nc, np, no;
    std::vector<double> K, meas, q0, cams, pts;
    std::vector<int> iidx, jidx;
    std::vector<unsigned char> fix;
    std::shared_ptr<psba::LocalIndex> index;  // built ahead of the commit (optional)
  };
  std::vector<Staged> staged;

  // committed sizes
  int P = 0;                           // problems
  long NC = 0, NP = 0, NO = 0, N = 0;  // cameras, points, observations, camera parameters
  long n_items = 0, S_len = 0, V_len = 0;
  int n_cchunk = 0, n_schunk = 0, nlb = 0, npairs = 0, n_rchunk = 0;
  bool committed = false;

  // device arena; the inputs occupy its first input_bytes and are uploaded
  // from a pinned staging image with one copy
  char* arena = nullptr;
  size_t arena_bytes = 0, input_bytes = 0;
  char* staging = nullptr;
  size_t staging_bytes = 0;
  hipEvent_t ev_staged = nullptr;

  // problem data
  double *K = nullptr, *q0 = nullptr, *meas = nullptr, *cams = nullptr, *pts = nullptr;
  double *cams0 = nullptr, *pts0 = nullptr;
  unsigned char* fix = nullptr;
  int *obs_cam = nullptr, *obs_pt = nullptr, *obs_prob = nullptr;
  int *pt_ptr = nullptr, *pt_prob = nullptr;
  int *cam_obs = nullptr, *cam_prob = nullptr;
  int *cchunk_b = nullptr, *cchunk_e = nullptr, *ccp = nullptr, *cchunk_prob = nullptr;
  int2* items = nullptr;
  int *schunk_b = nullptr, *schunk_e = nullptr, *psp = nullptr, *schunk_prob = nullptr;
  int *lb_prob = nullptr, *lb_k = nullptr, *lb_l = nullptr, *lb_pair = nullptr;
  int *rchunk_b = nullptr, *rchunk_e = nullptr, *rchunk_prob = nullptr, *rcp = nullptr;  // residual chunks
  int *cam_off = nullptr, *pt_off = nullptr, *obs_off = nullptr, *npad = nullptr, *vec_off = nullptr;
  long* S_off = nullptr;
  std::vector<int> h_cam_off, h_pt_off, h_obs_off;
  // dense solve: one workgroup per problem, or (chol_phased) one launch per
  // phase with the task tables chol_fo ((T_max-1) x (P+1)) and chol_bo
  // (T_max x (P+1)) and per-launch grid sizes h_Gf / h_Gb
  int T_max = 0;
  bool chol_phased = false;
  int *chol_fo = nullptr, *chol_bo = nullptr;
  std::vector<int> h_Gf, h_Gb;

  // work (R = double on the plain path, float on the mixed path)
  void *e = nullptr, *JA = nullptr, *JB = nullptr, *W = nullptr;
  void *ucpart = nullptr, *schpart = nullptr, *eapart = nullptr;
  double *U = nullptr, *V = nullptr, *g = nullptr, *uvdiag = nullptr, *Vinv = nullptr, *z = nullptr;
  double *S = nullptr, *Lf = nullptr, *ea = nullptr, *bw = nullptr, *yv = nullptr, *xs = nullptr;
  double *dp = nullptr, *eb = nullptr, *rpart = nullptr;

  // per-problem LM state
  double *cost = nullptr, *tcost = nullptr, *mu = nullptr, *p_L2 = nullptr;
  double *dp2 = nullptr, *rhoden = nullptr, *init_err = nullptr, *fok = nullptr;
  int *nu = nullptr, *itno = nullptr, *code = nullptr;
  unsigned char *active = nullptr, *need_jac = nullptr, *accept = nullptr, *first = nullptr;
  unsigned char* all_on = nullptr;  // every problem (initial cost)
  BatchStat* stat = nullptr;        // device, P + 1 entries
  BatchStat* h_stat = nullptr;      // pinned host copies, 2 x (P + 1) entries
  size_t h_stat_cap = 0;

  // solve in flight
  bool running = false, finished = false;
  int rounds = 0, max_iter = 100, max_rounds = 0;
  double rel_tol = 1e-8;
  int graph_iter = -1;
  double graph_tol = -1;
  // two captures of the round, identical but for the pinned status buffer
  // they write (round r writes buffer r % 2), so the status of a checked
  // round stays stable until the round two later is launched
  hipGraph_t graph[2] = {nullptr, nullptr};
  hipGraphExec_t graph_exec[2] = {nullptr, nullptr};
  // rounds enqueued as plain launches instead (PSBA_BATCH_EAGER=1 at
  // creation): the same kernels and arguments, for tools that do not follow
  // graph launches
  bool eager = false;
  static constexpr int kDepth = 2;  // rounds queued ahead of the host's check
  hipEvent_t ev_round[kDepth] = {nullptr, nullptr};
  int launched = 0, checked = 0;
  hipEvent_t ev_start = nullptr, ev_stop = nullptr;
  bool stop_pending = false;  // ev_stop recorded, elapsed time not read yet
  double last_ms = 0;

  // launch shapes chosen by psba_batch_autotune
  int obs_block = 256, pt_block = 256;

  bool mixed() const { return precision == PSBA_PREC_MIXED; }
};

namespace psba {

constexpr int kResidChunk = 1024;  // observations per residual chunk of a batch
constexpr int kPhasedMinT = 3;     // dense tiles from which a batch's solve runs phase by phase

// stage launchers for batches (defined next to the kernels they launch)
int launch_b_jacobian(psba_batch* b);       // e, JA, JB, W, U, V, g, uvdiag for need_jac problems
int launch_b_vinv(psba_batch* b);           // Vinv, z with each problem's mu
int launch_b_schur(psba_batch* b);          // S blocks (with padding), ea
int launch_b_factor_solve(psba_batch* b);   // dp camera part, fok
int launch_b_backsub(psba_batch* b);        // dp point part
int launch_b_lm(psba_batch* b, int stage);  // per-problem LM stages (k_batch.hip)

enum BatchLmStage { BLM_INIT = 0, BLM_MU0, BLM_STEP_STATS, BLM_TRIAL_COST, BLM_CONTROL, BLM_COMMIT, BLM_COUNT };

// host side, for the serving loop (batch.cpp)
int batch_add_staged(psba_batch* b, psba_batch::Staged&& s);  // checked by the caller
// Check the oldest queued round (blocking: wait for it), hand its status to
// cb (if non-null) while it is stable, then keep kDepth rounds queued.
// Returns 1 once every problem has stopped, 0 otherwise.
int batch_advance(psba_batch* b, bool blocking, void (*cb)(void*, const BatchStat*) = nullptr,
                  void* ctx = nullptr);
const BatchStat* batch_final_stat(const psba_batch* b);  // after the solve finished
long long staged_bytes(int cnp, int precision, const psba_batch::Staged& s);
int build_local_index(const psba_batch::Staged& s, int cnp, LocalIndex& L);
int batch_release(psba_batch* b);                             // free the arena (idle batch)

}  // namespace psba
