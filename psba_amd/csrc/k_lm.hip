// Levenberg-Marquardt control of one problem on the device.
//
// The decision of PSBA/levmar.cpp:124-245 (gain ratio :271-280, damping
// update :202-204, nu doubling :238-243, the |dp| tests :185-197) taken by
// one thread after each trial, so that a solve's trials are queued back to
// back with no host round trip between them.  The state lives in LmDev; three
// one-byte gates steer the kernels of the next trial:
//   gate[kGateActive] the solve is running (mu0, the control itself),
//   gate[kGateJac]    the trial opens an iteration (Jacobian + normal equations),
//   gate[kGateAccept] the trial was accepted (commit of p + dp).
// The arithmetic and the order of the tests are those of psba_levmar
// (drivers.cpp), which follows the reference line by line; both produce the
// same trace records.
#include "context.hpp"

// No FMA contraction in this file: the damping update and the tests round
// exactly as psba_levmar's host code does (x86-64 without FMA), so the device
// and host decisions agree bit for bit.
#pragma clang fp contract(off)

namespace psba {

namespace {

constexpr double kInitMu = 1e-3;       // PSBA/psba.h:6
constexpr double kStopThresh = 1e-12;  // PSBA/psba.h:7
constexpr double kEps = 1e-12;         // PSBA/psba.h:8

__device__ void lm_push(LmDev* s, psba_trace_rec* tr, const psba_trace_rec& r) {
  if (s->ntrace < s->trace_cap) tr[s->ntrace] = r;
  s->ntrace++;
}

// end of an LM iteration (PSBA/levmar.cpp:247-250 and the loop header :103)
__device__ void lm_end_iteration(LmDev* s, unsigned char* gate) {
  if (s->ex_L2 <= kStopThresh) s->flag = PSBA_ITER_ERR_SMALL_ENOUGH;
  s->itno++;
  const bool go = s->itno < s->max_iter && s->flag == PSBA_ITER_CONTINUE;
  gate[kGateActive] = go ? 1 : 0;
  gate[kGateJac] = go ? 1 : 0;
}

}  // namespace

// state at the entry of psba_levmar: the cost just evaluated at the current
// parameters (scal[SC_COST]); nothing runs when the iteration budget is spent
__global__ void k_lm_init(LmDev* s, unsigned char* gate, const double* __restrict__ scal, int itno, int max_iter,
                          double rel_tol, int no_handoff, int trace_cap) {
  s->ex_L2 = scal[SC_COST];
  s->entry_cost = s->ex_L2;
  s->mu = 0.0;
  s->p_L2 = 0.0;
  s->rel_tol = rel_tol;
  s->nu = 2;
  s->gooditer = 0;
  s->flag = PSBA_ITER_CONTINUE;
  s->itno = itno;
  s->max_iter = max_iter;
  s->first = 1;
  s->p_pending = 0;
  s->no_handoff = no_handoff;
  s->ntrace = 0;
  s->trace_cap = trace_cap;
  s->trials = 0;
  const bool go = itno < max_iter;
  gate[kGateActive] = go ? 1 : 0;
  gate[kGateJac] = go ? 1 : 0;
  gate[kGateAccept] = 0;
}

// mu0 = tau * max(UVdiag) on the solve's first iteration (PSBA/levmar.cpp:113-120)
// One workgroup of 1024 threads, eight loads in flight per thread (a maximum
// does not depend on the order in which it is taken).
__global__ __launch_bounds__(1024) void k_lm_mu0(LmDev* s, const unsigned char* __restrict__ gate,
                                                 const double* __restrict__ uvdiag, long ntot) {
  __shared__ double sh[16];
  if (!gate[kGateActive] || !s->first) return;
  double m = 0.0;
  for (long q0 = threadIdx.x; q0 < ntot; q0 += 8L * blockDim.x) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const long q = q0 + (long)u * blockDim.x;
      v[u] = q < ntot ? uvdiag[q] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) m = fmax(m, v[u]);
  }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    m = sh[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) m = fmax(m, sh[w]);
    s->mu = kInitMu * m;
    s->first = 0;
    s->p_L2 = 1e+3;
    s->nu = 2;
  }
}

// The trial's figures for the decision: the factorisation flag, |dp|^2, the
// gain-ratio denominator, the cost at p + dp and (after an accept) |p|^2.
struct TrialFig {
  double factor_ok, dp2, rhoden, cost, p2;
};

// the decision after one trial (PSBA/levmar.cpp:124-245)
__device__ void lm_decide(LmDev* s, unsigned char* gate, const TrialFig& f, psba_trace_rec* __restrict__ trace);

// One block: first the trial's step statistics (|dp|^2 and the gain-ratio
// denominator, from k_trial_eval's step partials) and cost (its cost
// partials) are summed into scal, as k_sum_partials<2> and <1> would, and,
// after an accepted trial (p_pending), |p|^2 from the commit's partials; then
// thread 0 decides.
// The state is read into LDS with the partials' loads (one word per thread),
// decided on there, and written back to `s` and to `out`, a slot of pinned
// host memory that the host reads once the trial's event has completed, one
// word per thread again.
__global__ __launch_bounds__(256) void k_lm_control(LmDev* s, unsigned char* gate, double* __restrict__ scal,
                                                    psba_trace_rec* __restrict__ trace, LmDev* out,
                                                    const double* __restrict__ step_part, int nb_step,
                                                    const double* __restrict__ cost_part, int nb_cost,
                                                    const double* __restrict__ p2_part, int nb_p2) {
  constexpr int NW = sizeof(LmDev) / 8;
  static_assert(sizeof(LmDev) % 8 == 0, "LmDev moves as 8-byte words");
  constexpr int B = 8;
  __shared__ double sh[16];
  __shared__ LmDev ls;
  const int t = threadIdx.x, T = blockDim.x;
  if (t < NW) ((unsigned long long*)&ls)[t] = ((const unsigned long long*)s)[t];
  TrialFig f{0.0, 0.0, 0.0, 0.0, 0.0};
  if (t == 0) f.factor_ok = scal[SC_FACTOR_OK];
  const bool active = gate[kGateActive];
  const bool want_p2 = active && s->p_pending;
  // The three sets of partials (|dp|^2 and the gain-ratio denominator, the
  // cost, |p|^2) summed as sum_partials_block sums each set (thread t adds
  // blocks t, t + T, .. in order, then the fixed-order block sum), with the
  // loads of all three in flight together.
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  const int nmax = max(nb_step, max(nb_cost, want_p2 ? nb_p2 : 0));
  for (int b0 = t; b0 < nmax; b0 += B * T) {
    double vs[B][2], vc[B], vp[B];
#pragma unroll
    for (int u = 0; u < B; ++u) {
      const int b = b0 + u * T;
      const int bs = min(b, nb_step - 1), bc = min(b, nb_cost - 1);
      vs[u][0] = step_part[(size_t)bs * 2];
      vs[u][1] = step_part[(size_t)bs * 2 + 1];
      vc[u] = cost_part[bc];
      vp[u] = want_p2 ? p2_part[min(b, nb_p2 - 1)] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < B; ++u) {
      const int b = b0 + u * T;
      const bool ls_ = b < nb_step, lc = b < nb_cost, lp = want_p2 && b < nb_p2;
      const double t0 = acc[0] + vs[u][0], t1 = acc[1] + vs[u][1], t2 = acc[2] + vc[u], t3 = acc[3] + vp[u];
      acc[0] = ls_ ? t0 : acc[0];
      acc[1] = ls_ ? t1 : acc[1];
      acc[2] = lc ? t2 : acc[2];
      acc[3] = lp ? t3 : acc[3];
    }
  }
  block_sum<4>(acc, sh);
  // (ls is visible here: the block sum passes workgroup barriers)
  if (t == 0) {
    scal[SC_DP2] = acc[0];
    scal[SC_RHODEN] = acc[1];
    scal[SC_COST] = acc[2];
    if (want_p2) scal[SC_P2] = acc[3];
    f.dp2 = acc[0];
    f.rhoden = acc[1];
    f.cost = acc[2];
    f.p2 = acc[3];
    lm_decide(&ls, gate, f, trace);
  }
  __syncthreads();
  if (t < NW) {
    const unsigned long long v = ((const unsigned long long*)&ls)[t];
    ((unsigned long long*)s)[t] = v;
    ((unsigned long long*)out)[t] = v;
  }
}

__device__ void lm_decide(LmDev* s, unsigned char* gate, const TrialFig& f, psba_trace_rec* __restrict__ trace) {
  gate[kGateAccept] = 0;
  if (!gate[kGateActive]) {
    gate[kGateJac] = 0;
    return;
  }
  s->trials++;
  gate[kGateJac] = 0;
  if (s->p_pending) {
    s->p_L2 = f.p2;
    s->p_pending = 0;
  }
  const double nan = __builtin_nan("");
  psba_trace_rec rec{s->itno, 0, 0, 0, nan, s->mu, nan, 0.0, nan};
  // a timed-out wait inside the dense solve is an execution error, not a
  // rejected trial
  if (f.factor_ok < 0.0) {
    s->flag = PSBA_ERR_HIP;
    gate[kGateActive] = 0;
    return;
  }
  const bool ok = f.factor_ok != 0.0;
  rec.factor_ok = ok;
  if (ok) {
    const double dp_L2 = f.dp2;
    rec.dp_sq = dp_L2;
    if (dp_L2 < s->p_L2 * kStopThresh * kStopThresh) {
      s->flag = PSBA_ITER_DP_NO_CHANGE;
      lm_push(s, trace, rec);
      lm_end_iteration(s, gate);
      return;
    }
    if (dp_L2 >= (s->p_L2 + kStopThresh) / (kEps * kEps)) {
      s->flag = PSBA_ITER_ERR;
      lm_push(s, trace, rec);
      lm_end_iteration(s, gate);
      return;
    }
    const double new_ex_L2 = f.cost;
    const double rho = (s->ex_L2 - new_ex_L2) / f.rhoden;
    rec.cost = new_ex_L2;
    rec.rho = rho;
    if (rho > 0) {
      double tmp = 2 * rho - 1;
      tmp = 1.0 - tmp * tmp * tmp;
      s->mu = s->mu * ((tmp >= (1.0 / 3.0)) ? tmp : (1.0 / 3.0));
      s->nu = 2;
      gate[kGateAccept] = 1;
      s->p_pending = 1;
      const double rel_dec = (s->ex_L2 - new_ex_L2) / s->ex_L2;
      s->ex_L2 = new_ex_L2;
      rec.accepted = 1;
      lm_push(s, trace, rec);
      if (s->rel_tol > 0 && rel_dec < s->rel_tol) {
        s->flag = PSBA_ITER_DP_NO_CHANGE;
      } else if (s->no_handoff) {
        s->gooditer = 0;
      } else if (fabs(rho - 1) < (1.0 / 5.0)) {
        s->gooditer++;
        if (s->gooditer >= 5) s->flag = PSBA_ITER_TURN_TO_TR;
      } else {
        s->gooditer = 0;
      }
      lm_end_iteration(s, gate);
      return;
    }
  } else {
    s->gooditer = 0;
  }
  lm_push(s, trace, rec);
  s->mu *= s->nu;
  if (s->nu >= (1 << 30)) {  // PSBA/levmar.cpp:238-243: int nu doubles until 2*nu wraps
    s->flag = PSBA_ITER_ERR;
    lm_end_iteration(s, gate);
    return;
  }
  s->nu = 2 * s->nu;
}

int launch_lm_init(psba_ctx* c, int itno, int max_iter, double rel_tol, int no_handoff) {
  hipLaunchKernelGGL(k_lm_init, dim3(1), dim3(1), 0, c->stream, c->lm, c->gate, c->d_scal, itno, max_iter, rel_tol,
                     no_handoff, c->lm_trace_cap);
  PSBA_LAUNCH_TRY();
  return PSBA_OK;
}

int launch_lm_mu0(psba_ctx* c) {
  hipLaunchKernelGGL(k_lm_mu0, dim3(1), dim3(1024), 0, c->stream, c->lm, c->gate, c->uvdiag, (long)c->ntot);
  PSBA_LAUNCH_TRY();
  return PSBA_OK;
}

int launch_lm_control(psba_ctx* c, LmDev* out) {
  hipLaunchKernelGGL(k_lm_control, dim3(1), dim3(256), 0, c->stream, c->lm, c->gate, c->d_scal, c->lm_trace, out,
                     (const double*)c->red_part, (int)ceil_div(c->ntot, 256),
                     (const double*)(c->red_part + trial_cost_part_off(c)), (int)ceil_div(c->no, 256),
                     (const double*)(c->red_part + commit_part_off(c)), (int)ceil_div(c->ntot, 256));
  PSBA_LAUNCH_TRY();
  return PSBA_OK;
}

}  // namespace psba
