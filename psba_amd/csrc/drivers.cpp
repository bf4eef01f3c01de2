// Levenberg-Marquardt and dogleg drivers of the C-ABI.
//
// Every trial runs as one batch of kernels on the context's stream; the host
// reads back one block of scalars per trial (factor status, |dp|^2, gain-ratio
// denominator, trial cost) and applies the reference's control logic:
//   LM      PSBA/levmar.cpp:45-256 (+ compute_rho :271-280)
//   dogleg  PSBA/trust_region.cpp:49-288, compute_PB :292-405,
//           compute_p_2 :520-595
//   hybrid  PSBA/main.cpp:193-208
// The damping is applied on the fly to the undamped U/V blocks, so the
// reference's update_UV / restore_UVdiag pair has no counterpart.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "context.hpp"

using namespace psba;

namespace {

constexpr double kInitMu = 1e-3;       // PSBA/psba.h:6
constexpr double kStopThresh = 1e-12;  // PSBA/psba.h:7
constexpr double kEps = 1e-12;         // PSBA/psba.h:8
constexpr double kEps2 = 1e-12;        // PSBA/psba.h:9
constexpr double kMaxDelta = 10000.0;  // PSBA/trust_region.cpp:18

int fetch(psba_ctx* c) {
  PSBA_HIP_TRY(hipMemcpyAsync(c->h_scal, c->d_scal, sizeof(double) * SC_COUNT, hipMemcpyDeviceToHost, c->stream));
  PSBA_HIP_TRY(hipStreamSynchronize(c->stream));
  return PSBA_OK;
}

// the dense solve's status after a fetch: a timed-out wait inside the launch
// (its workgroups were not all resident) is an execution error, not a failed
// factorisation
int factor_status(psba_ctx* c, bool* ok) {
  const double f = c->h_scal[SC_FACTOR_OK];
  if (f < 0.0) {
    set_last_error(__FILE__, __LINE__,
                   "dense solve: a grid-barrier wait timed out (the launch's workgroups were not all resident)");
    return PSBA_ERR_HIP;
  }
  *ok = f != 0.0;
  return PSBA_OK;
}

void push(psba_trace_rec* tr, int cap, int* len, const psba_trace_rec& r) {
  if (!tr || !len) return;
  if (*len < cap) tr[*len] = r;
  ++*len;
}

int opt_max_iter(const psba_opts* o) { return (o && o->max_iter > 0) ? o->max_iter : 50; }
int opt_max_pb(const psba_opts* o) { return (o && o->max_pb_tries > 0) ? o->max_pb_tries : 64; }

int fail_unsupported_mixed(const char* what) {
  char buf[200];
  snprintf(buf, sizeof(buf), "%s: the dogleg runs on FP64 contexts only (create with PSBA_PREC_FP64)", what);
  set_last_error(__FILE__, __LINE__, buf);
  return PSBA_ERR_UNSUPPORTED;
}

int check_ready(psba_ctx* c) {
  if (!c) return PSBA_ERR_ARG;
  if (!c->uploaded || !c->indexed) return PSBA_ERR_STATE;
  PSBA_HIP_TRY(hipSetDevice(c->device));
  return PSBA_OK;
}

// optional stage timing with events on the context's stream
struct StageTimer {
  psba_ctx* c;
  double* slot;
  StageTimer(psba_ctx* cc, double* s) : c(cc), slot(s) {
    if (c->timing_on) (void)hipEventRecord(c->ev0, c->stream);
  }
  ~StageTimer() {
    if (!c->timing_on) return;
    (void)hipEventRecord(c->ev1, c->stream);
    (void)hipEventSynchronize(c->ev1);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, c->ev0, c->ev1);
    *slot += ms;
  }
};

// damped reduced system + factor + solve + back-substitution: dp
int trial_solve(psba_ctx* c, double mu) {
  {
    StageTimer t(c, &c->timing.schur);
    PSBA_TRY(launch_vinv(c, mu));
    PSBA_TRY(launch_schur(c, mu));
  }
  {
    StageTimer t(c, &c->timing.factor);
    PSBA_TRY(launch_factor_solve(c));
  }
  {
    StageTimer t(c, &c->timing.backsub);
    PSBA_TRY(launch_backsub(c));
  }
  if (c->refine_steps > 0) {
    StageTimer t(c, &c->timing.other);
    PSBA_TRY(refine_step(c, mu, c->refine_steps));
  }
  c->timing.trials++;
  return PSBA_OK;
}

int jacobian(psba_ctx* c, double coeff, double coeff_g) {
  StageTimer t(c, &c->timing.jac);
  c->timing.jacobians++;
  return launch_jacobian_normal(c, coeff, coeff_g);
}

// ---------------------------------------------------------------------------
// psba_levmar with the decision on the device (k_lm.hip).  A trial is one
// fixed sequence of launches -- Jacobian + normal equations (gated: only when
// the trial opens an iteration), mu0 (the solve's first trial), V^-1, Schur
// complement, dense factor + solve, back substitution, step statistics,
// trial cost, the control kernel (which also writes the LM state to one of
// two pinned host slots), the commit (gated: accepted trials).  Two
// trials are kept queued: while the GPU runs trial t+1 the host waits for
// trial t's event and reads its state, so the device never waits for the
// host.  Only the kernels whose work depends on the decision are gated (the
// Jacobian stage, mu0, the commit): a gate read in front of every workgroup
// of the latency-bound Schur and dense-solve kernels cost more over a solve
// (14 us a trial at the 88-camera scene) than the one trial queued behind the
// stop, which runs those kernels to no effect.  The trace records are read
// back once, at the end.
// ---------------------------------------------------------------------------
constexpr int kLmDepth = 2;

int lm_device_alloc(psba_ctx* c, int trace_cap) {
  if (!c->lm) {
    PSBA_HIP_TRY(hipMalloc((void**)&c->lm, sizeof(LmDev)));
    PSBA_HIP_TRY(hipMalloc((void**)&c->gate, kGateCount));
    PSBA_HIP_TRY(hipHostMalloc((void**)&c->h_lm, sizeof(LmDev) * kLmDepth, hipHostMallocDefault));
    for (auto& e : c->ev_lm) PSBA_HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  }
  if (trace_cap > c->lm_trace_cap) {
    if (c->lm_trace) (void)hipFree(c->lm_trace);
    c->lm_trace = nullptr;
    c->lm_trace_cap = 0;
    PSBA_HIP_TRY(hipMalloc((void**)&c->lm_trace, sizeof(psba_trace_rec) * (size_t)trace_cap));
    c->lm_trace_cap = trace_cap;
  }
  return PSBA_OK;
}

int lm_enqueue_trial(psba_ctx* c, int slot, bool first) {
  if (first) {
    PSBA_TRY(launch_jacobian_normal(c, 1.0, 1.0));
    PSBA_TRY(launch_lm_mu0(c));
    PSBA_TRY(launch_vinv(c, 0.0));
  } else {
    PSBA_TRY(launch_jacobian_normal_vinv(c));
  }
  PSBA_TRY(launch_schur(c, 0.0));
  PSBA_TRY(launch_factor_solve(c));
  PSBA_TRY(launch_backsub(c));
  PSBA_TRY(launch_trial_eval(c));
  PSBA_TRY(launch_lm_control(c, c->h_lm + slot));
  PSBA_TRY(launch_commit(c));
  PSBA_HIP_TRY(hipEventRecord(c->ev_lm[slot], c->stream));
  return PSBA_OK;
}

struct DevCtl {
  psba_ctx* c;
  explicit DevCtl(psba_ctx* cc) : c(cc) { c->devctl = true; }
  ~DevCtl() { c->devctl = false; }
};

int lm_device(psba_ctx* c, const psba_opts* o, int* itno, double* finalErr, psba_trace_rec* trace, int cap,
              int* len) {
  const int max_iter = opt_max_iter(o);
  const bool no_handoff = o && (o->flags & PSBA_FLAG_NO_HANDOFF);
  const double rel_tol = o ? o->rel_tol : 0.0;
  // at most 31 rejected trials per iteration (nu doubles from 2 to 2^31)
  const long max_trials = (long)std::max(0, max_iter - *itno) * 32 + 1;
  PSBA_TRY(lm_device_alloc(c, (int)std::min<long>(max_trials, 1L << 20)));
  PSBA_TRY(launch_residual(c, false));
  PSBA_TRY(launch_lm_init(c, *itno, max_iter, rel_tol, no_handoff ? 1 : 0));
  {
    DevCtl on(c);
    long launched = 0, checked = 0;
    if (*itno < max_iter) {
      for (;;) {
        while (launched - checked < kLmDepth && launched < max_trials) {
          PSBA_TRY(lm_enqueue_trial(c, (int)(launched % kLmDepth), launched == 0));
          launched++;
        }
        const int slot = (int)(checked % kLmDepth);
        PSBA_HIP_TRY(hipEventSynchronize(c->ev_lm[slot]));
        const LmDev st = c->h_lm[slot];
        checked++;
        if (st.flag < 0 || st.flag != PSBA_ITER_CONTINUE || st.itno >= st.max_iter || checked >= max_trials) break;
      }
    }
  }
  PSBA_HIP_TRY(hipMemcpyAsync(c->h_lm, c->lm, sizeof(LmDev), hipMemcpyDeviceToHost, c->stream));
  PSBA_HIP_TRY(hipStreamSynchronize(c->stream));
  const LmDev st = c->h_lm[0];
  if (st.flag < 0) {
    set_last_error(__FILE__, __LINE__,
                   "dense solve: a grid-barrier wait timed out (the launch's workgroups were not all resident)");
    return st.flag;
  }
  if (trace && len) {
    const int base = *len;
    const int take = std::max(0, std::min(st.ntrace, std::min(st.trace_cap, cap - base)));
    if (take > 0)
      PSBA_HIP_TRY(hipMemcpy(trace + base, c->lm_trace, sizeof(psba_trace_rec) * (size_t)take,
                             hipMemcpyDeviceToHost));
    *len = base + st.ntrace;
  }
  c->entry_cost = st.entry_cost;
  c->timing.trials += st.trials;
  *itno = st.itno;
  *finalErr = st.ex_L2;
  return st.flag;
}

}  // namespace


namespace psba {

// FP64 iterative refinement of the mixed path's step dp (at damping mu):
// corrections delta_k from the FP64 residual of the damped normal equations,
// solved with the FP32-assembled blocks and their FP64 factor.  Refinement
// contracts while that solve's relative error times the conditioning stays
// below 1; at small damping it can exceed 1 and the corrections grow.  The
// sweeps therefore stop at the first correction larger than half the previous
// one, and if that happens at the second sweep the step returns to the
// unrefined one (the first correction was not a contraction either).
int refine_step(psba_ctx* c, double mu, int sweeps) {
  PSBA_TRY(launch_axpby(c, c->vecD, 1.0, c->dp, 0.0, c->dp, c->ntot));  // the unrefined step
  double prev = 0.0;
  for (int k = 0; k < sweeps + 1; ++k) {
    PSBA_TRY(launch_refine(c, mu, 1));
    PSBA_TRY(launch_dot(c, c->vecB, c->vecB, SC_D5));
    PSBA_TRY(fetch(c));
    const double d = c->h_scal[SC_D5];
    if (!std::isfinite(d) || (k > 0 && d > 0.25 * prev)) {
      if (k == 1) PSBA_TRY(launch_axpby(c, c->dp, 1.0, c->vecD, 0.0, c->vecD, c->ntot));
      return PSBA_OK;
    }
    if (k == sweeps) return PSBA_OK;  // the last pass only measured the correction
    PSBA_TRY(launch_axpby(c, c->dp, 1.0, c->dp, 1.0, c->vecB, c->ntot));
    prev = d;
  }
  return PSBA_OK;
}

}  // namespace psba

extern "C" {

int psba_levmar(psba_ctx* c, const psba_opts* o, int* itno, double* finalErr, psba_trace_rec* trace, int cap,
                int* len) {
  PSBA_TRY(check_ready(c));
  if (!itno || !finalErr) return PSBA_ERR_ARG;
  const int max_iter = opt_max_iter(o);
  const bool no_handoff = o && (o->flags & PSBA_FLAG_NO_HANDOFF);
  const double rel_tol = o ? o->rel_tol : 0.0;
  c->refine_steps = (o && c->mixed() && o->refine_steps > 0) ? o->refine_steps : 0;
  // the decision on the device, unless the solve needs the host between
  // kernels: refinement sweeps (a host contraction test per sweep), per-stage
  // timers (an event wait per stage) or the host-loop option
  if (c->refine_steps == 0 && !c->timing_on && !(o && (o->flags & PSBA_FLAG_HOST_CONTROL)))
    return lm_device(c, o, itno, finalErr, trace, cap, len);
  double mu = 0.0, p_L2 = 0.0;
  int gooditer_cnt = 0, nu = 2;
  bool first = true, p_pending = false;
  PSBA_TRY(launch_residual(c, false));
  PSBA_TRY(fetch(c));
  double ex_L2 = c->h_scal[SC_COST];
  c->entry_cost = ex_L2;
  int flag = PSBA_ITER_CONTINUE;
  for (; *itno < max_iter && flag == PSBA_ITER_CONTINUE; (*itno)++) {
    PSBA_TRY(jacobian(c, 1.0, 1.0));
    if (first) {
      // with refinement the mixed path takes mu0 from the FP64 diagonal too
      if (c->refine_steps > 0) PSBA_TRY(launch_refine(c, 0.0, -1));
      PSBA_TRY(launch_maxdiag(c));
      PSBA_TRY(fetch(c));
      mu = kInitMu * c->h_scal[SC_MAXDIAG];
      first = false;
      p_L2 = 1e+3;
      nu = 2;
    }
    while (true) {
      psba_trace_rec rec{*itno, 0, 0, 0, NAN, mu, NAN, 0.0, NAN};
      PSBA_TRY(trial_solve(c, mu));
      {
        StageTimer t(c, &c->timing.resid);
        PSBA_TRY(launch_step_stats(c, mu));
        PSBA_TRY(launch_residual(c, true));
      }
      PSBA_TRY(fetch(c));
      if (p_pending) {
        p_L2 = c->h_scal[SC_P2];
        p_pending = false;
      }
      bool ok = false;
      PSBA_TRY(factor_status(c, &ok));
      rec.factor_ok = ok;
      if (ok) {
        const double dp_L2 = c->h_scal[SC_DP2];
        rec.dp_sq = dp_L2;
        if (dp_L2 < p_L2 * kStopThresh * kStopThresh) {
          flag = PSBA_ITER_DP_NO_CHANGE;
          push(trace, cap, len, rec);
          break;
        }
        if (dp_L2 >= (p_L2 + kStopThresh) / (kEps * kEps)) {
          flag = PSBA_ITER_ERR;
          push(trace, cap, len, rec);
          break;
        }
        const double new_ex_L2 = c->h_scal[SC_COST];
        const double rho = (ex_L2 - new_ex_L2) / c->h_scal[SC_RHODEN];
        rec.cost = new_ex_L2;
        rec.rho = rho;
        if (rho > 0) {
          double tmp = 2 * rho - 1;
          tmp = 1.0 - tmp * tmp * tmp;
          mu = mu * ((tmp >= (1.0 / 3.0)) ? tmp : (1.0 / 3.0));
          nu = 2;
          PSBA_TRY(launch_commit(c));
          p_pending = true;
          const double rel_dec = (ex_L2 - new_ex_L2) / ex_L2;
          ex_L2 = new_ex_L2;
          rec.accepted = 1;
          push(trace, cap, len, rec);
          if (rel_tol > 0 && rel_dec < rel_tol) {
            flag = PSBA_ITER_DP_NO_CHANGE;
            break;
          }
          if (no_handoff) {
            gooditer_cnt = 0;
          } else if (std::fabs(rho - 1) < (1.0 / 5.0)) {
            gooditer_cnt++;
            if (gooditer_cnt >= 5) {
              flag = PSBA_ITER_TURN_TO_TR;
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
      push(trace, cap, len, rec);
      mu *= nu;
      if (nu >= (1 << 30)) {  // levmar.cpp:238-243, int nu doubles until 2*nu wraps
        flag = PSBA_ITER_ERR;
        break;
      }
      nu = 2 * nu;
    }
    if (ex_L2 <= kStopThresh) flag = PSBA_ITER_ERR_SMALL_ENOUGH;
  }
  PSBA_HIP_TRY(hipStreamSynchronize(c->stream));
  *finalErr = ex_L2;
  return flag;
}

}  // extern "C"

namespace {

// PSBA/trust_region.cpp:292-405.  Returns 1 when P_B (in vecB) was produced.
int compute_PB(psba_ctx* c, double* lambda, int* solved) {
  *solved = 0;
  {
    StageTimer t(c, &c->timing.schur);
    PSBA_TRY(launch_vinv(c, *lambda));
    PSBA_TRY(launch_schur(c, *lambda));
  }
  if (!c->Saux) PSBA_HIP_TRY(hipMalloc((void**)&c->Saux, sizeof(double) * (size_t)c->n_pad * c->n_pad));
  PSBA_TRY(launch_copy_S(c, c->Saux, c->S));
  {
    StageTimer t(c, &c->timing.factor);
    PSBA_TRY(launch_factor_solve(c));
  }
  PSBA_TRY(fetch(c));
  c->timing.trials++;
  bool factored = false;
  PSBA_TRY(factor_status(c, &factored));
  if (!factored) {
    if (*lambda == 0.0) {
      StageTimer t(c, &c->timing.other);
      PSBA_TRY(launch_modchol(c, c->Saux));
      PSBA_TRY(fetch(c));
      *lambda = c->h_scal[SC_LAMBDA];
      return PSBA_OK;
    }
    *lambda = 2 * (*lambda);
    return PSBA_OK;
  }
  {
    StageTimer t(c, &c->timing.backsub);
    PSBA_TRY(launch_backsub(c));
  }
  PSBA_TRY(launch_axpby(c, c->vecB, -1.0, c->dp, 0.0, nullptr, c->ntot));
  *solved = 1;
  return PSBA_OK;
}

// PSBA/trust_region.cpp:520-595: the step P (in dp); returns its "norm" as the
// reference reports it (including the bookkeeping of :564-569).
int compute_p_2(psba_ctx* c, double pUtBpU, double pUtBpB, double pBtBpB, double delta, double* p_norm_out) {
  const double* PU = c->vecA;
  const double* PB = c->vecB;
  PSBA_TRY(launch_tr_dots(c, PU, PB, c->g, 0));
  PSBA_TRY(fetch(c));
  const double pUg = c->h_scal[SC_D0], pBg = c->h_scal[SC_D1], pUpU = c->h_scal[SC_D2], pBpB = c->h_scal[SC_D4];
  const double den = -pUtBpB * pUtBpB + pBtBpB * pUtBpU;
  const double eta1 = (pBg * pUtBpB) / den - (pBtBpB * pUg) / den;
  const double eta2 = (pUg * pUtBpB) / den - (pBg * pUtBpU) / den;
  PSBA_TRY(launch_comb(c, c->dp, PU, PB, eta1, eta2, 0));
  PSBA_TRY(launch_dot(c, c->dp, c->dp, SC_D5));
  PSBA_TRY(fetch(c));
  double p_norm = std::sqrt(c->h_scal[SC_D5]);
  if (p_norm > delta) {
    const double pU_norm = std::sqrt(pUpU), pB_norm = std::sqrt(pBpB);
    if (pU_norm > delta) {
      PSBA_TRY(launch_comb(c, c->dp, PU, PU, delta, pU_norm, 1));
      *p_norm_out = delta;
      return PSBA_OK;
    } else if (pB_norm <= delta) {
      PSBA_TRY(launch_axpby(c, c->dp, 1.0, PB, 0.0, nullptr, c->ntot));
      *p_norm_out = std::sqrt(p_norm + pBpB);
      return PSBA_OK;
    } else {
      PSBA_TRY(launch_tr_dots(c, PU, PB, c->g, 1));
      PSBA_TRY(fetch(c));
      double a = c->h_scal[SC_D0], b = c->h_scal[SC_D1], cc = c->h_scal[SC_D2];
      b = 2 * b;
      cc = cc - delta * delta;
      double disc = b * b - 4 * a * cc;
      if (std::fabs(disc) < 1e-12) disc = 0;
      const double tau = (-b + std::sqrt(disc)) / (2 * a);
      PSBA_TRY(launch_comb(c, c->dp, PU, PB, tau, 0.0, 2));
      *p_norm_out = delta;
      return PSBA_OK;
    }
  }
  *p_norm_out = p_norm;
  return PSBA_OK;
}

}  // namespace

extern "C" {

int psba_levmar_entry_cost(psba_ctx* c, double* sse) {
  if (!c || !sse) return PSBA_ERR_ARG;
  *sse = c->entry_cost;
  return PSBA_OK;
}

int psba_trust_region(psba_ctx* c, const psba_opts* o, int* itno, double* finalErr, psba_trace_rec* trace, int cap,
                      int* len) {
  PSBA_TRY(check_ready(c));
  if (c->mixed()) return fail_unsupported_mixed("psba_trust_region");
  if (!itno || !finalErr) return PSBA_ERR_ARG;
  const int max_iter = opt_max_iter(o), max_pb = opt_max_pb(o);
  int notgood_cnt = 0, good_iters = 0, nu = 2;
  double dk = 1, lambda = 0, origin_lambda = 0.0;
  PSBA_TRY(launch_residual(c, false));
  PSBA_TRY(fetch(c));
  double ex_L2 = c->h_scal[SC_COST];
  int flag = PSBA_ITER_CONTINUE;
  double* P_U = c->vecA;
  double* P_B = c->vecB;
  for (; *itno < max_iter; (*itno)++) {
    PSBA_TRY(jacobian(c, 2.0, -2.0));
    PSBA_TRY(launch_jmul_dots(c, c->g, nullptr));
    PSBA_TRY(launch_dot(c, c->g, c->g, SC_GG));
    PSBA_TRY(fetch(c));
    const double gtBg = 2 * c->h_scal[SC_D0];
    const double gtg = c->h_scal[SC_GG];
    PSBA_TRY(launch_cauchy(c, P_U, c->g, gtg, gtBg));
    int solved = 0, tries = 0;
    bool ret_early = false;
    while (!solved) {
      PSBA_TRY(compute_PB(c, &lambda, &solved));
      if (!solved) {
        if (origin_lambda != 0.0) {
          if (nu > 4) {
            flag = PSBA_ITER_TURN_TO_LM;
            *finalErr = ex_L2;
            ret_early = true;
            break;
          }
          lambda = lambda * nu;
          nu = nu * 2;
        }
        if (++tries >= max_pb) {
          flag = PSBA_ITER_TURN_TO_LM;
          *finalErr = ex_L2;
          ret_early = true;
          break;
        }
      } else {
        nu = 2;
        origin_lambda = lambda;
      }
    }
    if (ret_early) break;
    PSBA_TRY(launch_jmul_dots(c, P_U, P_B));
    PSBA_TRY(fetch(c));
    const double pUtBpU = 2 * c->h_scal[SC_D0];
    const double pUtBpB = 2 * c->h_scal[SC_D1];
    const double pBtBpB = 2 * c->h_scal[SC_D2];
    flag = PSBA_ITER_CONTINUE;
    while (flag == PSBA_ITER_CONTINUE) {
      psba_trace_rec rec{*itno, 1, 1, 0, NAN, dk, NAN, lambda, NAN};
      double p_norm = 0;
      PSBA_TRY(compute_p_2(c, pUtBpU, pUtBpB, pBtBpB, dk, &p_norm));
      rec.dp_sq = p_norm * p_norm;
      PSBA_TRY(launch_residual(c, true));
      PSBA_TRY(fetch(c));
      const double act_ex_L2 = c->h_scal[SC_COST];
      rec.cost = act_ex_L2;
      if (std::fabs((ex_L2 - act_ex_L2) / ex_L2) < kEps2) {
        flag = PSBA_ITER_DP_NO_CHANGE;
        push(trace, cap, len, rec);
        break;
      }
      PSBA_TRY(launch_jmul_dots(c, c->dp, nullptr));
      PSBA_TRY(launch_dot(c, c->g, c->dp, SC_D5));
      PSBA_TRY(fetch(c));
      const double Jx_norm = 2 * c->h_scal[SC_D0];
      double pred_ex_L2 = c->h_scal[SC_D5];
      pred_ex_L2 += ex_L2 + Jx_norm / 2;
      const double rho = (ex_L2 - act_ex_L2) / (ex_L2 - pred_ex_L2);
      rec.rho = rho;
      if (rho < (1.0 / 4.0) || act_ex_L2 > ex_L2) {
        dk = dk / 4;
      } else if (rho >= (3.0 / 4.0) && act_ex_L2 < ex_L2) {
        flag = PSBA_ITER_PASS;
        PSBA_TRY(launch_commit(c));
        *finalErr = act_ex_L2;
        dk = std::fmin(2 * dk, kMaxDelta);
        rec.accepted = 1;
      } else if (rho >= (1.0 / 4.0) && rho < (3.0 / 4.0) && act_ex_L2 < ex_L2) {
        flag = PSBA_ITER_PASS;
        PSBA_TRY(launch_commit(c));
        *finalErr = act_ex_L2;
        rec.accepted = 1;
      } else if (std::isnan(rho)) {
        flag = PSBA_ITER_TURN_TO_LM;
        *finalErr = ex_L2;
        push(trace, cap, len, rec);
        ret_early = true;
        break;
      }
      push(trace, cap, len, rec);
      if (std::fabs((act_ex_L2 - ex_L2) / ex_L2) <= kEps2) {
        flag = PSBA_ITER_ERR_SMALL_ENOUGH;
        break;
      }
      if (rho < 1.0 / 4) {
        notgood_cnt++;
        if (notgood_cnt >= 5) {
          flag = PSBA_ITER_TURN_TO_LM;
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
    if (flag != PSBA_ITER_PASS) break;
  }
  PSBA_HIP_TRY(hipStreamSynchronize(c->stream));
  return flag;
}

int psba_hybrid(psba_ctx* c, const psba_opts* o, int* itno, double* initErr, double* finalErr,
                psba_trace_rec* trace, int cap, int* len) {
  PSBA_TRY(check_ready(c));
  if (c->mixed()) return fail_unsupported_mixed("psba_hybrid");
  if (initErr) {
    PSBA_TRY(launch_residual(c, false));
    PSBA_TRY(fetch(c));
    *initErr = c->h_scal[SC_COST];
  }
  int flag;
  while (true) {
    flag = psba_levmar(c, o, itno, finalErr, trace, cap, len);
    if (flag != PSBA_ITER_TURN_TO_TR) break;
    flag = psba_trust_region(c, o, itno, finalErr, trace, cap, len);
    if (flag != PSBA_ITER_TURN_TO_LM) break;
  }
  return flag;
}

int psba_hybrid_report(psba_ctx* c, const psba_opts* o, int* itno, psba_report* out) {
  PSBA_TRY(check_ready(c));
  if (!itno || !out) return PSBA_ERR_ARG;
  std::memset(out, 0, sizeof(*out));
  const bool was_on = c->timing_on;
  PSBA_TRY(psba_enable_timing(c, 1));
  c->timing = Timing();
  std::vector<psba_trace_rec> tr(1 << 16);
  int len = 0;
  double e0 = 0, e1 = 0;
  const auto t0 = std::chrono::steady_clock::now();
  const int code = psba_hybrid(c, o, itno, &e0, &e1, tr.data(), (int)tr.size(), &len);
  const auto t1 = std::chrono::steady_clock::now();
  PSBA_TRY(psba_enable_timing(c, was_on ? 1 : 0));
  if (code < 0) return code;
  out->code = code;
  out->itno = *itno;
  out->init_sse = e0;
  out->final_sse = e1;
  out->init_error_ref = std::sqrt(e0) / c->no;
  out->final_error_ref = std::sqrt(e1) / c->no;
  out->init_rms_px = std::sqrt(e0 / c->no);
  out->final_rms_px = std::sqrt(e1 / c->no);
  out->wall_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  const int n = std::min(len, (int)tr.size());
  int prev_kind = -1, prev_it = -1;
  for (int k = 0; k < n; ++k) {
    const psba_trace_rec& r = tr[k];
    if (r.kind == 0)
      out->lm_trials++;
    else
      out->tr_steps++;
    if (r.kind != prev_kind || r.itno != prev_it) {
      if (r.kind == 0)
        out->lm_iterations++;
      else
        out->tr_iterations++;
    }
    if (prev_kind >= 0 && r.kind != prev_kind) out->handoffs++;
    prev_kind = r.kind;
    prev_it = r.itno;
  }
  out->stages.jacobian_normal = c->timing.jac;
  out->stages.schur = c->timing.schur;
  out->stages.factor = c->timing.factor;
  out->stages.solve = c->timing.solve;
  out->stages.backsub = c->timing.backsub;
  out->stages.residual = c->timing.resid;
  out->stages.other = c->timing.other;
  out->stages.trials = c->timing.trials;
  out->stages.jacobians = c->timing.jacobians;
  return code;
}

}  // extern "C"
