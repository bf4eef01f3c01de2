// Per-problem Levenberg-Marquardt stages of a batch: residual cost over fixed
// observation chunks, damping start, step statistics, the accept / reject
// decision and the commit, each problem's state held on the device.
//
// The decision restates psba_levmar with PSBA_FLAG_NO_HANDOFF
// (PSBA/levmar.cpp:45-256, compute_rho :271-280): mu0 = 1e-3 max diag
// (levmar.cpp:85-92), the |dp| tests against |p| (:141-156), the gain ratio and
// the mu update max(1/3, 1 - (2 rho - 1)^3) (:171-181), nu doubling with the
// 2^30 overflow stop (:232-244), and the cost test at the end of an iteration
// (:250-252); a relative cost decrease below rel_tol on an accepted step also
// stops (the serving convergence rule).
//
// Every reduction has a fixed order that depends only on the problem itself
// (chunks are cut per problem, blocks reduce with a fixed tree), so a problem's
// result does not depend on which other problems share the batch.
#include "batch.hpp"
#include "model.hpp"

namespace psba {

namespace {

constexpr double kInitMu = 1e-3;       // PSBA/psba.h:6
constexpr double kStopThresh = 1e-12;  // PSBA/psba.h:7
constexpr double kEps = 1e-12;         // PSBA/psba.h:8

__device__ inline void prob_ranges(const int* cam_off, const int* pt_off, int cnp, long N, int p, long* a0, long* a1,
                                   long* b0, long* b1) {
  *a0 = (long)cam_off[p] * cnp;
  *a1 = (long)cam_off[p + 1] * cnp;
  *b0 = N + 3L * pt_off[p];
  *b1 = N + 3L * pt_off[p + 1];
}

// The indices lo + threadIdx.x + s blockDim.x (s = 0, 1, ..) below hi, U at a
// time in ascending order: fn(q, cnt) gets U indices of which the first cnt
// are in range, so it can issue all of their loads before using any and
// still accumulate in the plain strided loop's order.  Whole groups pass the
// constant U (the range tests fold away); only the last group is partial.
template <int U, typename F>
__device__ inline void strided(long lo, long hi, F fn) {
  const long step = blockDim.x;
  long q0 = lo + threadIdx.x;
  for (; q0 + (U - 1) * step < hi; q0 += U * step) {
    long q[U];
#pragma unroll
    for (int u = 0; u < U; ++u) q[u] = q0 + u * step;
    fn(q, U);
  }
  if (q0 < hi) {
    long q[U];
    int cnt = 0;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      q[u] = q0 + u * step;
      cnt += q[u] < hi;
    }
    fn(q, cnt);
  }
}

}  // namespace

// sum of squared residuals over one observation chunk; with dp the parameters
// are p + dp (the same additions the commit performs)
template <int CNP>
__global__ __launch_bounds__(256) void k_b_resid_chunks(const double* __restrict__ K, const double* __restrict__ q0,
                                                        const double* __restrict__ cams,
                                                        const double* __restrict__ pts,
                                                        const double* __restrict__ dp, long N,
                                                        const double* __restrict__ meas,
                                                        const int* __restrict__ obs_cam,
                                                        const int* __restrict__ obs_pt,
                                                        const int* __restrict__ rb, const int* __restrict__ re,
                                                        const int* __restrict__ rprob,
                                                        const unsigned char* __restrict__ run,
                                                        double* __restrict__ rpart) {
  __shared__ double sh[4];
  const int b = blockIdx.x;
  if (!run[rprob[b]]) return;
  double acc[1] = {0.0};
  for (int o = rb[b] + threadIdx.x; o < re[b]; o += blockDim.x) {
    const int j = obs_cam[o], i = obs_pt[o];
    double cam[CNP], M[3], uv[2];
#pragma unroll
    for (int k = 0; k < CNP; ++k) cam[k] = dp ? cams[j * CNP + k] + dp[j * CNP + k] : cams[j * CNP + k];
#pragma unroll
    for (int k = 0; k < 3; ++k) M[k] = dp ? pts[3 * i + k] + dp[N + 3 * i + k] : pts[3 * i + k];
    project<double, CNP, false>(K + 5 * j, q0 + 4 * j, cam, M, uv, nullptr, nullptr);
    const double e0 = meas[2 * o] - uv[0], e1 = meas[2 * o + 1] - uv[1];
    acc[0] += e0 * e0 + e1 * e1;
  }
  block_sum<1>(acc, sh);
  if (threadIdx.x == 0) rpart[b] = acc[0];
}

// initial state; cost = the summed chunk partials (one thread per problem)
__global__ void k_b_init(int P, const int* __restrict__ rcp, const double* __restrict__ rpart,
                         double* __restrict__ cost, double* __restrict__ init_err, double* __restrict__ mu,
                         int* __restrict__ nu, int* __restrict__ itno, int* __restrict__ code,
                         unsigned char* __restrict__ active, unsigned char* __restrict__ need_jac,
                         unsigned char* __restrict__ first, unsigned char* __restrict__ accept, int max_iter,
                         BatchStat* __restrict__ stat) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  double s = 0.0;
  for (int q = rcp[p]; q < rcp[p + 1]; ++q) s += rpart[q];
  cost[p] = s;
  init_err[p] = s;
  mu[p] = 0.0;
  nu[p] = 2;
  itno[p] = 0;
  code[p] = PSBA_ITER_CONTINUE;
  const bool go = max_iter > 0;
  active[p] = go;
  need_jac[p] = go;
  first[p] = 1;
  accept[p] = 0;
  stat[p + 1] = BatchStat{s, s, 0, PSBA_ITER_CONTINUE, go ? 1 : 0, 0};
  if (p == 0) stat[0].active = go ? P : 0;
}

// mu0 = 1e-3 max diag(J^T J) on a problem's first iteration (one block per problem)
__global__ __launch_bounds__(256) void k_b_mu0(const double* __restrict__ uvdiag, const int* __restrict__ cam_off,
                                               const int* __restrict__ pt_off, int cnp, long N,
                                               const unsigned char* __restrict__ active,
                                               unsigned char* __restrict__ first, double* __restrict__ mu,
                                               double* __restrict__ p_L2, int* __restrict__ nu) {
  __shared__ double sh[4];
  const int p = blockIdx.x;
  if (!active[p] || !first[p]) return;
  long a0, a1, b0, b1;
  prob_ranges(cam_off, pt_off, cnp, N, p, &a0, &a1, &b0, &b1);
  double m = 0.0;
  strided<8>(a0, a1, [&](const long* q, int cnt) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = u < cnt ? uvdiag[q[u]] : 0.0;
#pragma unroll
    for (int u = 0; u < 8; ++u) m = fmax(m, v[u]);
  });
  strided<8>(b0, b1, [&](const long* q, int cnt) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = u < cnt ? uvdiag[q[u]] : 0.0;
#pragma unroll
    for (int u = 0; u < 8; ++u) m = fmax(m, v[u]);
  });
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) m = fmax(m, __shfl_xor(m, off, kWave));
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) r = fmax(r, sh[w]);
    mu[p] = kInitMu * r;
    p_L2[p] = 1e+3;
    nu[p] = 2;
    first[p] = 0;
  }
}

// |dp|^2 and sum dp (mu dp + g) per problem (one block per problem)
__global__ __launch_bounds__(256) void k_b_step_stats(const double* __restrict__ dp, const double* __restrict__ g,
                                                      const int* __restrict__ cam_off, const int* __restrict__ pt_off,
                                                      int cnp, long N, const double* __restrict__ mu,
                                                      const unsigned char* __restrict__ active,
                                                      double* __restrict__ dp2, double* __restrict__ rhoden) {
  __shared__ double sh[8];
  const int p = blockIdx.x;
  if (!active[p]) return;
  long a0, a1, b0, b1;
  prob_ranges(cam_off, pt_off, cnp, N, p, &a0, &a1, &b0, &b1);
  const double m = mu[p];
  double acc[2] = {0.0, 0.0};
  auto part = [&](const long* q, int cnt) {
    double d[8], gq[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      d[u] = u < cnt ? dp[q[u]] : 0.0;
      gq[u] = u < cnt ? g[q[u]] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (u < cnt) {
        acc[0] += d[u] * d[u];
        acc[1] += d[u] * (m * d[u] + gq[u]);
      }
  };
  strided<8>(a0, a1, part);
  strided<8>(b0, b1, part);
  block_sum<2>(acc, sh);
  if (threadIdx.x == 0) {
    dp2[p] = acc[0];
    rhoden[p] = acc[1];
  }
}

// the accept / reject decision of one trial (one thread per problem)
__global__ void k_b_control(int P, const int* __restrict__ rcp, const double* __restrict__ rpart,
                            const double* __restrict__ fok, const double* __restrict__ dp2,
                            const double* __restrict__ rhoden, const double* __restrict__ p_L2,
                            double* __restrict__ cost, double* __restrict__ tcost, double* __restrict__ mu,
                            int* __restrict__ nu, int* __restrict__ itno, int* __restrict__ code,
                            unsigned char* __restrict__ active, unsigned char* __restrict__ need_jac,
                            unsigned char* __restrict__ accept, int max_iter, double rel_tol,
                            const double* __restrict__ init_err, BatchStat* __restrict__ stat) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  accept[p] = 0;
  need_jac[p] = 0;
  if (!active[p]) return;
  double tc = 0.0;
  for (int q = rcp[p]; q < rcp[p + 1]; ++q) tc += rpart[q];
  tcost[p] = tc;
  int flag = PSBA_ITER_CONTINUE;
  bool end_iter = false;
  if (fok[p] != 0.0) {
    const double d2 = dp2[p], pl2 = p_L2[p];
    if (d2 < pl2 * kStopThresh * kStopThresh) {
      flag = PSBA_ITER_DP_NO_CHANGE;
      end_iter = true;
    } else if (d2 >= (pl2 + kStopThresh) / (kEps * kEps)) {
      flag = PSBA_ITER_ERR;
      end_iter = true;
    } else {
      const double c0 = cost[p];
      const double rho = (c0 - tc) / rhoden[p];
      if (rho > 0) {
        double tmp = 2 * rho - 1;
        tmp = 1.0 - tmp * tmp * tmp;
        mu[p] = mu[p] * ((tmp >= (1.0 / 3.0)) ? tmp : (1.0 / 3.0));
        nu[p] = 2;
        accept[p] = 1;
        const double rel_dec = (c0 - tc) / c0;
        cost[p] = tc;
        if (rel_tol > 0 && rel_dec < rel_tol) flag = PSBA_ITER_DP_NO_CHANGE;
        end_iter = true;
      }
    }
  }
  if (!end_iter) {
    mu[p] = mu[p] * nu[p];
    if (nu[p] >= (1 << 30)) {
      flag = PSBA_ITER_ERR;
      end_iter = true;
    } else {
      nu[p] = 2 * nu[p];
    }
  }
  bool still = true;
  if (end_iter) {
    if (cost[p] <= kStopThresh) flag = PSBA_ITER_ERR_SMALL_ENOUGH;
    const int it = itno[p] + 1;
    itno[p] = it;
    if (flag != PSBA_ITER_CONTINUE || it >= max_iter) {
      code[p] = flag;
      active[p] = 0;
      still = false;
    } else {
      need_jac[p] = 1;
    }
  }
  stat[p + 1] = BatchStat{cost[p], init_err[p], itno[p], still ? PSBA_ITER_CONTINUE : code[p], still ? 1 : 0, 0};
  if (still) atomicAdd(&stat[0].active, 1);
}

// p += dp on accepted problems; p_L2 = |p|^2 for the next |dp| test
// (one block per problem)
__global__ __launch_bounds__(256) void k_b_commit(double* __restrict__ cams, double* __restrict__ pts,
                                                  const double* __restrict__ dp, const int* __restrict__ cam_off,
                                                  const int* __restrict__ pt_off, int cnp, long N,
                                                  const unsigned char* __restrict__ accept,
                                                  double* __restrict__ p_L2) {
  __shared__ double sh[4];
  const int p = blockIdx.x;
  if (!accept[p]) return;
  long a0, a1, b0, b1;
  prob_ranges(cam_off, pt_off, cnp, N, p, &a0, &a1, &b0, &b1);
  double acc[1] = {0.0};
  // camera parameters (q < N) live in cams, point coordinates in pts
  auto part = [&](double* base, long shift, const long* q, int cnt) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = u < cnt ? base[q[u] - shift] + dp[q[u]] : 0.0;
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (u < cnt) {
        base[q[u] - shift] = v[u];
        acc[0] += v[u] * v[u];
      }
  };
  strided<8>(a0, a1, [&](const long* q, int cnt) { part(cams, 0, q, cnt); });
  strided<8>(b0, b1, [&](const long* q, int cnt) { part(pts, N, q, cnt); });
  block_sum<1>(acc, sh);
  if (threadIdx.x == 0) p_L2[p] = acc[0];
}

__global__ void k_b_zero(BatchStat* st) { st->active = 0; }

template <int CNP>
static void run_b_resid(psba_batch* b, bool trial) {
  hipLaunchKernelGGL(k_b_resid_chunks<CNP>, dim3(b->n_rchunk), dim3(256), 0, b->stream, b->K, b->q0, b->cams, b->pts,
                     trial ? b->dp : nullptr, b->N, b->meas, b->obs_cam, b->obs_pt, b->rchunk_b, b->rchunk_e,
                     b->rchunk_prob, trial ? b->active : b->all_on, b->rpart);
}

int launch_b_lm(psba_batch* b, int stage) {
  const unsigned pb = ceil_div(b->P, 64);
  switch (stage) {
    case BLM_INIT:
      if (b->n_rchunk > 0) PSBA_DISPATCH_CNP(b->cnp, run_b_resid, b, false);
      hipLaunchKernelGGL(k_b_init, dim3(pb), dim3(64), 0, b->stream, b->P, b->rcp, b->rpart, b->cost, b->init_err,
                         b->mu, b->nu, b->itno, b->code, b->active, b->need_jac, b->first, b->accept, b->max_iter,
                         b->stat);
      break;
    case BLM_MU0:
      hipLaunchKernelGGL(k_b_mu0, dim3(b->P), dim3(256), 0, b->stream, b->uvdiag, b->cam_off, b->pt_off, b->cnp, b->N,
                         b->active, b->first, b->mu, b->p_L2, b->nu);
      break;
    case BLM_STEP_STATS:
      hipLaunchKernelGGL(k_b_step_stats, dim3(b->P), dim3(256), 0, b->stream, b->dp, b->g, b->cam_off, b->pt_off,
                         b->cnp, b->N, b->mu, b->active, b->dp2, b->rhoden);
      break;
    case BLM_TRIAL_COST:
      if (b->n_rchunk > 0) PSBA_DISPATCH_CNP(b->cnp, run_b_resid, b, true);
      break;
    case BLM_CONTROL:
      hipLaunchKernelGGL(k_b_zero, dim3(1), dim3(1), 0, b->stream, b->stat);
      hipLaunchKernelGGL(k_b_control, dim3(pb), dim3(64), 0, b->stream, b->P, b->rcp, b->rpart, b->fok, b->dp2,
                         b->rhoden, b->p_L2, b->cost, b->tcost, b->mu, b->nu, b->itno, b->code, b->active,
                         b->need_jac, b->accept, b->max_iter, b->rel_tol, b->init_err, b->stat);
      break;
    case BLM_COMMIT:
      hipLaunchKernelGGL(k_b_commit, dim3(b->P), dim3(256), 0, b->stream, b->cams, b->pts, b->dp, b->cam_off,
                         b->pt_off, b->cnp, b->N, b->accept, b->p_L2);
      break;
    default:
      return PSBA_ERR_ARG;
  }
  PSBA_LAUNCH_TRY();
  return PSBA_OK;
}

}  // namespace psba
