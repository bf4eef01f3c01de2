// Observation-, point- and camera-level kernels of the plain FP64 path:
// residual + cost, Jacobian + normal-equation blocks (U, V, W, g), damped
// point-block inverses, back-substitution, step statistics and commit.
//
// Reference stages: PSBA/CL_files/compute_exQT.cl (residual),
// compute_jacobiQT.cl (Jacobian), compute_U.cl, compute_V.cl,
// compute_Wblks.cl, compute_g.cl (normal equations), update_UV.cl +
// compute_Vinv.cl (damped V^-1), compute_eb.cl + compute_dpb.cl
// (back-substitution), compute_newp.cl + update_p.cl (commit),
// compute_Jmultiply.cl (J x for the dogleg's model decrease).
#include <algorithm>

#include "context.hpp"
#include "model.hpp"

namespace psba {

constexpr int kJacBlock = 64;  // threads per Jacobian workgroup (context launch; 64, 128 and 256 measured)

// ---------------------------------------------------------------------------
// residual and cost
// ---------------------------------------------------------------------------
// One thread per observation.  With dp != nullptr the parameters are p + dp
// (the trial point), computed with the same additions the commit performs.
// The block's work as a routine of its block index `blk` (k_residual, and
// k_trial_eval below, which runs it beside the step statistics).
template <int CNP>
__device__ __forceinline__ void residual_block(const double* __restrict__ K, const double* __restrict__ q0,
                                               const double* __restrict__ cams, const double* __restrict__ pts,
                                               const double* __restrict__ dp, int n, const double* __restrict__ meas,
                                               const int* __restrict__ obs_cam, const int* __restrict__ obs_pt,
                                               int no, double* __restrict__ e_out, double* partials, int loss,
                                               double loss_d, int blk, double* sh) {
  double acc[1] = {0.0};
  int o = blk * blockDim.x + threadIdx.x;
  if (o < no) {
    int j = obs_cam[o], i = obs_pt[o];
    double cam[CNP], M[3], uv[2];
#pragma unroll
    for (int k = 0; k < CNP; ++k) cam[k] = dp ? cams[j * CNP + k] + dp[j * CNP + k] : cams[j * CNP + k];
#pragma unroll
    for (int k = 0; k < 3; ++k) M[k] = dp ? pts[3 * i + k] + dp[n + 3 * i + k] : pts[3 * i + k];
    project<double, CNP, false>(K + 5 * j, q0 + 4 * j, cam, M, uv, nullptr, nullptr);
    double e0 = meas[2 * o] - uv[0], e1 = meas[2 * o + 1] - uv[1];
    if (e_out) {
      e_out[2 * o] = e0;
      e_out[2 * o + 1] = e1;
    }
    acc[0] = e0 * e0 + e1 * e1;
    if (loss) {
      double drho;
      loss_eval(loss, loss_d, acc[0], acc[0], drho);
    }
  }
  grid_sum_partials<1>(acc, partials, sh, blk);
}

template <int CNP>
__global__ __launch_bounds__(256) void k_residual(const double* __restrict__ K, const double* __restrict__ q0,
                                                  const double* __restrict__ cams, const double* __restrict__ pts,
                                                  const double* __restrict__ dp, int n,
                                                  const double* __restrict__ meas, const int* __restrict__ obs_cam,
                                                  const int* __restrict__ obs_pt, int no, double* __restrict__ e_out,
                                                  double* partials, int loss, double loss_d) {
  __shared__ double sh[4];
  residual_block<CNP>(K, q0, cams, pts, dp, n, meas, obs_cam, obs_pt, no, e_out, partials, loss, loss_d,
                      blockIdx.x, sh);
}

// |dp|^2 and dp.(mu dp + g) over 256 parameters per block (block `blk`)
__device__ __forceinline__ void step_stats_block(const double* __restrict__ dp, const double* __restrict__ g,
                                                 long ntot, double mu, const double* __restrict__ mu_p,
                                                 double* partials, long blk, double* sh) {
  double acc[2] = {0.0, 0.0};
  if (mu_p) mu = *mu_p;
  long q = blk * blockDim.x + threadIdx.x;
  if (q < ntot) {
    double d = dp[q];
    acc[0] = d * d;
    acc[1] = d * (mu * d + g[q]);
  }
  grid_sum_partials<2>(acc, partials, sh, blk);
}

// A device-controlled trial's step statistics and trial cost in one launch:
// the last nb_step blocks are k_step_stats' blocks, the others k_residual's
// (the same per-block arithmetic, so the same partials), each set of
// partials where the LM control kernel sums it.
template <int CNP>
__global__ __launch_bounds__(256) void k_trial_eval(const double* __restrict__ g, long ntot,
                                                    const double* __restrict__ mu_p, double* step_part, int nb_step,
                                                    const double* __restrict__ K, const double* __restrict__ q0,
                                                    const double* __restrict__ cams, const double* __restrict__ pts,
                                                    const double* __restrict__ dp, int n,
                                                    const double* __restrict__ meas, const int* __restrict__ obs_cam,
                                                    const int* __restrict__ obs_pt, int no, double* cost_part,
                                                    int loss, double loss_d) {
  __shared__ double sh[8];
  // the residual blocks (the longer ones) first, then the step statistics
  const int b = blockIdx.x, nb_cost = (int)gridDim.x - nb_step;
  if (b < nb_cost) {
    residual_block<CNP>(K, q0, cams, pts, dp, n, meas, obs_cam, obs_pt, no, nullptr, cost_part, loss, loss_d, b, sh);
  } else {
    step_stats_block(dp, g, ntot, 0.0, mu_p, step_part, b - nb_cost, sh);
  }
}

template <int CNP>
static void run_trial_eval(psba_ctx* c) {
  const unsigned nb_step = ceil_div(c->ntot, 256), nb_cost = ceil_div(c->no, 256);
  hipLaunchKernelGGL(k_trial_eval<CNP>, dim3(nb_step + nb_cost), dim3(256), 0, c->stream, c->g, (long)c->ntot,
                     mu_of(c), c->red_part, (int)nb_step, c->K, c->q0, c->cams, c->pts, c->dp, c->n, c->meas,
                     c->obs_cam, c->obs_pt, c->no, c->red_part + trial_cost_part_off(c), c->loss, c->loss_scale);
}

int launch_trial_eval(psba_ctx* c) {
  if (!c->devctl) return PSBA_ERR_ARG;
  PSBA_DISPATCH_CNP(c->cnp, run_trial_eval, c);
  PSBA_LAUNCH_TRY();
  return PSBA_OK;
}

template <int CNP>
static void run_residual(psba_ctx* c, bool trial) {
  unsigned nb = ceil_div(c->no, 256);
  // (a trial under device control takes k_trial_eval instead)
  hipLaunchKernelGGL(k_residual<CNP>, dim3(nb), dim3(256), 0, c->stream, c->K, c->q0, c->cams, c->pts,
                     trial ? c->dp : nullptr, c->n, c->meas, c->obs_cam, c->obs_pt, c->no,
                     trial ? nullptr : c->e, c->red_part, c->loss, c->loss_scale);
  launch_sum_partials<1>(c->stream, c->red_part, (int)nb, c->d_scal + SC_COST);
}

int launch_residual(psba_ctx* c, bool trial) {
  PSBA_DISPATCH_CNP(c->cnp, run_residual, c, trial);
  PSBA_LAUNCH_TRY();
  return PSBA_OK;
}

// ---------------------------------------------------------------------------
// Jacobian, W and residual per observation
// ---------------------------------------------------------------------------
// R is the storage (and accumulation) type of the normal-equation stages:
// double on the plain path, float on the mixed path; the projection and its
// derivatives are always evaluated in FP64.

// One observation's rows in FP64: the residual e = m - f(p) and the
// Jacobian blocks (fixed parameters' columns zeroed), reweighted by
// sqrt(rho'(|e|^2)) under a robust loss.  k_jacobian and the camera chunks
// that form U and g_a from it (without reading J_A back) both take the rows
// from here, so they get the same values.
template <int CNP>
__device__ __forceinline__ void jac_row(int o, int j, int i, const double* __restrict__ K,
                                        const double* __restrict__ q0, const double* __restrict__ cams,
                                        const double* __restrict__ pts, const double* __restrict__ meas,
                                        const unsigned char* __restrict__ fix, int loss, double loss_d,
                                        double (&ja)[2 * CNP], double (&jb)[6], double& e0, double& e1) {
  double cam[CNP], M[3], uv[2];
#pragma unroll
  for (int k = 0; k < CNP; ++k) cam[k] = cams[j * CNP + k];
#pragma unroll
  for (int k = 0; k < 3; ++k) M[k] = pts[3 * i + k];
  project<double, CNP, true>(K + 5 * j, q0 + 4 * j, cam, M, uv, ja, jb);
#pragma unroll
  for (int k = 0; k < CNP; ++k)
    if (fix[j * CNP + k]) ja[k] = ja[CNP + k] = 0.0;
  e0 = meas[2 * o] - uv[0];
  e1 = meas[2 * o + 1] - uv[1];
  if (loss) {
    // reweighted rows: residual and Jacobian scaled by sqrt(rho'(s))
    double rho, drho;
    loss_eval(loss, loss_d, e0 * e0 + e1 * e1, rho, drho);
    const double w = sqrt(drho);
#pragma unroll
    for (int k = 0; k < 2 * CNP; ++k) ja[k] *= w;
#pragma unroll
    for (int k = 0; k < 6; ++k) jb[k] *= w;
    e0 *= w;
    e1 *= w;
  }
}

// What jac_row reads, for kernels that recompute rows
struct JacModel {
  const double *K, *q0, *cams, *pts, *meas;
  const int* obs_cam;
  const unsigned char* fix;
  int loss;
  double loss_d;
};

// With write_ja false the J_A rows are not stored (the device-controlled LM
// forms U and g_a from recomputed rows, k_normal_chunks with RC).
template <int CNP, typename R>
__global__ __launch_bounds__(256) void k_jacobian(const double* __restrict__ K, const double* __restrict__ q0,
                                                  const double* __restrict__ cams, const double* __restrict__ pts,
                                                  const double* __restrict__ meas, const int* __restrict__ obs_cam,
                                                  const int* __restrict__ obs_pt, const unsigned char* __restrict__ fix,
                                                  int no, double coeff, R* __restrict__ e, R* __restrict__ JA,
                                                  R* __restrict__ JB, R* __restrict__ W,
                                                  const int* __restrict__ ownr, const unsigned char* __restrict__ run,
                                                  int loss, double loss_d, int write_ja) {
  // The workgroup's rows of e, J_A, J_B and W go out through LDS: each thread
  // puts its row in the staging buffer (stride L + 1 against bank
  // conflicts), then the workgroup writes the block of rows as one contiguous
  // range, so every store instruction covers consecutive addresses instead of
  // one element in each of 64 rows.  Rows of batched problems that do not
  // need a Jacobian (run) are left as they are.
  extern __shared__ double sJac[];
  // one problem (no owner map): run[0] gates the whole launch
  if (run && !ownr && !run[0]) return;
  R* const sb = (R*)sJac;
  unsigned char* const sAct = (unsigned char*)(sJac) + (size_t)blockDim.x * (3 * CNP + 1) * sizeof(R);
  const int o0 = blockIdx.x * blockDim.x, nrow = min((int)blockDim.x, no - o0);
  const int o = o0 + threadIdx.x;
  const bool act = o < no && (!run || run[ownr ? ownr[o] : 0]);
  sAct[threadIdx.x] = act;
  auto put = [&](int L, const R* v) {
    if (act)
      for (int k = 0; k < L; ++k) sb[threadIdx.x * (L + 1) + k] = v[k];
  };
  auto flush = [&](int L, R* dst) {
    __syncthreads();
    for (int f = threadIdx.x; f < nrow * L; f += blockDim.x) {
      const int rr = f / L;
      if (sAct[rr]) dst[(size_t)o0 * L + f] = sb[rr * (L + 1) + (f - rr * L)];
    }
    __syncthreads();
  };
  // every thread reaches the same barriers (a wave holds active and inactive
  // rows at once); only active threads compute and stage a row
  R jr[2 * CNP], br[6], er[2], wr[3 * CNP];
  if (act) {
    double ja[2 * CNP], jb[6], e0, e1;
    jac_row<CNP>(o, obs_cam[o], obs_pt[o], K, q0, cams, pts, meas, fix, loss, loss_d, ja, jb, e0, e1);
    er[0] = (R)e0;
    er[1] = (R)e1;
#pragma unroll
    for (int k = 0; k < 2 * CNP; ++k) jr[k] = (R)ja[k];
#pragma unroll
    for (int k = 0; k < 6; ++k) br[k] = (R)jb[k];
#pragma unroll
    for (int r = 0; r < CNP; ++r)
#pragma unroll
      for (int cc = 0; cc < 3; ++cc) wr[3 * r + cc] = (R)coeff * (jr[r] * br[cc] + jr[CNP + r] * br[3 + cc]);
  }
  put(2, er);
  flush(2, e);
  if (write_ja) {
    put(2 * CNP, jr);
    flush(2 * CNP, JA);
  }
  put(6, br);
  flush(6, JB);
  put(3 * CNP, wr);
  flush(3 * CNP, W);
}

// dynamic LDS of k_jacobian for a block of bs threads
template <int CNP, typename R>
constexpr size_t jacobian_lds(int bs) {
  return (size_t)bs * (3 * CNP + 1) * sizeof(R) + (size_t)bs;
}

// V_i (6 unique entries), g_b,i and the V part of UVdiag, one thread per point
// over its observations in ascending camera order (workgroup `blk`).
template <typename R>
__device__ __forceinline__ void point_normal_block(int blk, const int* __restrict__ pt_ptr,
                                                   const R* __restrict__ JB, const R* __restrict__ e, int np,
                                                   double coeff, double coeff_g, double* __restrict__ V,
                                                   double* __restrict__ gb, double* __restrict__ uvd,
                                                   const int* __restrict__ ownr,
                                                   const unsigned char* __restrict__ run) {
  if (run && !ownr && !run[0]) return;  // one problem, gated off
  // a quad per point: lane p sums a quarter of the point's observations (a
  // long track is split four ways), the quad adds the quarters in a fixed order
  const long gt = (long)blk * blockDim.x + threadIdx.x;
  const int i = (int)(gt >> 2), p = (int)(gt & 3);
  const bool ok = i < np && (!run || run[ownr ? ownr[i] : 0]);
  R v[6] = {0, 0, 0, 0, 0, 0}, g[3] = {0, 0, 0};
  if (ok) {
    const int beg = pt_ptr[i], cnt = pt_ptr[i + 1] - beg;
    const int o0 = beg + ((cnt * p) >> 2), o1 = beg + ((cnt * (p + 1)) >> 2);
#pragma unroll 4
    for (int o = o0; o < o1; ++o) {
      const R* b = JB + (size_t)o * 6;
      R b0 = b[0], b1 = b[1], b2 = b[2], b3 = b[3], b4 = b[4], b5 = b[5];
      R e0 = e[2 * o], e1 = e[2 * o + 1];
      v[0] += b0 * b0 + b3 * b3;
      v[1] += b0 * b1 + b3 * b4;
      v[2] += b0 * b2 + b3 * b5;
      v[3] += b1 * b1 + b4 * b4;
      v[4] += b1 * b2 + b4 * b5;
      v[5] += b2 * b2 + b5 * b5;
      g[0] += b0 * e0 + b3 * e1;
      g[1] += b1 * e0 + b4 * e1;
      g[2] += b2 * e0 + b5 * e1;
    }
  }
#pragma unroll
  for (int k = 0; k < 6; ++k) v[k] = quad_sum(v[k]);
#pragma unroll
  for (int k = 0; k < 3; ++k) g[k] = quad_sum(g[k]);
  if (!ok || p != 0) return;
#pragma unroll
  for (int k = 0; k < 6; ++k) V[(size_t)i * 6 + k] = coeff * (double)v[k];
#pragma unroll
  for (int k = 0; k < 3; ++k) gb[(size_t)i * 3 + k] = coeff_g * (double)g[k];
  uvd[(size_t)i * 3 + 0] = coeff * (double)v[0];
  uvd[(size_t)i * 3 + 1] = coeff * (double)v[3];
  uvd[(size_t)i * 3 + 2] = coeff * (double)v[5];
}

template <typename R>
__global__ __launch_bounds__(256) void k_point_normal(const int* __restrict__ pt_ptr, const R* __restrict__ JB,
                                                      const R* __restrict__ e, int np, double coeff,
                                                      double coeff_g, double* __restrict__ V,
                                                      double* __restrict__ gb, double* __restrict__ uvd,
                                                      const int* __restrict__ ownr,
                                                      const unsigned char* __restrict__ run) {
  point_normal_block<R>(blockIdx.x, pt_ptr, JB, e, np, coeff, coeff_g, V, gb, uvd, ownr, run);
}

// Camera-segmented sums of the U_j upper triangle and g_a,j: stage 1.  Each
// block takes one chunk (<= kCamChunk observations of one camera, ascending
// point), stages JA and e in LDS, and thread t sums entry t over the chunk.
template <int CNP>
constexpr int cam_chunk_obs() { return (CNP > 11) ? 128 : kCamChunk; }

// chunk b, staging in sJ (cam_chunk_obs() (2 CNP + 2) values of LDS).
// With RC (one problem, rec given) the rows are recomputed by jac_row from
// the parameters in `mdl` rather than read from J_A and e: the same values
// k_jacobian stores, rounded to R the same way.
template <int CNP, typename R, bool RC = false>
__device__ __forceinline__ void cam_chunk(int b, const int* __restrict__ cam_obs, const int* __restrict__ cb,
                                          const int* __restrict__ ce, const R* __restrict__ JA,
                                          const R* __restrict__ e, R* __restrict__ part,
                                          const int* __restrict__ ownr, const unsigned char* __restrict__ run,
                                          const int2* __restrict__ rec, R* sJ, const JacModel mdl = JacModel{}) {
  constexpr int NU = CNP * (CNP + 1) / 2, NE = NU + CNP, REC = 2 * CNP + 2;
  constexpr int CH = cam_chunk_obs<CNP>();
  static_assert(CH <= 256, "one observation per thread");
  // with rec (one problem) the observation comes from the chunk's record at
  // an address the workgroup index gives, loaded beside the chunk's extent
  // and the gate
  const int2 r_rec = (rec && (int)threadIdx.x < CH) ? rec[(size_t)b * CH + threadIdx.x] : make_int2(-1, -1);
  const int o_rec = r_rec.x;
  const int beg = cb[b], cnt = ce[b] - beg;
  if (gated_off(run, ownr, b)) return;
  // one thread per observation stages its Jacobian rows and residual (one
  // round of dependent loads for the whole chunk, all of a record's loads in
  // flight together)
  for (int s = threadIdx.x; s < cnt; s += blockDim.x) {
    const int o = rec ? o_rec : cam_obs[beg + s];
    R* dst = sJ + s * REC;
    if constexpr (RC) {
      double ja[2 * CNP], jb[6], e0, e1;
      jac_row<CNP>(o, mdl.obs_cam[o], r_rec.y, mdl.K, mdl.q0, mdl.cams, mdl.pts, mdl.meas, mdl.fix, mdl.loss,
                   mdl.loss_d, ja, jb, e0, e1);
#pragma unroll
      for (int el = 0; el < 2 * CNP; ++el) dst[el] = (R)ja[el];
      dst[2 * CNP] = (R)e0;
      dst[2 * CNP + 1] = (R)e1;
    } else {
      const R* ja = JA + (size_t)o * 2 * CNP;
#pragma unroll
      for (int el = 0; el < 2 * CNP; ++el) dst[el] = ja[el];
      dst[2 * CNP] = e[2 * o];
      dst[2 * CNP + 1] = e[2 * o + 1];
    }
  }
  __syncthreads();
  // a quad per output: lane p of the quad sums observations p, p + 4, ..
  // (neighbouring records: no LDS bank conflicts inside the quad), the quad
  // adds the four in a fixed order
  const int p = threadIdx.x & 3;
  for (int base = 0; base < NE; base += (int)blockDim.x >> 2) {
    const int t = base + ((int)threadIdx.x >> 2);
    R acc = 0;
    if (t < NU) {
      int r = 0, rem = t;
      while (rem >= CNP - r) {
        rem -= CNP - r;
        ++r;
      }
      int cc = r + rem;
      for (int s = p; s < cnt; s += 4) {
        const R* a = sJ + s * REC;
        acc += a[r] * a[cc] + a[CNP + r] * a[CNP + cc];
      }
    } else if (t < NE) {
      int r = t - NU;
      for (int s = p; s < cnt; s += 4) {
        const R* a = sJ + s * REC;
        acc += a[r] * a[2 * CNP] + a[CNP + r] * a[2 * CNP + 1];
      }
    }
    acc = quad_sum(acc);
    if (t < NE && p == 0) part[(size_t)b * NE + t] = acc;
  }
}

template <int CNP, typename R>
__global__ __launch_bounds__(256) void k_cam_chunks(const int* __restrict__ cam_obs, const int* __restrict__ cb,
                                                    const int* __restrict__ ce, const R* __restrict__ JA,
                                                    const R* __restrict__ e, R* __restrict__ part,
                                                    const int* __restrict__ ownr,
                                                    const unsigned char* __restrict__ run,
                                                    const int2* __restrict__ rec) {
  __shared__ R sJ[cam_chunk_obs<CNP>() * (2 * CNP + 2)];
  cam_chunk<CNP, R>(blockIdx.x, cam_obs, cb, ce, JA, e, part, ownr, run, rec, sJ);
}

// One problem's point normal equations and camera chunks in one launch (both
// read only the Jacobian stage's rows): the last nb_pt workgroups are
// k_point_normal's, the others k_cam_chunks' (the same per-workgroup work;
// with RC the camera chunks recompute their J_A rows, which k_jacobian then
// does not store).
template <int CNP, typename R, bool RC>
__global__ __launch_bounds__(256) void k_normal_chunks(const int* __restrict__ pt_ptr, const R* __restrict__ JB,
                                                       const R* __restrict__ e, int np, double coeff, double coeff_g,
                                                       double* __restrict__ V, double* __restrict__ gb,
                                                       double* __restrict__ uvd_b, int nb_pt,
                                                       const int* __restrict__ cb, const int* __restrict__ ce,
                                                       const R* __restrict__ JA, R* __restrict__ part,
                                                       const int2* __restrict__ rec,
                                                       const unsigned char* __restrict__ run, const JacModel mdl) {
  __shared__ R sJ[cam_chunk_obs<CNP>() * (2 * CNP + 2)];
  // the camera chunks first (the longer workgroups), then the points
  const int ncc = (int)gridDim.x - nb_pt;
  if ((int)blockIdx.x < ncc)
    cam_chunk<CNP, R, RC>((int)blockIdx.x, nullptr, cb, ce, JA, e, part, nullptr, run, rec, sJ, mdl);
  else
    point_normal_block<R>((int)blockIdx.x - ncc, pt_ptr, JB, e, np, coeff, coeff_g, V, gb, uvd_b, nullptr, run);
}

// stage 2: per camera, sum its chunk partials in chunk order (camera j).
template <int CNP, typename R>
__device__ __forceinline__ void cam_finish_block(int j, const int* __restrict__ ccp, const R* __restrict__ part,
                                                 int nc, double coeff, double coeff_g, double* __restrict__ U,
                                                 double* __restrict__ ga, double* __restrict__ uvd,
                                                 const int* __restrict__ ownr,
                                                 const unsigned char* __restrict__ run) {
  constexpr int NU = CNP * (CNP + 1) / 2, NE = NU + CNP;
  if (run && !run[ownr ? ownr[j] : 0]) return;
  for (int t = threadIdx.x; t < NE; t += blockDim.x) {
    const double acc = ordered_sum<16>(ccp[j], ccp[j + 1], [&](int q) { return (double)part[(size_t)q * NE + t]; });
    if (t < NU) {
      int r = 0, rem = t;
      while (rem >= CNP - r) {
        rem -= CNP - r;
        ++r;
      }
      int cc = r + rem;
      double val = coeff * acc;
      U[(size_t)j * CNP * CNP + r * CNP + cc] = val;
      U[(size_t)j * CNP * CNP + cc * CNP + r] = val;
      if (r == cc) uvd[j * CNP + r] = val;
    } else {
      ga[j * CNP + (t - NU)] = coeff_g * acc;
    }
  }
}

template <int CNP, typename R>
__global__ __launch_bounds__(256) void k_cam_finish(const int* __restrict__ ccp, const R* __restrict__ part,
                                                    int nc, double coeff, double coeff_g, double* __restrict__ U,
                                                    double* __restrict__ ga, double* __restrict__ uvd,
                                                    const int* __restrict__ ownr,
                                                    const unsigned char* __restrict__ run) {
  cam_finish_block<CNP, R>(blockIdx.x, ccp, part, nc, coeff, coeff_g, U, ga, uvd, ownr, run);
}

// ---------------------------------------------------------------------------
// damped point-block inverse and z = V^-1 g_b
// ---------------------------------------------------------------------------
// Regular branch: adjugate over T = -det (PSBA/CL_files/compute_Vinv.cl:29,76-86).
// |T| < 1e-16: the reference's pivoted-elimination divisor (compute_Vinv.cl:31-73).
// Batched problems pass mu_p (damping per problem) with ownr (point -> problem)
// and run (problem -> still iterating); the single-problem path passes mu.
__device__ __forceinline__ void vinv_one(int i, const double* __restrict__ V, const double* __restrict__ gb, int np,
                                         double mu, double* __restrict__ Vinv, double* __restrict__ z,
                                         const double* __restrict__ mu_p, const int* __restrict__ ownr,
                                         const unsigned char* __restrict__ run) {
  if (i >= np) return;
  if (mu_p) {
    const int p = ownr ? ownr[i] : 0;
    if (run && !run[p]) return;
    mu = mu_p[p];
  }
  const double* v = V + (size_t)i * 6;
  double a11 = v[0] + mu, a12 = v[1], a13 = v[2], a22 = v[3] + mu, a23 = v[4], a33 = v[5] + mu;
  double T = a33 * a12 * a12 - 2 * a12 * a13 * a23 + a22 * a13 * a13 + a11 * a23 * a23 - a11 * a22 * a33;
  double c00 = a22 * a33 - a23 * a23, c01 = a13 * a23 - a12 * a33, c02 = a12 * a23 - a13 * a22;
  double c11 = a11 * a33 - a13 * a13, c12 = a12 * a13 - a11 * a23, c22 = a11 * a22 - a12 * a12;
  double d;
  if (fabs(T) < 1e-16) {
    double m[3][3] = {{a11, a12, a13}, {a12, a22, a23}, {a13, a23, a33}};
    int p = 0;
    if (m[0][0] < m[1][0]) p = 1;
    if (m[p][0] < m[2][0]) p = 2;
    if (p != 0)
      for (int q = 0; q < 3; ++q) {
        double t = m[0][q]; m[0][q] = m[p][q]; m[p][q] = t;
      }
    m[1][0] /= m[0][0];
    m[2][0] /= m[0][0];
    m[1][1] -= m[1][0] * m[0][1];
    m[1][2] -= m[1][0] * m[0][2];
    m[2][1] -= m[2][0] * m[0][1];
    m[2][2] -= m[2][0] * m[0][2];
    if (m[1][1] < m[2][1])
      for (int q = 0; q < 3; ++q) {
        double t = m[1][q]; m[1][q] = m[2][q]; m[2][q] = t;
      }
    if (m[1][1] != 0.0) {
      m[2][1] /= m[1][1];
      m[2][2] -= m[2][1] * m[1][2];
    }
    d = m[0][0] * m[1][1] * m[2][2];
  } else {
    d = -T;
  }
  double w[6] = {c00 / d, c01 / d, c02 / d, c11 / d, c12 / d, c22 / d};
  double* out = Vinv + (size_t)i * 6;
#pragma unroll
  for (int k = 0; k < 6; ++k) out[k] = w[k];
  double g0 = gb[3 * i], g1 = gb[3 * i + 1], g2 = gb[3 * i + 2];
  z[3 * i + 0] = w[0] * g0 + w[1] * g1 + w[2] * g2;
  z[3 * i + 1] = w[1] * g0 + w[3] * g1 + w[4] * g2;
  z[3 * i + 2] = w[2] * g0 + w[4] * g1 + w[5] * g2;
}

__global__ __launch_bounds__(256) void k_vinv(const double* __restrict__ V, const double* __restrict__ gb, int np,
                                              double mu, double* __restrict__ Vinv, double* __restrict__ z,
                                              const double* __restrict__ mu_p, const int* __restrict__ ownr,
                                              const unsigned char* __restrict__ run) {
  vinv_one(blockIdx.x * blockDim.x + threadIdx.x, V, gb, np, mu, Vinv, z, mu_p, ownr, run);
}

// One problem under device control: k_cam_finish (gated like the Jacobian
// stage) and k_vinv (every trial, the damping from the LM state) in one
// launch, workgroups [0, nc) the cameras, the rest 256 points each.  (The
// first trial of a solve keeps them apart: mu0 needs U before V^-1.)
template <int CNP, typename R>
__global__ __launch_bounds__(256) void k_finish_vinv(const int* __restrict__ ccp, const R* __restrict__ part, int nc,
                                                     double coeff, double coeff_g, double* __restrict__ U,
                                                     double* __restrict__ ga, double* __restrict__ uvd,
                                                     const unsigned char* __restrict__ run,
                                                     const double* __restrict__ V, const double* __restrict__ gb,
                                                     int np, double* __restrict__ Vinv, double* __restrict__ z,
                                                     const double* __restrict__ mu_p) {
  if ((int)blockIdx.x < nc)
    cam_finish_block<CNP, R>(blockIdx.x, ccp, part, nc, coeff, coeff_g, U, ga, uvd, nullptr, run);
  else
    vinv_one(((int)blockIdx.x - nc) * blockDim.x + threadIdx.x, V, gb, np, 0.0, Vinv, z, mu_p, nullptr, nullptr);
}

template <int CNP, typename R>
static void run_jac_normal_t(psba_ctx* c, double coeff, double coeff_g, R* e, R* JA, R* JB, R* W, R* ucpart,
                             bool with_vinv) {
  const unsigned char* gate = gate_of(c, kGateJac);
  // under device control (the LM's trials) the camera chunks recompute the
  // J_A rows, which are then not stored: 2 C values per observation neither
  // written nor read back
  const bool rc = c->devctl;
  probe_begin(c, PSBA_PROBE_JACOBIAN);
  hipLaunchKernelGGL((k_jacobian<CNP, R>), dim3(ceil_div(c->no, kJacBlock)), dim3(kJacBlock),
                     (jacobian_lds<CNP, R>(kJacBlock)), c->stream, c->K, c->q0,
                     c->cams, c->pts, c->meas, c->obs_cam, c->obs_pt, c->fix, c->no, coeff, e, JA, JB, W,
                     (const int*)nullptr, gate, c->loss, c->loss_scale, rc ? 0 : 1);
  probe_end(c, PSBA_PROBE_JACOBIAN);
  const int nb_pt = (int)ceil_div(4L * c->np, 256);
  const JacModel mdl{c->K, c->q0, c->cams, c->pts, c->meas, c->obs_cam, c->fix, c->loss, c->loss_scale};
  if (rc)
    hipLaunchKernelGGL((k_normal_chunks<CNP, R, true>), dim3(nb_pt + c->n_cchunk), dim3(256), 0, c->stream, c->pt_ptr,
                       JB, e, c->np, coeff, coeff_g, c->V, c->g + c->n, c->uvdiag + c->n, nb_pt, c->cchunk_begin,
                       c->cchunk_end, JA, ucpart, (const int2*)c->cchunk_rec, gate, mdl);
  else
    hipLaunchKernelGGL((k_normal_chunks<CNP, R, false>), dim3(nb_pt + c->n_cchunk), dim3(256), 0, c->stream,
                       c->pt_ptr, JB, e, c->np, coeff, coeff_g, c->V, c->g + c->n, c->uvdiag + c->n, nb_pt,
                       c->cchunk_begin, c->cchunk_end, JA, ucpart, (const int2*)c->cchunk_rec, gate, mdl);
  if (with_vinv)
    hipLaunchKernelGGL((k_finish_vinv<CNP, R>), dim3(c->nc + ceil_div(c->np, 256)), dim3(256), 0, c->stream,
                       c->cam_cchunk_ptr, ucpart, c->nc, coeff, coeff_g, c->U, c->g, c->uvdiag, gate, c->V,
                       c->g + c->n, c->np, c->Vinv, c->z, mu_of(c));
  else
    hipLaunchKernelGGL((k_cam_finish<CNP, R>), dim3(c->nc), dim3(256), 0, c->stream, c->cam_cchunk_ptr, ucpart,
                       c->nc, coeff, coeff_g, c->U, c->g, c->uvdiag, (const int*)nullptr, gate);
}

template <int CNP>
static void run_jac_normal(psba_ctx* c, double coeff, double coeff_g, bool with_vinv) {
  if (c->mixed())
    run_jac_normal_t<CNP, float>(c, coeff, coeff_g, c->e32, c->JA32, c->JB32, c->W32, c->ucpart32, with_vinv);
  else
    run_jac_normal_t<CNP, double>(c, coeff, coeff_g, c->e, c->JA, c->JB, c->W, c->ucpart, with_vinv);
}

int launch_jacobian_normal(psba_ctx* c, double coeff, double coeff_g) {
  PSBA_DISPATCH_CNP(c->cnp, run_jac_normal, c, coeff, coeff_g, false);
  PSBA_LAUNCH_TRY();
  return PSBA_OK;
}

int launch_jacobian_normal_vinv(psba_ctx* c) {
  if (!c->devctl) return PSBA_ERR_ARG;
  PSBA_DISPATCH_CNP(c->cnp, run_jac_normal, c, 1.0, 1.0, true);
  PSBA_LAUNCH_TRY();
  return PSBA_OK;
}

// max over UVdiag (single block; once per LM entry)
__global__ __launch_bounds__(1024) void k_maxdiag(const double* __restrict__ d, int n, double* out) {
  __shared__ double sh[16];
  double m = 0.0;
  for (int q = threadIdx.x; q < n; q += blockDim.x) m = fmax(m, d[q]);
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) m = fmax(m, __shfl_xor(m, off, kWave));
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x / 64] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = 0.0;
    for (int w = 0; w < (int)(blockDim.x / 64); ++w) r = fmax(r, sh[w]);
    *out = r;
  }
}

int launch_maxdiag(psba_ctx* c) {
  hipLaunchKernelGGL(k_maxdiag, dim3(1), dim3(1024), 0, c->stream, c->uvdiag, c->ntot, c->d_scal + SC_MAXDIAG);
  PSBA_LAUNCH_TRY();
  return PSBA_OK;
}

int launch_vinv(psba_ctx* c, double mu) {
  // in device-control mode the damping comes from the LM state
  hipLaunchKernelGGL(k_vinv, dim3(ceil_div(c->np, 256)), dim3(256), 0, c->stream, c->V, c->g + c->n, c->np, mu,
                     c->Vinv, c->z, mu_of(c), (const int*)nullptr, (const unsigned char*)nullptr);
  PSBA_LAUNCH_TRY();
  return PSBA_OK;
}

// ---------------------------------------------------------------------------
// back-substitution: eb_i = g_b,i - sum_j W_ij^T dpa_j ; dpb_i = V_i^-1 eb_i
// ---------------------------------------------------------------------------
// Three lanes per point, one per component of eb_i: each lane keeps the single
// running sum over the point's observations (ascending camera order) and
// their CNP products, the order of PSBA/CL_files/compute_eb.cl:20-36, so the
// result does not depend on the launch shape.  A wavefront holds 21 points
// (lane 63 idles); the three eb values meet through lane shuffles for
// dpb_i = V_i^-1 eb_i (V_i^-1 packed as its 6 unique entries).
constexpr int kBacksubPtsPerWave = 21;
constexpr int kBacksubLdsDpa = 4096;  // camera parameters that go through LDS (32 KB)

// With LDS_DPA the camera step dp[0..n) is first copied into LDS (dynamic
// shared memory of n doubles), so that a long track waits for the W rows only.
// SORTED: `order` holds the points by descending track length.
template <int CNP, typename R, bool LDS_DPA, bool SORTED>
__global__ __launch_bounds__(256) void k_backsub(const int* __restrict__ pt_ptr, const int* __restrict__ obs_cam,
                                                 const R* __restrict__ W, const double* __restrict__ Vinv,
                                                 const double* __restrict__ gb, int np, int n,
                                                 double* __restrict__ dp, double* __restrict__ eb,
                                                 const int* __restrict__ ownr, const unsigned char* __restrict__ run,
                                                 const int* __restrict__ order) {
  extern __shared__ double sDa[];
  if constexpr (LDS_DPA) {
    for (int q = threadIdx.x; q < n; q += blockDim.x) sDa[q] = dp[q];
    __syncthreads();
  }
  const int lane = threadIdx.x & 63;
  const int grp = lane / 3;
  const int cc = lane - 3 * grp;
  // slot -> point: with `order` (points by descending track length) the
  // points of a wave, which step through their observations together, have
  // tracks of similar length
  const int slot = (blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * kBacksubPtsPerWave + grp;
  const bool in = grp < kBacksubPtsPerWave && slot < np;
  const int i = (in && SORTED) ? order[slot] : slot;
  const bool ok = in && (!run || run[ownr[i]]);
  R s = 0;
  if constexpr (!SORTED) {
    // points in id order (tracks of any length side by side): each lane
    // walks its own track; camera indices are fetched two observations
    // ahead, so each step waits for one round of loads (W and dpa)
    if (ok) {
      const int beg = pt_ptr[i], end = pt_ptr[i + 1];
      int jc = beg < end ? obs_cam[beg] : 0;
      int jn = beg + 1 < end ? obs_cam[beg + 1] : 0;
#pragma unroll 2
      for (int o = beg; o < end; ++o) {
        const int jnn = o + 2 < end ? obs_cam[o + 2] : 0;
        const R* w = W + (size_t)o * 3 * CNP + cc;
        const double* da;
        if constexpr (LDS_DPA) da = sDa + jc * CNP;
        else da = dp + (size_t)jc * CNP;
#pragma unroll
        for (int k = 0; k < CNP; ++k) s += w[3 * k] * (R)da[k];
        jc = jn;
        jn = jnn;
      }
    }
  } else {
    // points by track length (the lanes of a wave have tracks of similar
    // length): uniform control flow, every lane of the wave steps through the
    // wave's longest track, a lane past its own track (or without a point)
    // loading a valid row and keeping its sum.  The W values and camera index
    // of the next PF steps are in registers before they are used
    // (unconditional loads, so the waits count exactly), and the running sum
    // keeps its order.
    constexpr int PF = 2;
    const int beg = ok ? pt_ptr[i] : 0, len = ok ? pt_ptr[i + 1] - beg : 0;
    int lmax = len;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) lmax = max(lmax, __shfl_xor(lmax, off, 64));
    R wb[PF][CNP];
    int cb[PF];
    auto fetch = [&](int u, int st) {
      const int o = beg + max(0, min(st, len - 1));
      cb[u] = obs_cam[o];
      const R* w = W + (size_t)o * 3 * CNP + cc;
#pragma unroll
      for (int k = 0; k < CNP; ++k) wb[u][k] = w[3 * k];
    };
#pragma unroll
    for (int u = 0; u < PF; ++u) fetch(u, u);
    for (int st0 = 0; st0 < lmax; st0 += PF) {
#pragma unroll
      for (int u = 0; u < PF; ++u) {
        const int st = st0 + u;
        const double* da;
        if constexpr (LDS_DPA) da = sDa + cb[u] * CNP;
        else da = dp + (size_t)cb[u] * CNP;
        const bool live = st < len;
#pragma unroll
        for (int k = 0; k < CNP; ++k) {
          const R t = s + wb[u][k] * (R)da[k];
          s = live ? t : s;
        }
        fetch(u, st + PF);
      }
    }
  }
  const double e = ok ? gb[3 * (size_t)i + cc] - (double)s : 0.0;
  const int base = 3 * grp;
  const double e0 = __shfl(e, base), e1 = __shfl(e, base + 1), e2 = __shfl(e, base + 2);
  if (!ok) return;
  eb[3 * (size_t)i + cc] = e;
  // row cc of the symmetric 3x3 from its packed entries (0 1 2 / 1 3 4 / 2 4 5)
  const double* v = Vinv + (size_t)i * 6;
  const double v0 = v[cc], v1 = v[cc == 0 ? 1 : cc + 2], v2 = v[cc == 0 ? 2 : cc + 3];
  dp[n + 3 * (size_t)i + cc] = v0 * e0 + v1 * e1 + v2 * e2;
}

template <int CNP, typename R>
static void backsub_launch(hipStream_t st, const int* pt_ptr, const int* obs_cam, const R* W, const double* Vinv,
                           const double* gb, int np, int n, double* dp, double* eb, const int* pt_ownr,
                           const unsigned char* run, int block, const int* order = nullptr) {
  const int per_block = (block / 64) * kBacksubPtsPerWave;
  const size_t lds = n <= kBacksubLdsDpa ? sizeof(double) * (size_t)n : 0;
  auto kern = n <= kBacksubLdsDpa ? (order ? k_backsub<CNP, R, true, true> : k_backsub<CNP, R, true, false>)
                                  : (order ? k_backsub<CNP, R, false, true> : k_backsub<CNP, R, false, false>);
  hipLaunchKernelGGL(kern, dim3(ceil_div(np, per_block)), dim3(block), lds, st, pt_ptr, obs_cam, W, Vinv, gb, np, n,
                     dp, eb, pt_ownr, run, order);
}

template <int CNP>
static void run_backsub(psba_ctx* c) {
  probe_begin(c, PSBA_PROBE_BSOLVE);
  if (c->mixed())
    // (FP32 W rows: the id order's locality is worth more than the balance,
    // 33 against 64-71 us at the 88-camera scene)
    backsub_launch<CNP, float>(c->stream, c->pt_ptr, c->obs_cam, c->W32, c->Vinv, c->g + c->n, c->np, c->n, c->dp, c->eb, (const int*)nullptr, (const unsigned char*)nullptr, 256);
  else
    backsub_launch<CNP, double>(c->stream, c->pt_ptr, c->obs_cam, c->W, c->Vinv, c->g + c->n, c->np, c->n, c->dp, c->eb, (const int*)nullptr, (const unsigned char*)nullptr, 256, c->pt_order);
  probe_end(c, PSBA_PROBE_BSOLVE);
}

int launch_backsub(psba_ctx* c) {
  PSBA_DISPATCH_CNP(c->cnp, run_backsub, c);
  PSBA_LAUNCH_TRY();
  return PSBA_OK;
}

// ---------------------------------------------------------------------------
// step statistics, commit, dot products
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_step_stats(const double* __restrict__ dp, const double* __restrict__ g,
                                                    long ntot, double mu, const double* __restrict__ mu_p,
                                                    double* partials) {
  __shared__ double sh[8];
  step_stats_block(dp, g, ntot, mu, mu_p, partials, blockIdx.x, sh);
}

int launch_step_stats(psba_ctx* c, double mu) {
  const unsigned nb = ceil_div(c->ntot, 256);
  hipLaunchKernelGGL(k_step_stats, dim3(nb), dim3(256), 0, c->stream, c->dp, c->g, (long)c->ntot, mu, mu_of(c),
                     c->red_part);
  launch_sum_partials<2>(c->stream, c->red_part, (int)nb, c->d_scal + SC_DP2);
  PSBA_LAUNCH_TRY();
  return PSBA_OK;
}

// With `gate`, nothing is written unless *gate is set (the LM control accepted
// the trial); the |p|^2 partials are then left as they were.
__global__ __launch_bounds__(256) void k_commit(double* __restrict__ cams, double* __restrict__ pts,
                                                const double* __restrict__ dp, int n, long ntot, double* partials,
                                                const unsigned char* __restrict__ gate) {
  if (gate && !*gate) return;
  __shared__ double sh[4];
  double acc[1] = {0.0};
  long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < ntot) {
    double v;
    if (q < n) {
      v = cams[q] + dp[q];
      cams[q] = v;
    } else {
      v = pts[q - n] + dp[q];
      pts[q - n] = v;
    }
    acc[0] = v * v;
  }
  grid_sum_partials<1>(acc, partials, sh);
}

int launch_commit(psba_ctx* c) {
  const unsigned nb = ceil_div(c->ntot, 256);
  // under device control the partials stay where the next trial's LM control
  // kernel sums them, when this commit ran (an accept leaves p_pending set)
  double* part = c->devctl ? c->red_part + commit_part_off(c) : c->red_part;
  hipLaunchKernelGGL(k_commit, dim3(nb), dim3(256), 0, c->stream, c->cams, c->pts, c->dp, c->n, (long)c->ntot, part,
                     gate_of(c, kGateAccept));
  if (!c->devctl) launch_sum_partials<1>(c->stream, c->red_part, (int)nb, c->d_scal + SC_P2);
  PSBA_LAUNCH_TRY();
  return PSBA_OK;
}

// (J x1).(J x1), (J x1).(J x2), (J x2).(J x2) over the observations, J the
// projection Jacobian (coefficient 1), PSBA/CL_files/compute_Jmultiply.cl:6-52
// followed by the host dot products of PSBA/trust_region.cpp:126,174-176,210.
template <int CNP>
__global__ __launch_bounds__(256) void k_jmul_dots(const double* __restrict__ JA, const double* __restrict__ JB,
                                                   const int* __restrict__ obs_cam, const int* __restrict__ obs_pt,
                                                   int no, int n, const double* __restrict__ x1,
                                                   const double* __restrict__ x2, double* partials) {
  __shared__ double sh[12];
  double acc[3] = {0.0, 0.0, 0.0};
  int o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o < no) {
    int j = obs_cam[o], i = obs_pt[o];
    const double* ja = JA + (size_t)o * 2 * CNP;
    const double* jb = JB + (size_t)o * 6;
    double y1[2], y2[2] = {0, 0};
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      double s = 0, t = 0;
#pragma unroll
      for (int k = 0; k < CNP; ++k) {
        s += ja[r * CNP + k] * x1[j * CNP + k];
        if (x2) t += ja[r * CNP + k] * x2[j * CNP + k];
      }
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        s += jb[3 * r + k] * x1[n + 3 * i + k];
        if (x2) t += jb[3 * r + k] * x2[n + 3 * i + k];
      }
      y1[r] = s;
      y2[r] = t;
    }
    acc[0] = y1[0] * y1[0] + y1[1] * y1[1];
    acc[1] = y1[0] * y2[0] + y1[1] * y2[1];
    acc[2] = y2[0] * y2[0] + y2[1] * y2[1];
  }
  grid_sum_partials<3>(acc, partials, sh);
}

template <int CNP>
static void run_jmul(psba_ctx* c, const double* x1, const double* x2) {
  const unsigned nb = ceil_div(c->no, 256);
  hipLaunchKernelGGL(k_jmul_dots<CNP>, dim3(nb), dim3(256), 0, c->stream, c->JA, c->JB, c->obs_cam, c->obs_pt, c->no,
                     c->n, x1, x2, c->red_part);
  launch_sum_partials<3>(c->stream, c->red_part, (int)nb, c->d_scal + SC_D0);
}

int launch_jmul_dots(psba_ctx* c, const double* x1, const double* x2) {
  PSBA_DISPATCH_CNP(c->cnp, run_jmul, c, x1, x2);
  PSBA_LAUNCH_TRY();
  return PSBA_OK;
}

__global__ __launch_bounds__(256) void k_dot(const double* __restrict__ a, const double* __restrict__ b, long n,
                                             double* partials) {
  __shared__ double sh[4];
  double acc[1] = {0.0};
  long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) acc[0] = a[q] * b[q];
  grid_sum_partials<1>(acc, partials, sh);
}

int launch_dot(psba_ctx* c, const double* a, const double* b, int slot) {
  const unsigned nb = ceil_div(c->ntot, 256);
  hipLaunchKernelGGL(k_dot, dim3(nb), dim3(256), 0, c->stream, a, b, (long)c->ntot, c->red_part);
  launch_sum_partials<1>(c->stream, c->red_part, (int)nb, c->d_scal + slot);
  PSBA_LAUNCH_TRY();
  return PSBA_OK;
}

__global__ __launch_bounds__(256) void k_axpby(double* __restrict__ out, double alpha, const double* __restrict__ a,
                                               double beta, const double* __restrict__ b, long n) {
  long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  double v = alpha * a[q];
  if (b) v += beta * b[q];
  out[q] = v;
}

int launch_axpby(psba_ctx* c, double* out, double alpha, const double* a, double beta, const double* b, long n) {
  hipLaunchKernelGGL(k_axpby, dim3(ceil_div(n, 256)), dim3(256), 0, c->stream, out, alpha, a, beta, b, n);
  PSBA_LAUNCH_TRY();
  return PSBA_OK;
}

// ---------------------------------------------------------------------------
// FP64 iterative refinement of the mixed path's step
// ---------------------------------------------------------------------------
// The mixed path assembles the reduced system from FP32 blocks, so its step dp
// solves the damped normal equations (J^T J + mu I) dp = g only to FP32
// accuracy.  A refinement sweep recomputes, per observation in FP64, the
// Jacobian rows and residual (with the loss weights), forms
//   r = J^T (e - J dp) - mu dp        (and the FP64 gradient g = J^T e),
// reduces r with the existing factor and blocks (the same Schur steps as the
// step itself) and adds the correction to dp.
template <int CNP>
__global__ __launch_bounds__(256) void k_refine_obs(const double* __restrict__ K, const double* __restrict__ q0,
                                                    const double* __restrict__ cams, const double* __restrict__ pts,
                                                    const double* __restrict__ meas, const int* __restrict__ obs_cam,
                                                    const int* __restrict__ obs_pt,
                                                    const unsigned char* __restrict__ fix, int no, int n,
                                                    const double* __restrict__ dp, int loss, double loss_d,
                                                    double* __restrict__ rA, double* __restrict__ rB,
                                                    double* __restrict__ gA, double* __restrict__ gB, int diag) {
  // rows go out through LDS as contiguous blocks (as in k_jacobian); every
  // thread reaches the same barriers, only rows o < no are computed
  extern __shared__ double sRef[];
  const int o0 = blockIdx.x * blockDim.x, nrow = min((int)blockDim.x, no - o0);
  const int o = o0 + threadIdx.x;
  const bool act = o < no;
  auto stage = [&](int L, const double* v, double* dst) {
    if (act)
      for (int k = 0; k < L; ++k) sRef[threadIdx.x * (L + 1) + k] = v[k];
    __syncthreads();
    for (int f = threadIdx.x; f < nrow * L; f += blockDim.x) {
      const int rr = f / L;
      dst[(size_t)o0 * L + f] = sRef[rr * (L + 1) + (f - rr * L)];
    }
    __syncthreads();
  };
  double ra[CNP], ga[CNP], rb[3], gbv[3];
  if (act) {
    const int j = obs_cam[o], i = obs_pt[o];
    double cam[CNP], M[3], uv[2], ja[2 * CNP], jb[6];
#pragma unroll
    for (int k = 0; k < CNP; ++k) cam[k] = cams[j * CNP + k];
#pragma unroll
    for (int k = 0; k < 3; ++k) M[k] = pts[3 * i + k];
    project<double, CNP, true>(K + 5 * j, q0 + 4 * j, cam, M, uv, ja, jb);
#pragma unroll
    for (int k = 0; k < CNP; ++k)
      if (fix[j * CNP + k]) ja[k] = ja[CNP + k] = 0.0;
    double e0 = meas[2 * o] - uv[0], e1 = meas[2 * o + 1] - uv[1];
    if (loss) {
      double rho, drho;
      loss_eval(loss, loss_d, e0 * e0 + e1 * e1, rho, drho);
      const double w = sqrt(drho);
#pragma unroll
      for (int k = 0; k < 2 * CNP; ++k) ja[k] *= w;
#pragma unroll
      for (int k = 0; k < 6; ++k) jb[k] *= w;
      e0 *= w;
      e1 *= w;
    }
    if (diag) {
      // diagonal of J^T J (for mu0) in place of r; g as usual
#pragma unroll
      for (int k = 0; k < CNP; ++k) {
        ra[k] = ja[k] * ja[k] + ja[CNP + k] * ja[CNP + k];
        ga[k] = ja[k] * e0 + ja[CNP + k] * e1;
      }
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        rb[k] = jb[k] * jb[k] + jb[3 + k] * jb[3 + k];
        gbv[k] = jb[k] * e0 + jb[3 + k] * e1;
      }
    } else {
      double t0 = 0.0, t1 = 0.0;
#pragma unroll
      for (int k = 0; k < CNP; ++k) {
        const double d = dp[j * CNP + k];
        t0 += ja[k] * d;
        t1 += ja[CNP + k] * d;
      }
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const double d = dp[n + 3 * i + k];
        t0 += jb[k] * d;
        t1 += jb[3 + k] * d;
      }
      const double u0 = e0 - t0, u1 = e1 - t1;
#pragma unroll
      for (int k = 0; k < CNP; ++k) {
        ra[k] = ja[k] * u0 + ja[CNP + k] * u1;
        ga[k] = ja[k] * e0 + ja[CNP + k] * e1;
      }
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        rb[k] = jb[k] * u0 + jb[3 + k] * u1;
        gbv[k] = jb[k] * e0 + jb[3 + k] * e1;
      }
    }
  }
  stage(CNP, ra, rA);
  stage(CNP, ga, gA);
  stage(3, rb, rB);
  stage(3, gbv, gB);
}

// per camera chunk: sums of rA and gA over the chunk's observations.  Output
// t (2 cnp of them) is summed by eight threads p = 0..7 over observations p,
// p + 8, .. (their loads independent, so many are in flight), and the eight
// partial sums are added in p order.
template <int CNP>
__global__ __launch_bounds__(256) void k_refine_cam_chunks(const int* __restrict__ cam_obs, const int* __restrict__ cb,
                                                           const int* __restrict__ ce, const double* __restrict__ rA,
                                                           const double* __restrict__ gA, double* __restrict__ part) {
  static_assert(8 * 2 * CNP <= 256, "eight threads per output");
  __shared__ int sO[kCamChunk];
  __shared__ double sP[8][2 * CNP];
  const int b = blockIdx.x, tid = threadIdx.x;
  const int beg = cb[b], cnt = ce[b] - beg;
  for (int s = tid; s < cnt; s += blockDim.x) sO[s] = cam_obs[beg + s];
  __syncthreads();
  if (tid < 8 * 2 * CNP) {
    const int p = tid / (2 * CNP), t = tid - p * (2 * CNP);
    const double* src = t < CNP ? rA : gA;
    const int k = t < CNP ? t : t - CNP;
    double acc = 0.0;
#pragma unroll 4
    for (int s = p; s < cnt; s += 8) acc += src[(size_t)sO[s] * CNP + k];
    sP[p][t] = acc;
  }
  __syncthreads();
  if (tid < 2 * CNP) {
    double a = sP[0][tid];
#pragma unroll
    for (int p = 1; p < 8; ++p) a += sP[p][tid];
    part[(size_t)b * 2 * CNP + tid] = a;
  }
}

template <int CNP>
__global__ __launch_bounds__(64) void k_refine_cam_finish(const int* __restrict__ ccp, const double* __restrict__ part,
                                                          int nc, double mu, const double* __restrict__ dp,
                                                          double* __restrict__ r, double* __restrict__ g) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nc * CNP) return;
  const int j = q / CNP, k = q - j * CNP;
  double ar = 0.0, ag = 0.0;
  for (int c = ccp[j]; c < ccp[j + 1]; ++c) {
    ar += part[(size_t)c * 2 * CNP + k];
    ag += part[(size_t)c * 2 * CNP + CNP + k];
  }
  r[q] = mu != 0.0 ? ar - mu * dp[q] : ar;
  g[q] = ag;
}

__global__ __launch_bounds__(256) void k_refine_pt(const int* __restrict__ pt_ptr, const double* __restrict__ rB,
                                                   const double* __restrict__ gB, int np, int n, double mu,
                                                   const double* __restrict__ dp, const double* __restrict__ Vinv,
                                                   double* __restrict__ r, double* __restrict__ g,
                                                   double* __restrict__ z) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  double ar[3] = {0, 0, 0}, ag[3] = {0, 0, 0};
  for (int o = pt_ptr[i]; o < pt_ptr[i + 1]; ++o)
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      ar[k] += rB[(size_t)o * 3 + k];
      ag[k] += gB[(size_t)o * 3 + k];
    }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    if (mu != 0.0) ar[k] -= mu * dp[n + 3 * i + k];
    r[n + 3 * i + k] = ar[k];
    g[n + 3 * i + k] = ag[k];
  }
  if (!z) return;
  // z = V^-1 r_b for the reduced right-hand side
  const double* w = Vinv + (size_t)i * 6;
  z[3 * i + 0] = w[0] * ar[0] + w[1] * ar[1] + w[2] * ar[2];
  z[3 * i + 1] = w[1] * ar[0] + w[3] * ar[1] + w[4] * ar[2];
  z[3 * i + 2] = w[2] * ar[0] + w[4] * ar[1] + w[5] * ar[2];
}

template <int CNP>
static void refine_pass(psba_ctx* c, double mu, int diag, double* r, double* z) {
  const int C = CNP;
  double* rA = c->ref_obs;
  double* gA = rA + (size_t)C * c->no;
  double* rB = gA + (size_t)C * c->no;
  double* gB = rB + 3 * (size_t)c->no;
  hipLaunchKernelGGL(k_refine_obs<CNP>, dim3(ceil_div(c->no, 256)), dim3(256), sizeof(double) * 256 * (CNP + 1),
                     c->stream, c->K, c->q0, c->cams,
                     c->pts, c->meas, c->obs_cam, c->obs_pt, c->fix, c->no, c->n, c->dp, c->loss, c->loss_scale, rA, rB,
                     gA, gB, diag);
  if (c->n_cchunk > 0)
    hipLaunchKernelGGL(k_refine_cam_chunks<CNP>, dim3(c->n_cchunk), dim3(256), 0, c->stream, c->cam_obs,
                       c->cchunk_begin, c->cchunk_end, rA, gA, c->ref_part);
  hipLaunchKernelGGL(k_refine_cam_finish<CNP>, dim3(ceil_div((long)c->nc * CNP, 64)), dim3(64), 0, c->stream,
                     c->cam_cchunk_ptr, c->ref_part, c->nc, mu, c->dp, r, c->g);
  hipLaunchKernelGGL(k_refine_pt, dim3(ceil_div(c->np, 256)), dim3(256), 0, c->stream, c->pt_ptr, rB, gB, c->np, c->n,
                     mu, c->dp, c->Vinv, r, c->g, z);
}

template <int CNP>
static int run_refine(psba_ctx* c, double mu, int sweeps) {
  if (sweeps < 0) {
    // FP64 diagonal of J^T J into uvdiag (and the FP64 gradient into g)
    refine_pass<CNP>(c, 0.0, 1, c->uvdiag, nullptr);
    return PSBA_OK;
  }
  // one correction: delta (vecB) solving the reduced system for
  // r = J^T (e - J dp) - mu dp; dp is not changed here
  refine_pass<CNP>(c, mu, 0, c->vecA, c->vecC);
  PSBA_TRY(launch_ea_rhs(c, c->vecA, c->vecC, c->mc_work));
  PSBA_TRY(launch_tri_solve(c, c->mc_work, c->vecB));
  PSBA_TRY(launch_backsub_rhs(c, c->vecA + c->n, c->vecB));
  return PSBA_OK;
}

// sweeps > 0: the correction of dp into vecB (and the FP64 gradient into g);
// sweeps < 0: FP64 diag(J^T J) into uvdiag (for mu0) and the FP64 gradient
// into g, at the current parameters
int launch_refine(psba_ctx* c, double mu, int sweeps) {
  if (sweeps == 0) return PSBA_OK;
  if (!c->ref_obs) {
    if (hipMalloc((void**)&c->ref_obs, sizeof(double) * 2 * (c->cnp + 3) * (size_t)c->no) != hipSuccess) {
      c->ref_obs = nullptr;
      set_last_error(__FILE__, __LINE__, "hipMalloc of the refinement buffers failed");
      return PSBA_ERR_NOMEM;
    }
  }
  if (c->ref_part_cap < (long)c->n_cchunk) {
    if (c->ref_part) (void)hipFree(c->ref_part);
    c->ref_part = nullptr;
    c->ref_part_cap = 0;
    if (hipMalloc((void**)&c->ref_part, sizeof(double) * 2 * c->cnp * (size_t)std::max(c->n_cchunk, 1)) != hipSuccess) {
      c->ref_part = nullptr;
      set_last_error(__FILE__, __LINE__, "hipMalloc of the refinement buffers failed");
      return PSBA_ERR_NOMEM;
    }
    c->ref_part_cap = c->n_cchunk;
  }
  int rc = PSBA_OK;
  switch (c->cnp) {
    case 6: rc = run_refine<6>(c, mu, sweeps); break;
    case 9: rc = run_refine<9>(c, mu, sweeps); break;
    case 11: rc = run_refine<11>(c, mu, sweeps); break;
    case 16: rc = run_refine<16>(c, mu, sweeps); break;
    default: return PSBA_ERR_UNSUPPORTED;
  }
  if (rc != PSBA_OK) return rc;
  PSBA_LAUNCH_TRY();
  return PSBA_OK;
}

// dp_out[n..) = V^-1 (gb - W^T dp_out[0..n)) with a caller-given right-hand side
template <int CNP>
static void run_backsub_rhs(psba_ctx* c, const double* gb, double* dp_out) {
  if (c->mixed())
    backsub_launch<CNP, float>(c->stream, c->pt_ptr, c->obs_cam, c->W32, c->Vinv, gb, c->np, c->n, dp_out, c->eb, (const int*)nullptr, (const unsigned char*)nullptr, 256);
  else
    backsub_launch<CNP, double>(c->stream, c->pt_ptr, c->obs_cam, c->W, c->Vinv, gb, c->np, c->n, dp_out, c->eb, (const int*)nullptr, (const unsigned char*)nullptr, 256, c->pt_order);
}

int launch_backsub_rhs(psba_ctx* c, const double* gb, double* dp_out) {
  PSBA_DISPATCH_CNP(c->cnp, run_backsub_rhs, c, gb, dp_out);
  PSBA_LAUNCH_TRY();
  return PSBA_OK;
}

}  // namespace psba

// ---------------------------------------------------------------------------
// batched problems: the same stage kernels over the concatenated arrays,
// masked per problem
// ---------------------------------------------------------------------------
#include "batch.hpp"

namespace psba {

template <int CNP, typename R>
static void run_b_jac_t(psba_batch* b) {
  R *e = (R*)b->e, *JA = (R*)b->JA, *JB = (R*)b->JB, *W = (R*)b->W, *ucpart = (R*)b->ucpart;
  // staging buffer within 64 KB (cnp 16 in FP64 with 256 threads would not be)
  const int bs = jacobian_lds<CNP, R>(b->obs_block) <= 65536 ? b->obs_block : 128;
  hipLaunchKernelGGL((k_jacobian<CNP, R>), dim3(ceil_div(b->NO, bs)), dim3(bs), (jacobian_lds<CNP, R>(bs)), b->stream,
                     b->K, b->q0, b->cams,
                     b->pts, b->meas, b->obs_cam, b->obs_pt, b->fix, (int)b->NO, 1.0, e, JA, JB, W, b->obs_prob,
                     b->need_jac, PSBA_LOSS_NONE, 1.0, 1);
  hipLaunchKernelGGL(k_point_normal<R>, dim3(ceil_div(4L * b->NP, b->pt_block)), dim3(b->pt_block), 0, b->stream, b->pt_ptr, JB, e,
                     (int)b->NP, 1.0, 1.0, b->V, b->g + b->N, b->uvdiag + b->N, b->pt_prob, b->need_jac);
  if (b->n_cchunk > 0)
    hipLaunchKernelGGL((k_cam_chunks<CNP, R>), dim3(b->n_cchunk), dim3(256), 0, b->stream, b->cam_obs, b->cchunk_b,
                       b->cchunk_e, JA, e, ucpart, b->cchunk_prob, b->need_jac, (const int2*)nullptr);
  hipLaunchKernelGGL((k_cam_finish<CNP, R>), dim3(b->NC), dim3(256), 0, b->stream, b->ccp, ucpart, (int)b->NC, 1.0,
                     1.0, b->U, b->g, b->uvdiag, b->cam_prob, b->need_jac);
}

template <int CNP>
static void run_b_jac(psba_batch* b) {
  if (b->mixed())
    run_b_jac_t<CNP, float>(b);
  else
    run_b_jac_t<CNP, double>(b);
}

int launch_b_jacobian(psba_batch* b) {
  PSBA_DISPATCH_CNP(b->cnp, run_b_jac, b);
  PSBA_LAUNCH_TRY();
  return PSBA_OK;
}

int launch_b_vinv(psba_batch* b) {
  hipLaunchKernelGGL(k_vinv, dim3(ceil_div(b->NP, b->pt_block)), dim3(b->pt_block), 0, b->stream, b->V, b->g + b->N, (int)b->NP, 0.0,
                     b->Vinv, b->z, b->mu, b->pt_prob, b->active);
  PSBA_LAUNCH_TRY();
  return PSBA_OK;
}

template <int CNP>
static void run_b_backsub(psba_batch* b) {
  if (b->mixed())
    backsub_launch<CNP, float>(b->stream, b->pt_ptr, b->obs_cam, (const float*)b->W, b->Vinv, b->g + b->N, (int)b->NP, (int)b->N, b->dp, b->eb, b->pt_prob, b->active, b->pt_block);
  else
    backsub_launch<CNP, double>(b->stream, b->pt_ptr, b->obs_cam, (const double*)b->W, b->Vinv, b->g + b->N, (int)b->NP, (int)b->N, b->dp, b->eb, b->pt_prob, b->active, b->pt_block);
}

int launch_b_backsub(psba_batch* b) {
  PSBA_DISPATCH_CNP(b->cnp, run_b_backsub, b);
  PSBA_LAUNCH_TRY();
  return PSBA_OK;
}

}  // namespace psba
