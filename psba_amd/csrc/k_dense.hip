// Dense FP64 linear algebra on the reduced camera matrix S (n_pad x n_pad,
// row-major, lower triangle; rows/cols >= n are the identity):
//   - blocked right-looking Cholesky S = L L^T in place (32x32 tiles):
//     diagonal-tile factor, panel triangular solve, trailing update;
//     replaces the reference's explicit inverse (PSBA/CL_files/SPD_inv.cl,
//     kern_cholesky/_s2 + kern_trigMat_inv/_mul + kern_fill_rest);
//   - forward/back substitution for S dpa = e_a (replaces matVec_mul.cl);
//   - the dogleg's modified Cholesky and lambda = |sum E| / n
//     (PSBA/CL_files/cholmod_blk.cl, PSBA/cl_cholmod.cpp:109-202).
// Failure of the factorisation (a pivot that is not finite and positive) is
// recorded in d_scal[SC_FACTOR_OK]; later kernels of the same factorisation
// then do nothing.
#include "context.hpp"

namespace psba {

constexpr int NB = kNB;

// factor the diagonal tile kb; thread 0 of kb == 0 resets the status flag
__global__ __launch_bounds__(64) void k_potrf_tile(double* __restrict__ S, int ld, int kb, double* ok) {
  __shared__ double A[NB][NB + 1];
  __shared__ int good;
  if (kb == 0) {
    if (threadIdx.x == 0) *ok = 1.0;
  } else if (*ok == 0.0) {
    return;
  }
  const int base = kb * NB;
  for (int f = threadIdx.x; f < NB * NB; f += blockDim.x) {
    int r = f / NB, c = f - r * NB;
    A[r][c] = (c <= r) ? S[(size_t)(base + r) * ld + base + c] : 0.0;
  }
  if (threadIdx.x == 0) good = 1;
  __syncthreads();
  for (int j = 0; j < NB; ++j) {
    if (threadIdx.x == 0) {
      double d = A[j][j];
      if (!(d > 0.0) || !isfinite(d)) {
        good = 0;
      } else {
        A[j][j] = sqrt(d);
      }
    }
    __syncthreads();
    if (!good) break;
    const double ljj = A[j][j];
    for (int i = j + 1 + threadIdx.x; i < NB; i += blockDim.x) A[i][j] = A[i][j] / ljj;
    __syncthreads();
    const int m = NB - 1 - j;  // trailing size
    for (int f = threadIdx.x; f < m * m; f += blockDim.x) {
      int i = j + 1 + f / m, c = j + 1 + f % m;
      if (c <= i) A[i][c] -= A[i][j] * A[c][j];
    }
    __syncthreads();
  }
  if (!good) {
    if (threadIdx.x == 0) *ok = 0.0;
    return;
  }
  for (int f = threadIdx.x; f < NB * NB; f += blockDim.x) {
    int r = f / NB, c = f - r * NB;
    if (c <= r) {
      double v = A[r][c];
      if (!isfinite(v)) good = 0;
      S[(size_t)(base + r) * ld + base + c] = v;
    }
  }
  __syncthreads();
  if (!good && threadIdx.x == 0) *ok = 0.0;
}

// L_ib,kb = A_ib,kb L_kk^-T for every tile row ib > kb (one block per tile)
__global__ __launch_bounds__(64) void k_trsm_tiles(double* __restrict__ S, int ld, int kb, const double* ok) {
  __shared__ double Lk[NB][NB + 1];
  __shared__ double X[NB][NB + 1];
  if (*ok == 0.0) return;
  const int ib = kb + 1 + blockIdx.x;
  const int kbase = kb * NB, ibase = ib * NB;
  for (int f = threadIdx.x; f < NB * NB; f += blockDim.x) {
    int r = f / NB, c = f - r * NB;
    Lk[r][c] = (c <= r) ? S[(size_t)(kbase + r) * ld + kbase + c] : 0.0;
    X[r][c] = S[(size_t)(ibase + r) * ld + kbase + c];
  }
  __syncthreads();
  if (threadIdx.x < NB) {
    const int r = threadIdx.x;
    for (int c = 0; c < NB; ++c) {
      double s = X[r][c];
      for (int m = 0; m < c; ++m) s -= X[r][m] * Lk[c][m];
      X[r][c] = s / Lk[c][c];
    }
  }
  __syncthreads();
  for (int f = threadIdx.x; f < NB * NB; f += blockDim.x) {
    int r = f / NB, c = f - r * NB;
    S[(size_t)(ibase + r) * ld + kbase + c] = X[r][c];
  }
}

// A_ib,jb -= L_ib,kb L_jb,kb^T for kb < jb <= ib (one block per tile)
__global__ __launch_bounds__(256) void k_syrk_tiles(double* __restrict__ S, int ld, int kb, int T,
                                                    const double* ok) {
  __shared__ double Li[NB][NB + 1];
  __shared__ double Lj[NB][NB + 1];
  if (*ok == 0.0) return;
  // linear tile index -> (a, b) with 0 <= b <= a < T
  const int t = blockIdx.x;
  int a = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
  while ((a + 1) * (a + 2) / 2 <= t) ++a;
  while (a * (a + 1) / 2 > t) --a;
  const int b = t - a * (a + 1) / 2;
  const int ib = kb + 1 + a, jb = kb + 1 + b;
  const int kbase = kb * NB, ibase = ib * NB, jbase = jb * NB;
  for (int f = threadIdx.x; f < NB * NB; f += blockDim.x) {
    int r = f / NB, c = f - r * NB;
    Li[r][c] = S[(size_t)(ibase + r) * ld + kbase + c];
    Lj[r][c] = S[(size_t)(jbase + r) * ld + kbase + c];
  }
  __syncthreads();
  // each thread: 4 outputs (rows r0, r0+8, r0+16, r0+24 ; column c)
  const int c = threadIdx.x & (NB - 1), r0 = threadIdx.x / NB;
  double acc[4] = {0, 0, 0, 0};
  for (int m = 0; m < NB; ++m) {
    double lj = Lj[c][m];
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[q] += Li[r0 + 8 * q][m] * lj;
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    int r = r0 + 8 * q;
    if (ib != jb || c <= r) S[(size_t)(ibase + r) * ld + jbase + c] -= acc[q];
  }
}

int launch_factor(psba_ctx* c) {
  const int nbk = c->n_pad / NB;
  double* ok = c->d_scal + SC_FACTOR_OK;
  for (int kb = 0; kb < nbk; ++kb) {
    hipLaunchKernelGGL(k_potrf_tile, dim3(1), dim3(64), 0, c->stream, c->S, c->n_pad, kb, ok);
    const int T = nbk - kb - 1;
    if (T > 0) {
      hipLaunchKernelGGL(k_trsm_tiles, dim3(T), dim3(64), 0, c->stream, c->S, c->n_pad, kb, ok);
      hipLaunchKernelGGL(k_syrk_tiles, dim3(T * (T + 1) / 2), dim3(256), 0, c->stream, c->S, c->n_pad, kb, T, ok);
    }
  }
  PSBA_LAUNCH_TRY();
  return PSBA_OK;
}

// Forward then back substitution with the factor in S: x = L^-T L^-1 b.
// Single block of 256 threads (4 waves); the solution vector lives in LDS
// (n_pad doubles of dynamic shared memory), y is overwritten by x in place.
__global__ __launch_bounds__(256) void k_trsv(const double* __restrict__ L, int ld, int nbk,
                                              const double* __restrict__ b, double* __restrict__ xout,
                                              const double* ok) {
  extern __shared__ double vsh[];
  __shared__ double rhs[NB];
  __shared__ double part[8][NB];
  if (*ok == 0.0) return;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  // ---- forward: L y = b ----
  for (int kb = 0; kb < nbk; ++kb) {
    const int base = kb * NB;
    // rows base..base+31: rhs = b - L[row][0:base] y[0:base]; wave w takes 8 rows
    for (int rr = 0; rr < 8; ++rr) {
      const int r = wid * 8 + rr;
      const double* Lr = L + (size_t)(base + r) * ld;
      double s = 0.0;
      for (int cidx = lane; cidx < base; cidx += 64) s += Lr[cidx] * vsh[cidx];
      s = wave_sum(s);
      if (lane == 0) rhs[r] = b[base + r] - s;
    }
    __syncthreads();
    if (wid == 0) {
      // lane r < 32 owns row r of the diagonal tile
      double Lrow[NB];
      double br = 0.0;
      if (lane < NB) {
        br = rhs[lane];
#pragma unroll
        for (int m = 0; m < NB; ++m) Lrow[m] = L[(size_t)(base + lane) * ld + base + m];
      } else {
#pragma unroll
        for (int m = 0; m < NB; ++m) Lrow[m] = 1.0;
      }
      double yv = 0.0;
#pragma unroll
      for (int m = 0; m < NB; ++m) {
        double cand = br / Lrow[m];
        double ym = __shfl(cand, m, 64);
        if (lane == m) yv = ym;
        if (lane > m) br -= Lrow[m] * ym;
      }
      if (lane < NB) vsh[base + lane] = yv;
    }
    __syncthreads();
  }
  // ---- backward: L^T x = y ----
  for (int kb = nbk - 1; kb >= 0; --kb) {
    const int base = kb * NB;
    // rhs[col] = y[col] - sum_{row >= base+NB} L[row][col] x[row]
    {
      const int col = threadIdx.x & (NB - 1), grp = threadIdx.x / NB;  // 8 groups
      double s = 0.0;
      for (int row = base + NB + grp; row < nbk * NB; row += 8) s += L[(size_t)row * ld + base + col] * vsh[row];
      part[grp][col] = s;
    }
    __syncthreads();
    if (threadIdx.x < NB) {
      double s = 0.0;
      for (int q = 0; q < 8; ++q) s += part[q][threadIdx.x];
      rhs[threadIdx.x] = vsh[base + threadIdx.x] - s;
    }
    __syncthreads();
    if (wid == 0) {
      // lane r < 32 owns column r of the diagonal tile: Lcol[m] = L[base+m][base+r]
      double Lcol[NB];
      double br = 0.0;
      if (lane < NB) {
        br = rhs[lane];
#pragma unroll
        for (int m = 0; m < NB; ++m) Lcol[m] = (m >= lane) ? L[(size_t)(base + m) * ld + base + lane] : 0.0;
      } else {
#pragma unroll
        for (int m = 0; m < NB; ++m) Lcol[m] = 1.0;
      }
      double xv = 0.0;
#pragma unroll
      for (int m = NB - 1; m >= 0; --m) {
        double cand = br / Lcol[m];
        double xm = __shfl(cand, m, 64);
        if (lane == m) xv = xm;
        if (lane < m) br -= Lcol[m] * xm;
      }
      if (lane < NB) vsh[base + lane] = xv;
    }
    __syncthreads();
  }
  for (int q = threadIdx.x; q < nbk * NB; q += blockDim.x) xout[q] = vsh[q];
}

__global__ void k_copy_n(double* __restrict__ dst, const double* __restrict__ src, int n, const double* ok) {
  if (*ok == 0.0) return;
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) dst[q] = src[q];
}

int launch_solve(psba_ctx* c) {
  const int nbk = c->n_pad / NB;
  const double* ok = c->d_scal + SC_FACTOR_OK;
  if (c->n_pad > kMaxSolveN) return PSBA_ERR_UNSUPPORTED;
  hipLaunchKernelGGL(k_trsv, dim3(1), dim3(256), sizeof(double) * c->n_pad, c->stream, c->S, c->n_pad, nbk, c->ea,
                     c->xsol, ok);
  hipLaunchKernelGGL(k_copy_n, dim3(ceil_div(c->n, 256)), dim3(256), 0, c->stream, c->dp, c->xsol, c->n, ok);
  PSBA_LAUNCH_TRY();
  return PSBA_OK;
}

__global__ void k_copy_mat(double* __restrict__ dst, const double* __restrict__ src, long count) {
  long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < count) dst[q] = src[q];
}

int launch_copy_S(psba_ctx* c, double* dst, const double* src) {
  long cnt = (long)c->n_pad * c->n_pad;
  hipLaunchKernelGGL(k_copy_mat, dim3(ceil_div(cnt, 256)), dim3(256), 0, c->stream, dst, src, cnt);
  PSBA_LAUNCH_TRY();
  return PSBA_OK;
}

// ---------------------------------------------------------------------------
// modified Cholesky (dogleg recovery path), single block of 256 threads.
// A (read-only original, lower triangle) -> factor in M; lambda to *out.
// ---------------------------------------------------------------------------
__device__ inline double block_max256(double v, double* sh) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
  __syncthreads();
  return r;
}

__device__ inline double block_sum256(double v, double* sh) {
  v = wave_sum(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = ((sh[0] + sh[1]) + sh[2]) + sh[3];
  __syncthreads();
  return r;
}

__device__ inline int block_or256(int v, int* sh) {
  __syncthreads();
  if (threadIdx.x == 0) *sh = 0;
  __syncthreads();
  if (v) atomicOr(sh, 1);
  __syncthreads();
  int r = *sh;
  __syncthreads();
  return r;
}

__global__ __launch_bounds__(256) void k_modchol(const double* __restrict__ A, double* __restrict__ M, int ld, int n,
                                                 double* __restrict__ Cc, double* out) {
  __shared__ double sh[4];
  __shared__ int shi;
  __shared__ double T[9], di[9];
  __shared__ int blk_ok;
  // delta / beta from the original matrix (lower triangle, symmetric)
  double xi = 0.0, gam = 0.0;
  for (long f = threadIdx.x; f < (long)n * n; f += blockDim.x) {
    int r = (int)(f / n), k = (int)(f - (long)r * n);
    if (k > r) continue;
    double v = fabs(A[(size_t)r * ld + k]);
    if (k == r) gam = fmax(gam, v);
    else xi = fmax(xi, v);
  }
  xi = block_max256(xi, sh);
  gam = block_max256(gam, sh);
  const double delta = 1e-15 * fmax(xi + gam, 1.0);
  double beta = fmax(gam, 1e-15);
  beta = fmax(beta, xi / sqrt((double)(n * n - 1)));
  beta = sqrt(beta);
  // clear the factor's lower triangle
  for (long f = threadIdx.x; f < (long)n * n; f += blockDim.x) {
    int r = (int)(f / n), k = (int)(f - (long)r * n);
    if (k <= r) M[(size_t)r * ld + k] = 0.0;
  }
  __syncthreads();
  const int nb = n / 3;
  for (int j = 0; j < nb; ++j) {
    const int j3 = 3 * j;
    if (threadIdx.x < 9) {
      int u = threadIdx.x / 3, v = threadIdx.x % 3;
      double s = A[(size_t)(j3 + u) * ld + j3 + v];
      if (v > u) s = A[(size_t)(j3 + v) * ld + j3 + u];
      for (int k = 0; k < j3; ++k) s -= M[(size_t)(j3 + u) * ld + k] * M[(size_t)(j3 + v) * ld + k];
      T[threadIdx.x] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      int okb = 1;
      double l0 = T[0];
      if (!isfinite(l0) || l0 <= 0) okb = 0;
      l0 = sqrt(l0);
      double l1 = T[3] / sqrt(T[0]);
      if (!isfinite(l1)) okb = 0;
      double l2 = T[4] - T[3] * T[3] / T[0];
      if (!isfinite(l2) || l2 <= 0) okb = 0;
      l2 = sqrt(l2);
      double l3 = T[6] / sqrt(T[0]);
      if (!isfinite(l3)) okb = 0;
      double l4 = sqrt(T[0] / (T[0] * T[4] - T[3] * T[3])) * (T[7] - T[3] * T[6] / T[0]);
      if (!isfinite(l4)) okb = 0;
      double t1 = -T[8] * T[3] * T[3], t2 = 2 * T[7] * T[3] * T[6], t3 = -T[4] * T[6] * T[6];
      double t4 = T[0] * (T[4] * T[8] - T[7] * T[7]), t5 = -T[3] * T[3] + T[0] * T[4];
      double l5 = (t1 + t2 + t3 + t4) / t5;
      if (!isfinite(l5) || l5 <= 0) okb = 0;
      l5 = sqrt(l5);
      blk_ok = okb;
      if (okb) {
        di[0] = 1 / l0; di[1] = 0; di[2] = 0;
        di[3] = -l1 / (l0 * l2); di[4] = 1 / l2; di[5] = 0;
        di[6] = (l1 * l4 - l2 * l3) / (l0 * l2 * l5); di[7] = -l4 / (l2 * l5); di[8] = 1 / l5;
        size_t bb = (size_t)j3 * ld + j3;
        M[bb] = l0; M[bb + ld] = l1; M[bb + ld + 1] = l2;
        M[bb + 2 * ld] = l3; M[bb + 2 * ld + 1] = l4; M[bb + 2 * ld + 2] = l5;
      }
    }
    __syncthreads();
    int over = 0;
    if (blk_ok) {
      for (int i = j + 1 + threadIdx.x; i < nb; i += blockDim.x) {
        const int i3 = 3 * i;
        double Tij[9];
        for (int u = 0; u < 3; ++u)
          for (int v = 0; v < 3; ++v) {
            double s = A[(size_t)(i3 + u) * ld + j3 + v];
            for (int k = 0; k < j3; ++k) s -= M[(size_t)(i3 + u) * ld + k] * M[(size_t)(j3 + v) * ld + k];
            Tij[3 * u + v] = s;
          }
        for (int u = 0; u < 3; ++u)
          for (int v = 0; v < 3; ++v) {
            double s = Tij[3 * u] * di[3 * v] + Tij[3 * u + 1] * di[3 * v + 1] + Tij[3 * u + 2] * di[3 * v + 2];
            M[(size_t)(i3 + u) * ld + j3 + v] = s;
            if (s > beta) over = 1;
          }
      }
      over = block_or256(over, &shi);
      if (!over) continue;
      // restore the block column: clear it, then take the column-wise path
      for (int f = threadIdx.x; f < (n - j3) * 3; f += blockDim.x) {
        int i = j3 + f / 3, v = f % 3;
        M[(size_t)i * ld + j3 + v] = 0.0;
      }
    } else {
      if (threadIdx.x < 9) {
        int u = threadIdx.x / 3, v = threadIdx.x % 3;
        M[(size_t)(j3 + u) * ld + j3 + v] = 0.0;
      }
    }
    __syncthreads();
    for (int col = 0; col < 3; ++col) {
      const int xj = j3 + col;
      double part = 0.0;
      for (int k = threadIdx.x; k < xj; k += blockDim.x) part += M[(size_t)xj * ld + k] * M[(size_t)xj * ld + k];
      double ssum = block_sum256(part, sh);
      double dj = fmax(fabs(A[(size_t)xj * ld + xj] - ssum), delta);
      double sd = sqrt(dj);
      if (threadIdx.x == 0) M[(size_t)xj * ld + xj] = sd;
      int ov = 0;
      for (int i = xj + 1 + threadIdx.x; i < n; i += blockDim.x) {
        double s = A[(size_t)i * ld + xj];
        for (int k = 0; k < xj; ++k) s = s - M[(size_t)i * ld + k] * M[(size_t)xj * ld + k];
        Cc[i] = s;
        double lij = s / sd;
        M[(size_t)i * ld + xj] = lij;
        if (lij > beta) ov = 1;
      }
      ov = block_or256(ov, &shi);
      if (ov) {
        double th = 0.0;
        for (int i = xj + 1 + threadIdx.x; i < n; i += blockDim.x) th = fmax(th, fabs(Cc[i]));
        th = block_max256(th, sh);
        double nd = th / beta;
        if (threadIdx.x == 0) M[(size_t)xj * ld + xj] = nd;
        for (int i = xj + 1 + threadIdx.x; i < n; i += blockDim.x) M[(size_t)i * ld + xj] = Cc[i] / nd;
      }
      __syncthreads();
    }
  }
  // E_i = (L L^T)_ii - A_ii ; lambda = |sum E| / n
  double es = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    double s = 0.0;
    for (int k = 0; k <= i; ++k) s += M[(size_t)i * ld + k] * M[(size_t)i * ld + k];
    es += s - A[(size_t)i * ld + i];
  }
  es = block_sum256(es, sh);
  if (threadIdx.x == 0) *out = fabs(es) / n;
}

int launch_modchol(psba_ctx* c, double* A) {
  hipLaunchKernelGGL(k_modchol, dim3(1), dim3(256), 0, c->stream, A, c->S, c->n_pad, c->n, c->mc_work,
                     c->d_scal + SC_LAMBDA);
  PSBA_LAUNCH_TRY();
  return PSBA_OK;
}

}  // namespace psba
