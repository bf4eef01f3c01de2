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

constexpr int NB = kNB;   // 64
constexpr int TP = NB + 1; // padded LDS row length

// ---------------------------------------------------------------------------
// blocked right-looking Cholesky, 64x64 tiles.  Step kb:
//   k_potrf64   factor the diagonal tile (one wave, rows in registers)
//   k_trsm64    L_ib,kb = A_ib,kb L_kk^-T for ib > kb (one wave per tile) and,
//               in one extra block, the forward substitution y_kb = L_kk^-1 b_kb
//   k_syrk64    A_ib,jb -= L_ib,kb L_jb,kb^T for kb < jb <= ib (tile GEMM) and,
//               in extra blocks, b_jb -= L_jb,kb y_kb
// so that after the last step b holds y = L^-1 b.
// ---------------------------------------------------------------------------
// Diagonal tile: 4 waves, lane = row, wave W owns columns 16W..16W+15 of all
// 64 rows (registers); the pivot column goes through double-buffered LDS with
// one barrier per pivot.  The wave index is a template parameter so every
// column test is resolved at compile time (no branches around the updates).
template <int W>
__device__ __forceinline__ void potrf_wave(double (&a)[16], double (*col)[NB], int lane, int* bad) {
#pragma unroll 1
  for (int kb = 0; kb < 4; ++kb) {
#pragma unroll
    for (int kk = 0; kk < 16; ++kk) {
      const int k = 16 * kb + kk;
      double* cb = col[kk & 1];
      if (kb == W) {
        const double dkk = __shfl(a[kk], k, 64);
        const double lkk = sqrt(dkk);
        const double v = (lane == k) ? lkk : a[kk] / lkk;
        a[kk] = v;
        cb[lane] = v;
        if (lane == 0 && !(dkk > 0.0 && isfinite(dkk))) *bad = 1;
      }
      __syncthreads();
      const double lik = cb[lane];
      if (W > kb) {
#pragma unroll
        for (int q = 0; q < 16; ++q) a[q] -= lik * cb[16 * W + q];
      } else if (W == kb) {
#pragma unroll
        for (int q = kk + 1; q < 16; ++q) a[q] -= lik * cb[16 * W + q];
      }
    }
  }
}

// A failed pivot leaves the tile untouched and clears *ok.
__global__ __launch_bounds__(256) void k_potrf64(double* __restrict__ S, int ld, int kb, double* ok) {
  __shared__ double T[NB * TP];
  __shared__ double col[2][NB];
  __shared__ int bad;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (kb == 0) {
    if (threadIdx.x == 0) *ok = 1.0;
  } else if (*ok == 0.0) {
    return;
  }
  const int base = kb * NB;
  for (int f = threadIdx.x; f < NB * NB; f += blockDim.x) {
    int r = f >> 6, c = f & 63;
    T[r * TP + c] = (c <= r) ? S[(size_t)(base + r) * ld + base + c] : 0.0;
  }
  if (threadIdx.x == 0) bad = 0;
  __syncthreads();
  double a[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) a[q] = T[lane * TP + 16 * w + q];
  switch (w) {
    case 0: potrf_wave<0>(a, col, lane, &bad); break;
    case 1: potrf_wave<1>(a, col, lane, &bad); break;
    case 2: potrf_wave<2>(a, col, lane, &bad); break;
    default: potrf_wave<3>(a, col, lane, &bad); break;
  }
#pragma unroll
  for (int q = 0; q < 16; ++q) T[lane * TP + 16 * w + q] = a[q];
  __syncthreads();
  if (bad) {
    if (threadIdx.x == 0) *ok = 0.0;
    return;
  }
  for (int f = threadIdx.x; f < NB * NB; f += blockDim.x) {
    int r = f >> 6, c = f & 63;
    if (c <= r) S[(size_t)(base + r) * ld + base + c] = T[r * TP + c];
  }
}

template <int W>
__device__ __forceinline__ void trsm_wave(double (&x)[16], double (*col)[NB], const double* L, const double* rd,
                                          int lane) {
#pragma unroll
  for (int c = 0; c < NB; ++c) {
    double* cb = col[c & 1];
    if (W == (c >> 4)) {
      x[c & 15] *= rd[c];
      cb[lane] = x[c & 15];
    }
    __syncthreads();
    const double xc = cb[lane];
#pragma unroll
    for (int q = 0; q < 16; ++q)
      if (16 * W + q > c) x[q] -= xc * L[(16 * W + q) * TP + c];
  }
}

// Panel: L_ib,kb = A_ib,kb L_kk^-T, one block (4 waves) per tile row ib > kb;
// lane = row, wave w owns columns 16w..16w+15.  The last block instead does
// the forward substitution y_kb = L_kk^-1 b_kb (wave 0).
__global__ __launch_bounds__(256) void k_trsm64(double* __restrict__ S, int ld, int kb, int T,
                                               double* __restrict__ b, const double* ok) {
  __shared__ double L[NB * TP];
  __shared__ double rd[NB];
  __shared__ double col[2][NB];
  if (*ok == 0.0) return;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, kbase = kb * NB;
  for (int f = threadIdx.x; f < NB * NB; f += blockDim.x) {
    int r = f >> 6, c = f & 63;
    L[r * TP + c] = (c <= r) ? S[(size_t)(kbase + r) * ld + kbase + c] : 0.0;
  }
  __syncthreads();
  if ((int)blockIdx.x == T) {
    if (w != 0) return;
    double Lrow[NB];
#pragma unroll
    for (int m = 0; m < NB; ++m) Lrow[m] = L[lane * TP + m];
    double br = b[kbase + lane], yv = 0.0;
#pragma unroll
    for (int m = 0; m < NB; ++m) {
      const double ym = __shfl(br / Lrow[m], m, 64);
      if (lane == m) yv = ym;
      if (lane > m) br -= Lrow[m] * ym;
    }
    b[kbase + lane] = yv;
    return;
  }
  if (threadIdx.x < NB) rd[threadIdx.x] = 1.0 / L[threadIdx.x * TP + threadIdx.x];
  const size_t ibase = (size_t)(kb + 1 + blockIdx.x) * NB;
  double x[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) x[q] = S[(ibase + lane) * ld + kbase + 16 * w + q];
  __syncthreads();
  switch (w) {
    case 0: trsm_wave<0>(x, col, L, rd, lane); break;
    case 1: trsm_wave<1>(x, col, L, rd, lane); break;
    case 2: trsm_wave<2>(x, col, L, rd, lane); break;
    default: trsm_wave<3>(x, col, L, rd, lane); break;
  }
#pragma unroll
  for (int q = 0; q < 16; ++q) S[(ibase + lane) * ld + kbase + 16 * w + q] = x[q];
}

// blocks [0, ntiles): trailing tile (ib, jb); blocks [ntiles, ntiles+T): b update
__global__ __launch_bounds__(256) void k_syrk64(double* __restrict__ S, int ld, int kb, int T, int ntiles,
                                               double* __restrict__ b, const double* ok) {
  constexpr int AP = NB + 4;  // transposed tiles: [m][r], padded
  __shared__ double As[NB * AP];
  __shared__ double Bs[NB * AP];
  if (*ok == 0.0) return;
  const int kbase = kb * NB;
  const int t = blockIdx.x;
  if (t >= ntiles) {
    // b_jb -= L_jb,kb y_kb : 4 threads per row
    const int jbase = (kb + 1 + (t - ntiles)) * NB;
    const int r = threadIdx.x >> 2, part = threadIdx.x & 3;
    const double* Lr = S + (size_t)(jbase + r) * ld + kbase + part * 16;
    const double* y = b + kbase + part * 16;
    double s = 0.0;
#pragma unroll
    for (int c = 0; c < 16; ++c) s += Lr[c] * y[c];
    s += __shfl_xor(s, 1, 64);
    s += __shfl_xor(s, 2, 64);
    if (part == 0) b[jbase + r] -= s;
    return;
  }
  int a = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
  while ((a + 1) * (a + 2) / 2 <= t) ++a;
  while (a * (a + 1) / 2 > t) --a;
  const int bb = t - a * (a + 1) / 2;
  const int ib = kb + 1 + a, jb = kb + 1 + bb;
  const int ibase = ib * NB, jbase = jb * NB;
  for (int f = threadIdx.x; f < NB * NB; f += blockDim.x) {
    int r = f >> 6, m = f & 63;
    As[m * AP + r] = S[(size_t)(ibase + r) * ld + kbase + m];
    Bs[m * AP + r] = S[(size_t)(jbase + r) * ld + kbase + m];
  }
  __syncthreads();
  const int r0 = 4 * (threadIdx.x >> 4), c0 = 4 * (threadIdx.x & 15);
  double acc[4][4];
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int v = 0; v < 4; ++v) acc[u][v] = 0.0;
#pragma unroll 8
  for (int m = 0; m < NB; ++m) {
    const double* ar = As + m * AP + r0;
    const double* bc = Bs + m * AP + c0;
    double a4[4] = {ar[0], ar[1], ar[2], ar[3]};
    double b4[4] = {bc[0], bc[1], bc[2], bc[3]};
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int v = 0; v < 4; ++v) acc[u][v] += a4[u] * b4[v];
  }
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    double* row = S + (size_t)(ibase + r0 + u) * ld + jbase + c0;
#pragma unroll
    for (int v = 0; v < 4; ++v)
      if (ib != jb || c0 + v <= r0 + u) row[v] -= acc[u][v];
  }
}

int launch_factor(psba_ctx* c) {
  const int nbk = c->n_pad / NB;
  double* ok = c->d_scal + SC_FACTOR_OK;
  probe_begin(c, PSBA_PROBE_FACTOR);
  for (int kb = 0; kb < nbk; ++kb) {
    hipLaunchKernelGGL(k_potrf64, dim3(1), dim3(256), 0, c->stream, c->S, c->n_pad, kb, ok);
    const int T = nbk - kb - 1;
    hipLaunchKernelGGL(k_trsm64, dim3(T + 1), dim3(256), 0, c->stream, c->S, c->n_pad, kb, T, c->ea, ok);
    if (T > 0) {
      const int ntiles = T * (T + 1) / 2;
      hipLaunchKernelGGL(k_syrk64, dim3(ntiles + T), dim3(256), 0, c->stream, c->S, c->n_pad, kb, T, ntiles, c->ea,
                         ok);
    }
  }
  probe_end(c, PSBA_PROBE_FACTOR);
  PSBA_LAUNCH_TRY();
  return PSBA_OK;
}

// Back substitution L^T x = y (y = forward-substituted right-hand side).
// One block of 1024 threads (16 waves); x lives in LDS.  The rows below each
// diagonal tile are streamed by all waves (4 independent row loads in flight
// per wave); the 64-step triangular solve of the tile is done by wave 0 with
// the tile staged in LDS.  Also writes dp[0..n).
__global__ __launch_bounds__(1024) void k_bsolve64(const double* __restrict__ L, int ld, int nbk,
                                                   const double* __restrict__ y, double* __restrict__ xout,
                                                   double* __restrict__ dp, int n, const double* ok) {
  constexpr int NW = 16;
  extern __shared__ double xs[];
  __shared__ double part[NW][NB];
  __shared__ double Lt[NB * TP];
  if (*ok == 0.0) return;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int np_ = nbk * NB;
  for (int q = threadIdx.x; q < np_; q += blockDim.x) xs[q] = y[q];
  __syncthreads();
  for (int kb = nbk - 1; kb >= 0; --kb) {
    const int base = kb * NB;
    for (int f = threadIdx.x; f < NB * NB; f += blockDim.x) {
      int r = f >> 6, c = f & 63;
      Lt[r * TP + c] = L[(size_t)(base + r) * ld + base + c];
    }
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int r = base + NB + wid;
    for (; r + 3 * NW < np_; r += 4 * NW) {
      s0 += L[(size_t)r * ld + base + lane] * xs[r];
      s1 += L[(size_t)(r + NW) * ld + base + lane] * xs[r + NW];
      s2 += L[(size_t)(r + 2 * NW) * ld + base + lane] * xs[r + 2 * NW];
      s3 += L[(size_t)(r + 3 * NW) * ld + base + lane] * xs[r + 3 * NW];
    }
    for (; r < np_; r += NW) s0 += L[(size_t)r * ld + base + lane] * xs[r];
    part[wid][lane] = (s0 + s1) + (s2 + s3);
    __syncthreads();
    if (wid == 0) {
      double acc = 0.0;
#pragma unroll
      for (int w = 0; w < NW; ++w) acc += part[w][lane];
      double br = xs[base + lane] - acc;
      double xv = 0.0;
#pragma unroll 8
      for (int m = NB - 1; m >= 0; --m) {
        const double lm = Lt[m * TP + lane];  // L[base+m][base+lane]
        const double xm = __shfl(br / lm, m, 64);
        if (lane == m) xv = xm;
        if (lane < m) br -= lm * xm;
      }
      xs[base + lane] = xv;
    }
    __syncthreads();
  }
  for (int q = threadIdx.x; q < np_; q += blockDim.x) {
    xout[q] = xs[q];
    if (q < n) dp[q] = xs[q];
  }
}

int launch_solve(psba_ctx* c) {
  const int nbk = c->n_pad / NB;
  const double* ok = c->d_scal + SC_FACTOR_OK;
  if (c->n_pad > kMaxSolveN) return PSBA_ERR_UNSUPPORTED;
  probe_begin(c, PSBA_PROBE_BSOLVE);
  hipLaunchKernelGGL(k_bsolve64, dim3(1), dim3(1024), sizeof(double) * c->n_pad, c->stream, c->S, c->n_pad, nbk,
                     c->ea, c->xsol, c->dp, c->n, ok);
  probe_end(c, PSBA_PROBE_BSOLVE);
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

// One column of the column-wise modified Cholesky (PSBA/CL_files/cholmod_blk.cl:
// 446-697): diagonal from the row's squared sum bounded below by delta, the
// column scaled by it, and rescaled by max|c| / beta when an entry exceeds beta.
__device__ void modchol_column(const double* __restrict__ A, double* __restrict__ M, int ld, int n,
                               double* __restrict__ Cc, int xj, double delta, double beta, double* sh, int* shi) {
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
  ov = block_or256(ov, shi);
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
      // row triplets below the diagonal block (the last one partial when n % 3 != 0)
      for (int i = j + 1 + threadIdx.x; 3 * i < n; i += blockDim.x) {
        const int i3 = 3 * i;
        const int nu = min(n - i3, 3);
        double Tij[9];
        for (int u = 0; u < nu; ++u)
          for (int v = 0; v < 3; ++v) {
            double s = A[(size_t)(i3 + u) * ld + j3 + v];
            for (int k = 0; k < j3; ++k) s -= M[(size_t)(i3 + u) * ld + k] * M[(size_t)(j3 + v) * ld + k];
            Tij[3 * u + v] = s;
          }
        for (int u = 0; u < nu; ++u)
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
    for (int col = 0; col < 3; ++col) modchol_column(A, M, ld, n, Cc, j3 + col, delta, beta, sh, &shi);
  }
  // columns past the last full 3-block (n not a multiple of 3): column-wise path
  for (int xj = 3 * nb; xj < n; ++xj) modchol_column(A, M, ld, n, Cc, xj, delta, beta, sh, &shi);
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
