// Dense FP64 linear algebra on the reduced camera matrix S (n_pad x n_pad,
// row-major, lower triangle; rows/cols >= n are the identity):
//   - Cholesky S = L L^T and the solve S dpa = e_a fused in one cooperative
//     launch over 64x64 tiles (k_chol_solve below); replaces the reference's
//     explicit inverse (PSBA/CL_files/SPD_inv.cl, kern_cholesky/_s2 +
//     kern_trigMat_inv/_mul + kern_fill_rest) and matVec_mul.cl;
//   - the dogleg's modified Cholesky and lambda = |sum E| / n
//     (PSBA/CL_files/cholmod_blk.cl, PSBA/cl_cholmod.cpp:109-202).
// Failure of the factorisation (a pivot that is not finite and positive) is
// recorded in d_scal[SC_FACTOR_OK].
#include <algorithm>

#include "context.hpp"

namespace psba {

constexpr int NB = kNB;   // 64
constexpr int TP = NB + 1; // padded LDS row length

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
  c->L_cur = c->S;
  return PSBA_OK;
}

// ---------------------------------------------------------------------------
// Fused Cholesky factor and solve of S dpa = e_a in ONE launch.
//
// S is viewed as T x T tiles of 64 (T = n_pad / 64).  Phase k (k = 0..T-2)
// applies panel k to the trailing matrix: for every tile (i, j) with
// k+1 <= j <= i, the owning workgroup recomputes the panel rows it needs,
//     L_ik = A_ik L_kk^-T,   L_jk = A_jk L_kk^-T        (row-parallel solves)
//     A_ij -= L_ik L_jk^T                               (FP64 MFMA 16x16x4)
// the task (i, k+1) stores L_ik, and the task (k+1, k+1) factors the new
// diagonal tile at once.  Recomputing L_ik per task trades a few redundant
// triangular solves (cheap) for a single grid barrier per phase.
// The forward substitution L y = b runs one phase behind, off the diagonal
// chain, on one more workgroup that has no tile tasks: in phase k it applies
// y_{k-1} to the rows below (b_i -= L_{i,k-1} y_{k-1}, i >= k; the panel
// L_{.,k-1} was stored in phase k-1) and forms y_k = L_kk^-1 b_k, so nothing
// else waits for it.  Workgroup 0 forms y_0 before the first barrier and the
// last y after the last one.  (With a single workgroup, the tasks (i, k+1)
// update b_i and the diagonal task forms y_{k+1}.)
// The back substitution L^T x = y runs in T more phases: workgroup 0 solves
// block k (after applying x_{k+1} to it), the others apply x_{k+1} to the
// blocks above.
//
// The factor goes to Lf (S is overwritten by the trailing updates); e_a is
// left intact.  Barriers: a 64-bit arrival counter that only grows; its value
// before a launch (the epoch) is kept on the device beside the flags
// (bar[2 + T]) and advanced by workgroup 0 at the end of the launch, so no
// host bookkeeping has to follow the launches.  Every wait is bounded.
// d_scal[SC_FACTOR_OK] becomes 1 on success, 0 when a pivot is not finite
// and positive, and -1 when a wait timed out (an execution error: the
// workgroups were not all resident).  The grid is at most one workgroup per
// CU, so on the context's in-order stream every workgroup is resident at once;
// should another stream's kernels hold CUs, the waiting workgroups are
// scheduled as those finish (the bounded waits cover the worst case).
// ---------------------------------------------------------------------------
typedef double dvec4 __attribute__((ext_vector_type(4)));

// The large unrolled tile routines below are kept out of line: inlined
// together into the kernel they exceed the register file and spill.
#define PSBA_OUTLINE __device__ __attribute__((noinline))

// LDS pointers for the out-of-line routines: with a generic pointer parameter
// every LDS access would carry a generic-to-local conversion (a null test and
// a select per address); callers cast their __shared__ addresses, which folds.
typedef __attribute__((address_space(3))) double lds_dbl;
#define LDS(p) ((lds_dbl*)(p))

struct CholSmem {
  double P[NB * TP];  // L_kk of the current phase (lower, zeros above)
  double X[NB * TP];  // L_ik
  double Y[NB * TP];  // L_jk, or the diagonal tile being factored
  double rd[NB];      // 1 / diag(L_kk)
  double v[NB];       // y_k of the phase / solution block
  double w[NB];       // y_{k+1} / x_{k+1}
  double rd16[16];
  double part[4][NB];
  int bad;
};
typedef __attribute__((address_space(3))) CholSmem lds_smem;


// value of v on lane l (l a compile-time constant in the unrolled loops)
__device__ __forceinline__ double bcast(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// partner values inside a quad of lanes (DPP quad_perm, no LDS round trip)
template <int CTRL>
__device__ __forceinline__ double quad_perm(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp((int)b, CTRL, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, true);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ double quad_xor1(double v) { return quad_perm<0xB1>(v); }  // [1,0,3,2]
__device__ __forceinline__ double quad_xor2(double v) { return quad_perm<0x4E>(v); }  // [2,3,0,1]

// Grid barrier over a counter that only grows: every workgroup adds one and
// waits until the counter reaches `target`.  The add returns the count, so the
// workgroup that arrives last (workgroup 0, which carries the diagonal chain)
// sees at once that the others are in and does not poll; the others poll
// with a bounded wait (~1 s, then fail bit 2; never reached when every
// workgroup of the launch is resident).
__device__ __forceinline__ void grid_barrier(unsigned long long* bar, unsigned long long target, int* fail) {
  __syncthreads();
  if (threadIdx.x == 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    const unsigned long long arrived =
        __hip_atomic_fetch_add(bar, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1;
    if (arrived < target) {
      unsigned spins = 0;
      while (__hip_atomic_load(bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
        __builtin_amdgcn_s_sleep(2);
        if (++spins > (1u << 24)) {
          __hip_atomic_fetch_or(fail, 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          break;
        }
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  __syncthreads();
}

// Point-to-point flags (back substitution).  A flag only grows, across
// launches too: a launch's values start above its epoch (its barrier base),
// which exceeds every value an earlier launch wrote, so nothing is reset.
// Publish: the calling wave's earlier stores become visible first (thread 0
// only, after the caller's __syncthreads when other waves stored).
__device__ __forceinline__ void flag_publish(unsigned long long* f, unsigned long long v) {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  __hip_atomic_store(f, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Wait until *f >= v (thread 0 polls, bounded like the grid barrier), then the whole
// workgroup may read what was published before.  `seen` is a value of *f that
// thread 0 read earlier (0 if none): when it already reaches v, thread 0 does
// not read the flag again and only takes the acquire fence.
__device__ __forceinline__ void flag_wait(unsigned long long* f, unsigned long long v, int* fail,
                                          unsigned long long seen = 0) {
  if (threadIdx.x == 0) {
    if (seen < v) {
      unsigned spins = 0;
      while (__hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < v) {
        __builtin_amdgcn_s_sleep(2);
        if (++spins > (1u << 24)) {
          __hip_atomic_fetch_or(fail, 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          break;
        }
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  __syncthreads();
}

// Tile transfers for the 256-thread kernels of this file: thread t moves the
// elements t + 256 q (q = 0..15) of a 64x64 tile.  All 16 loads are issued
// before the first LDS store, so a transfer waits for one memory latency.
__device__ __forceinline__ void tile_regs(double (&v)[16], const double* src, int ld) {
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int f = (int)threadIdx.x + 256 * q;
    v[q] = src[(size_t)(f >> 6) * ld + (f & 63)];
  }
}

// v -> LDS tile (stride TP); with `lower`, zeros above the diagonal
__device__ __forceinline__ void regs_to_lds(double* dst, const double (&v)[16], bool lower) {
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int f = (int)threadIdx.x + 256 * q, r = f >> 6, c = f & 63;
    dst[r * TP + c] = (!lower || c <= r) ? v[q] : 0.0;
  }
}

// lower triangle of a global tile into LDS (zeros above the diagonal)
__device__ __forceinline__ void load_lower(double* dst, const double* src, int ld) {
  double v[16];
  tile_regs(v, src, ld);
  regs_to_lds(dst, v, true);
}

// a full global tile into LDS (stride TP)
__device__ __forceinline__ void load_full(double* dst, const double* src, int ld) {
  double v[16];
  tile_regs(v, src, ld);
  regs_to_lds(dst, v, false);
}

// A tile by waves 1-3 only (192 threads, element f = tid - 64 + 192 q,
// q < 22), while wave 0 runs a tile solve: global -> registers, all loads
// issued before the first is used, and registers -> LDS (stride TP; with
// `lower`, zeros above the diagonal).
__device__ __forceinline__ void tile_regs_w123(double (&v)[22], const double* src, int ld) {
#pragma unroll
  for (int q = 0; q < 22; ++q) {
    const int f = (int)threadIdx.x - 64 + 192 * q;
    v[q] = (f < NB * NB) ? src[(size_t)(f >> 6) * ld + (f & 63)] : 0.0;
  }
}

__device__ __forceinline__ void regs_to_lds_w123(double* dst, const double (&v)[22], bool lower) {
#pragma unroll
  for (int q = 0; q < 22; ++q) {
    const int f = (int)threadIdx.x - 64 + 192 * q, r = f >> 6, c = f & 63;
    if (f < NB * NB) dst[r * TP + c] = (!lower || c <= r) ? v[q] : 0.0;
  }
}

// store_tile by waves 1-3 only (wave 0 is free for a concurrent tile solve)
__device__ __forceinline__ void store_tile_w123(const double* src, double* dst, int ld) {
  for (int f = (int)threadIdx.x - 64; f >= 0 && f < NB * NB; f += (int)blockDim.x - 64) {
    const int r = f >> 6, c = f & 63;
    dst[(size_t)r * ld + c] = src[r * TP + c];
  }
}

__device__ __forceinline__ void store_tile(const double* src, double* dst, int ld) {
  for (int f = threadIdx.x; f < NB * NB; f += blockDim.x) {
    const int r = f >> 6, c = f & 63;
    dst[(size_t)r * ld + c] = src[r * TP + c];
  }
}

// value of v on lane G0 of each quad (DPP quad_perm [G0, G0, G0, G0])
template <int G0>
__device__ __forceinline__ double quad_bcast(double v) {
  return quad_perm<G0 * 0x55>(v);
}

// X = A L^-T in place (A an LDS tile, L = sm.P lower with zeros above): rows
// are independent; wave w owns rows 16w..16w+15, the four lanes of a quad
// share a row and hold its columns c = 4q + g (g = lane % 4).  Branch-free;
// no workgroup barriers.
PSBA_OUTLINE void trsm_rows(lds_dbl* out, const lds_dbl* L, const lds_dbl* rd) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int r = 16 * w + (lane >> 2), g = lane & 3;
  const int li = lane & 15, lk = lane >> 4;
  double x[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) x[q] = out[r * TP + 4 * q + g];
  // Four blocks of 16 columns.  Inside block b, column c: the owner lane
  // (g == c % 4) scales its accumulated value, the quad shares it, and the
  // block's later columns take their terms (right-looking: a multiply, a quad
  // broadcast and one FMA on the chain).  The block then goes to LDS and the
  // later blocks take X_b L_{b',b}^T by FP64 MFMA (the wave's own 16 rows, so
  // no workgroup barrier), through LDS back into the quad layout.
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    // this lane's rows of the diagonal block L_bb, all loaded up front (the
    // column chain then waits on no LDS access)
    double rdv[4], Lb[4][16];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      rdv[q] = rd[16 * b + 4 * q + g];
#pragma unroll
      for (int cb = 0; cb < 16; ++cb) Lb[q][cb] = (4 * q + g > cb) ? L[(16 * b + 4 * q + g) * TP + 16 * b + cb] : 0.0;
    }
#pragma unroll
    for (int cb = 0; cb < 16; ++cb) {
      const int q0 = cb >> 2, g0 = cb & 3;
      const double xo = x[4 * b + q0] * rdv[q0];
      double xc;
      if (g0 == 0) xc = quad_bcast<0>(xo);
      else if (g0 == 1) xc = quad_bcast<1>(xo);
      else if (g0 == 2) xc = quad_bcast<2>(xo);
      else xc = quad_bcast<3>(xo);
      const double fm = fma(-xc, Lb[q0][cb], x[4 * b + q0]);
      x[4 * b + q0] = (g == g0) ? xc : ((g > g0) ? fm : x[4 * b + q0]);
#pragma unroll
      for (int q = q0 + 1; q < 4; ++q) x[4 * b + q] = fma(-xc, Lb[q][cb], x[4 * b + q]);
#pragma unroll
      for (int q = q0; q < 4; ++q) __asm__ volatile("" : "+v"(x[4 * b + q]));
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) out[r * TP + 16 * b + 4 * q + g] = x[4 * b + q];
    if (b < 3) {
      // P = X_b L_{b',b}^T for every later block b' (MFMA layout: register k of
      // lane l is row 4k + l/16, column l%16), stored over block b' of `out`
      // and read back in the quad layout
#pragma unroll
      for (int bp = b + 1; bp < 4; ++bp) {
        dvec4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int k0 = 0; k0 < 16; k0 += 4) {
          const double av = out[(16 * w + li) * TP + 16 * b + k0 + lk];
          const double bv = L[(16 * bp + li) * TP + 16 * b + k0 + lk];
          acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) out[(16 * w + 4 * k + lk) * TP + 16 * bp + li] = acc[k];
#pragma unroll
        for (int q = 0; q < 4; ++q) x[4 * bp + q] -= out[r * TP + 16 * bp + 4 * q + g];
      }
    }
  }
}

// The diagonal task's pair of steps in one routine: out = A_ik L^-T as
// trsm_rows, and D = A_ii - out out^T (A_ii from global, ld; only the lower
// quadrants updated, as acc_update with lower_only), with the product's
// k-steps for a 16-column block of out taken as soon as every wave has
// finished that block (one workgroup barrier per block), so that they run
// beside the next block's column chain.  Per accumulator the MFMA sequence is
// acc_update's, in the same order, so D is the same bits.
PSBA_OUTLINE void trsm_syrk(lds_dbl* out, const lds_dbl* L, const lds_dbl* rd, const double* A, int ld,
                            lds_dbl* D) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int r = 16 * w + (lane >> 2), g = lane & 3;
  const int li = lane & 15, lk = lane >> 4;
  const int R0 = 32 * (w >> 1), C0 = 32 * (w & 1);
  const bool upd = C0 <= R0;
  // A read as global memory: through a generic pointer these would be flat
  // loads, which count against the LDS wait counter too, so the first LDS
  // wait of the column chain would wait for them
  const __attribute__((address_space(1))) double* Ag = (const __attribute__((address_space(1))) double*)A;
  dvec4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int bq = 0; bq < 2; ++bq)
#pragma unroll
      for (int rr = 0; rr < 4; ++rr)
        acc[a][bq][rr] = Ag[(size_t)(R0 + 16 * a + 4 * rr + lk) * ld + C0 + 16 * bq + li];
  double x[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) x[q] = out[r * TP + 4 * q + g];
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    double rdv[4], Lb[4][16];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      rdv[q] = rd[16 * b + 4 * q + g];
#pragma unroll
      for (int cb = 0; cb < 16; ++cb) Lb[q][cb] = (4 * q + g > cb) ? L[(16 * b + 4 * q + g) * TP + 16 * b + cb] : 0.0;
    }
#pragma unroll
    for (int cb = 0; cb < 16; ++cb) {
      const int q0 = cb >> 2, g0 = cb & 3;
      const double xo = x[4 * b + q0] * rdv[q0];
      double xc;
      if (g0 == 0) xc = quad_bcast<0>(xo);
      else if (g0 == 1) xc = quad_bcast<1>(xo);
      else if (g0 == 2) xc = quad_bcast<2>(xo);
      else xc = quad_bcast<3>(xo);
      const double fm = fma(-xc, Lb[q0][cb], x[4 * b + q0]);
      x[4 * b + q0] = (g == g0) ? xc : ((g > g0) ? fm : x[4 * b + q0]);
#pragma unroll
      for (int q = q0 + 1; q < 4; ++q) x[4 * b + q] = fma(-xc, Lb[q][cb], x[4 * b + q]);
#pragma unroll
      for (int q = q0; q < 4; ++q) __asm__ volatile("" : "+v"(x[4 * b + q]));
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) out[r * TP + 16 * b + 4 * q + g] = x[4 * b + q];
    __syncthreads();  // block b of out complete for every row
    if (upd) {
#pragma unroll
      for (int k0 = 16 * b; k0 < 16 * b + 16; k0 += 4) {
        const double a0 = -out[(R0 + li) * TP + k0 + lk];
        const double a1 = -out[(R0 + 16 + li) * TP + k0 + lk];
        const double b0 = out[(C0 + li) * TP + k0 + lk];
        const double b1 = out[(C0 + 16 + li) * TP + k0 + lk];
        acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
        acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
        acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
        acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
      }
    }
    if (b < 3) {
#pragma unroll
      for (int bp = b + 1; bp < 4; ++bp) {
        dvec4 t = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int k0 = 0; k0 < 16; k0 += 4) {
          const double av = out[(16 * w + li) * TP + 16 * b + k0 + lk];
          const double bv = L[(16 * bp + li) * TP + 16 * b + k0 + lk];
          t = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, t, 0, 0, 0);
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) out[(16 * w + 4 * k + lk) * TP + 16 * bp + li] = t[k];
#pragma unroll
        for (int q = 0; q < 4; ++q) x[4 * bp + q] -= out[r * TP + 16 * bp + 4 * q + g];
      }
    }
  }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int bq = 0; bq < 2; ++bq)
#pragma unroll
      for (int rr = 0; rr < 4; ++rr) D[(R0 + 16 * a + 4 * rr + lk) * TP + C0 + 16 * bq + li] = acc[a][bq][rr];
}

// value of v on lane N of each 16-lane row (DPP row_newbcast; CDNA's 64-bit
// DPP takes this control, so the move is one v_mov_b64_dpp, not two halves)
template <int N>
__device__ __forceinline__ double row_bcast(double v) {
  return __builtin_amdgcn_mov_dpp(v, 0x150 + N, 0xF, 0xF, true);
}

// 1/x for a normal, finite x: hardware reciprocal estimate refined by two
// Newton steps (relative error ~1 ulp); keeps the division's long dependent
// sequence off the pivot chain
__device__ __forceinline__ double fast_rcp(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = fma(r, fma(-x, r, 1.0), r);
  r = fma(r, fma(-x, r, 1.0), r);
  return r;
}

template <int C, int Q>
struct Potrf16Row {
  static __device__ __forceinline__ void update(double (&a)[16], double ur) {
    a[Q] = fma(-ur, row_bcast<Q>(a[C]), a[Q]);
    Potrf16Row<C, Q + 1>::update(a, ur);
  }
};
template <int C>
struct Potrf16Row<C, 16> {
  static __device__ __forceinline__ void update(double (&)[16], double) {}
};

template <int C>
struct Potrf16Step {
  static __device__ __forceinline__ void run(double (&a)[16], int lane, __attribute__((address_space(3))) int* bad) {
    const double piv = row_bcast<C>(a[C]);
    if (lane == 0 && !(piv > 0.0)) *bad = 1;
    const double ur = a[C] * fast_rcp(piv);
    Potrf16Row<C, C + 1>::update(a, ur);
    Potrf16Step<C + 1>::run(a, lane, bad);
  }
};
template <>
struct Potrf16Step<16> {
  static __device__ __forceinline__ void run(double (&)[16], int, __attribute__((address_space(3))) int*) {}
};

template <int Q>
struct Potrf16Scale {
  static __device__ __forceinline__ void run(const double (&a)[16], lds_dbl* B, int r, int lane, double d, double rdl) {
    const double rq = row_bcast<Q>(rdl);
    const double v = (Q < r) ? a[Q] * rq : ((Q == r) ? d : 0.0);
    if (lane < 16) B[r * TP + Q] = v;
    Potrf16Scale<Q + 1>::run(a, B, r, lane, d, rdl);
  }
};
template <>
struct Potrf16Scale<16> {
  static __device__ __forceinline__ void run(const double (&)[16], lds_dbl*, int, int, double, double) {}
};

// the factor's row r in place of the LDL^T row a, left of the diagonal (what
// Potrf16Scale writes there); the entries from the diagonal on are left for
// the caller (a[Q] * rq there is not the diagonal's d, nor zero above it)
template <int Q>
struct Potrf16Rows {
  static __device__ __forceinline__ void run(double (&a)[16], double rdl) {
    a[Q] = a[Q] * row_bcast<Q>(rdl);
    Potrf16Rows<Q + 1>::run(a, rdl);
  }
};
template <>
struct Potrf16Rows<16> {
  static __device__ __forceinline__ void run(double (&)[16], double) {}
};

// 16x16 Cholesky of B (LDS, stride TP) by one wave, lane = row (the four
// 16-lane rows of the wave carry identical copies).  Outer-product LDL^T with
// the pivot column broadcast by DPP row_newbcast (one reciprocal per pivot on
// the dependent chain), then L = U sqrt(D) with one sqrt per lane.
// rd16 = 1 / diag(L).
__device__ __forceinline__ void potrf16(lds_dbl* B, lds_dbl* rd16, __attribute__((address_space(3))) int* bad) {
  const int lane = threadIdx.x & 63;
  const int r = lane & 15;
  double a[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) a[q] = B[r * TP + q];
  Potrf16Step<0>::run(a, lane, bad);
  double dd = a[0];
#pragma unroll
  for (int q = 1; q < 16; ++q) dd = (r == q) ? a[q] : dd;
  const double d = sqrt(dd), rdl = 1.0 / d;
  Potrf16Scale<0>::run(a, B, r, lane, d, rdl);
  if (lane < 16) rd16[r] = rdl;
}

// The panel solve below the 16x16 block, in registers: x = A_r L_bb^-T for
// this lane's panel row, right-looking (x_c *= 1/L_cc, then x_m -= x_c L_mc
// for m > c).  1/L_cc and L_mc come from the lanes that hold row c / row m
// of L_bb in the lane's own 16-lane row (DPP row_newbcast), not from LDS.
template <int C, int M>
struct PanelRow {
  static __device__ __forceinline__ void update(double (&x)[16], const double (&Lr)[16]) {
    x[M] = fma(-x[C], row_bcast<M>(Lr[C]), x[M]);
    PanelRow<C, M + 1>::update(x, Lr);
  }
};
template <int C>
struct PanelRow<C, 16> {
  static __device__ __forceinline__ void update(double (&)[16], const double (&)[16]) {}
};
template <int C>
struct PanelStep {
  static __device__ __forceinline__ void run(double (&x)[16], double (&Lr)[16], double rdl) {
    // one step at a time (the empty asm pins Lr[C] here and x after the
    // step): hoisting every step's broadcasts would hold 120 values in
    // registers
    __asm__ volatile("" : "+v"(Lr[C]));
    x[C] = x[C] * row_bcast<C>(rdl);
    PanelRow<C, C + 1>::update(x, Lr);
#pragma unroll
    for (int m = C; m < 16; ++m) __asm__ volatile("" : "+v"(x[m]));
    PanelStep<C + 1>::run(x, Lr, rdl);
  }
};
template <>
struct PanelStep<16> {
  static __device__ __forceinline__ void run(double (&)[16], double (&)[16], double) {}
};

// potrf16 on the diagonal block B (LDS, stride TP; its upper part is left
// unspecified, tile_potrf clears it) and the panel solve of the R rows below
// it, by one wave: lane l takes panel row 16 + l (l < R), whose
// entries are loaded before the factor starts; every 16-lane row of the wave
// keeps its copy of the factor (row l % 16 of L_bb) for the panel's
// broadcasts.  The same operations in the same order as potrf16 followed by
// the per-row panel solve from LDS, so the same results bit for bit.
__device__ __forceinline__ void potrf16_panel(lds_dbl* B, int R, lds_dbl* rd16,
                                              __attribute__((address_space(3))) int* bad) {
  const int lane = threadIdx.x & 63;
  const int r = lane & 15;
  double x[16];
#pragma unroll
  for (int c = 0; c < 16; ++c) x[c] = (lane < R) ? B[(16 + lane) * TP + c] : 0.0;
  double a[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) a[q] = B[r * TP + q];
  Potrf16Step<0>::run(a, lane, bad);
  double dd = a[0];
#pragma unroll
  for (int q = 1; q < 16; ++q) dd = (r == q) ? a[q] : dd;
  const double d = sqrt(dd), rdl = 1.0 / d;
  Potrf16Rows<0>::run(a, rdl);  // a = row r of L_bb left of the diagonal
  if (lane < 16) {
    // the diagonal is d; what lands above it is cleared at the end of
    // tile_potrf, and nothing reads it before
#pragma unroll
    for (int q = 0; q < 16; ++q) B[r * TP + q] = a[q];
    B[r * TP + r] = d;
    rd16[r] = rdl;
  }
  PanelStep<0>::run(x, a, rdl);
  if (lane < R) {
#pragma unroll
    for (int c = 0; c < 16; ++c) B[(16 + lane) * TP + c] = x[c];
  }
}

// Cholesky of the 64x64 LDS tile Yt (lower part), in place, zeros above; with
// write_rd also 1 / diag(L) into sm.rd.
// Right-looking over 16-column blocks: the diagonal block and the panel below
// it (one wave, potrf16_panel), the trailing lower blocks (MFMA, K = 16): the
// next panel's at once, the others while wave 0 factors the next block.
// Only the first nkb blocks are eliminated: the rows and columns from 16 * nkb
// on must be padding (unit diagonal, zeros elsewhere in the lower part), which
// the first blocks' panels and updates leave as they are (+0 products added to
// +0 and 1.0), so their factor is the identity already and 1/diag is 1.
PSBA_OUTLINE int tile_potrf(lds_smem& sm, lds_dbl* Yt, bool write_rd, int nkb) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int li = lane & 15, lk = lane >> 4;
  if (threadIdx.x == 0) sm.bad = 0;
  if (write_rd)
    for (int q = 16 * nkb + (int)threadIdx.x; q < NB; q += blockDim.x) sm.rd[q] = 1.0;
  __syncthreads();
  // block (ib, jc) -= L_{ib,cb} L_{jc,cb}^T, block column cb = b0 / 16 (MFMA, K = 16)
  auto update = [&](int ib, int jc, int b0) {
    dvec4 acc;
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) acc[rr] = Yt[(16 * ib + 4 * rr + lk) * TP + 16 * jc + li];
#pragma unroll
    for (int k0 = 0; k0 < 16; k0 += 4) {
      const double av = -Yt[(16 * ib + li) * TP + b0 + k0 + lk];
      const double bv = Yt[(16 * jc + li) * TP + b0 + k0 + lk];
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
    }
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) Yt[(16 * ib + 4 * rr + lk) * TP + 16 * jc + li] = acc[rr];
  };
#pragma unroll 1
  for (int kb = 0; kb < nkb; ++kb) {
    const int b0 = 16 * kb;
    if (w == 0) {
      potrf16_panel(Yt + b0 * TP + b0, 48 - b0, sm.rd16, &sm.bad);
    } else if (kb >= 1) {
      // while wave 0 factors, waves 1-3 apply block column kb - 1 to the
      // blocks right of this panel (columns kb + 1 on), which potrf16_panel
      // neither reads nor writes; each block still takes block column kb - 1
      // before block column kb
      const int m = nkb - 1 - kb, nblk = m * (m + 1) / 2;
      for (int bi = w - 1; bi < nblk; bi += 3) {
        int jb = 0, cum = 0;
        while (cum + (m - jb) <= bi) { cum += m - jb; ++jb; }
        update(kb + 1 + jb + (bi - cum), kb + 1 + jb, b0 - 16);
      }
    }
    __syncthreads();
    if (write_rd && threadIdx.x < 16) sm.rd[b0 + threadIdx.x] = sm.rd16[threadIdx.x];
    // the next panel (block column kb + 1) now; the blocks right of it wait
    // for the next factor step
    for (int ib = kb + 1 + w; ib < nkb; ib += 4) update(ib, kb + 1, b0);
    __syncthreads();
  }
  for (int f = threadIdx.x; f < NB * NB; f += blockDim.x) {
    const int rr = f >> 6, c = f & 63;
    if (c > rr) Yt[rr * TP + c] = 0.0;
  }
  __syncthreads();
  return sm.bad;
}

// acc (MFMA layout: reg r of lane l is row 4r + l/16, column l%16 of its 16x16
// block; wave w owns the 32x32 quadrant rows 32*(w>>1), cols 32*(w&1)) <- C
__device__ __forceinline__ void acc_load(dvec4 (&acc)[2][2], const double* C, int ld) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int R0 = 32 * (w >> 1), C0 = 32 * (w & 1);
  const int li = lane & 15, lk = lane >> 4;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[a][b][r] = C[(size_t)(R0 + 16 * a + 4 * r + lk) * ld + C0 + 16 * b + li];
}

__device__ __forceinline__ void acc_store(const dvec4 (&acc)[2][2], double* dst, int ld) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int R0 = 32 * (w >> 1), C0 = 32 * (w & 1);
  const int li = lane & 15, lk = lane >> 4;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) dst[(size_t)(R0 + 16 * a + 4 * r + lk) * ld + C0 + 16 * b + li] = acc[a][b][r];
}

// acc -= X Y^T (LDS tiles, K = 64); `lower_only` leaves the upper quadrant
__device__ __forceinline__ void acc_update(dvec4 (&acc)[2][2], const double* X, const double* Y, bool lower_only) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int R0 = 32 * (w >> 1), C0 = 32 * (w & 1);
  const int li = lane & 15, lk = lane >> 4;
  if (lower_only && C0 > R0) return;
#pragma unroll 4
  for (int k0 = 0; k0 < NB; k0 += 4) {
    const double a0 = -X[(R0 + li) * TP + k0 + lk];
    const double a1 = -X[(R0 + 16 + li) * TP + k0 + lk];
    const double b0 = Y[(C0 + li) * TP + k0 + lk];
    const double b1 = Y[(C0 + 16 + li) * TP + k0 + lk];
    acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
    acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
    acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
    acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
  }
}

// A tile task's inputs: A_ik (and A_jk) into sm.X (sm.Y) for the in-place
// solves and A_ij into the accumulators, every load issued before the first
// is waited for
__device__ __forceinline__ void task_load(dvec4 (&acc)[2][2], CholSmem& sm, const double* Aij, const double* Aik,
                                          const double* Ajk, int ld) {
  double va[16], vb[16];
  tile_regs(va, Aik, ld);
  if (Ajk) tile_regs(vb, Ajk, ld);
  acc_load(acc, Aij, ld);
  regs_to_lds(sm.X, va, false);
  if (Ajk) regs_to_lds(sm.Y, vb, false);
  __syncthreads();
}

// y = L^-1 b (L = lower tile in LDS), wave 0, lane = row; the lane's row of
// L is loaded first, the solved value is broadcast by readlane, and every
// lane updates unconditionally (a lane past its own row no longer matters,
// entries above the diagonal are zero).  y -> yg (global) and ys (LDS).
// The steps run in 16-row blocks, the blocks from row nr on only if nr does
// not cover them: rows from nr on must be padding (unit diagonal, zero
// coupling, zero right-hand side), whose steps give +0 and change no other
// row, so skipping them gives the same values.
PSBA_OUTLINE void tile_forward(const lds_dbl* L, const lds_dbl* b, double* yg, lds_dbl* ys, int nr) {
  if (threadIdx.x >= 64) return;
  const int lane = threadIdx.x;
  double Lr[NB];
#pragma unroll
  for (int m = 0; m < NB; ++m) Lr[m] = L[lane * TP + m];
  double br = b[lane], yv = 0.0;
  const double rdl = 1.0 / L[lane * TP + lane];
#pragma unroll
  for (int mb = 0; mb < 4; ++mb) {
    if (16 * mb >= nr) break;
#pragma unroll
    for (int m = 16 * mb; m < 16 * mb + 16; ++m) {
      const double ym = bcast(br * rdl, m);
      yv = (lane == m) ? ym : yv;
      br = fma(-Lr[m], ym, br);
    }
  }
  yg[lane] = yv;
  ys[lane] = yv;
}

// x = L^-T r (L lower tile in LDS), wave 0, same scheme with the lane's column;
// the padding rows from nr on (as for tile_forward) come first and are skipped
PSBA_OUTLINE void tile_backward(const lds_dbl* L, double br, double* xg, lds_dbl* xsm, int nr) {
  if (threadIdx.x >= 64) return;
  const int lane = threadIdx.x;
  double Lc[NB];
#pragma unroll
  for (int m = 0; m < NB; ++m) Lc[m] = L[m * TP + lane];
  double xv = 0.0;
  const double rdl = 1.0 / L[lane * TP + lane];
#pragma unroll
  for (int mb = 3; mb >= 0; --mb) {
    if (16 * mb >= nr) continue;
#pragma unroll
    for (int m = 16 * mb + 15; m >= 16 * mb; --m) {
      const double xm = bcast(br * rdl, m);
      xv = (lane == m) ? xm : xv;
      br = fma(-Lc[m], xm, br);
    }
  }
  xg[lane] = xv;
  xsm[lane] = xv;
}

// part[w][c] = sum over rows rr = w, w+4, .. of Lt[rr][c] * x[rr]; returns the
// fixed-order total for lane c (valid on all waves after the barrier inside)
__device__ __forceinline__ double tile_tmatvec(CholSmem& sm, const double* Lt, int ld, const double* x) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  double s = 0.0;
#pragma unroll
  for (int rr = w; rr < NB; rr += 4) s += Lt[(size_t)rr * ld + lane] * x[rr];
  sm.part[w][lane] = s;
  __syncthreads();
  const double t = ((sm.part[0][lane] + sm.part[1][lane]) + sm.part[2][lane]) + sm.part[3][lane];
  __syncthreads();
  return t;
}

// s = sum_c Lt[r][c] x[c] for row r = lane of a 64x64 tile, in a fixed order
__device__ __forceinline__ double tile_matvec(CholSmem& sm, const double* Lt, int ld, const double* x) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  double s = 0.0;
#pragma unroll
  for (int c = 16 * w; c < 16 * w + 16; ++c) s += Lt[(size_t)lane * ld + c] * x[c];
  sm.part[w][lane] = s;
  __syncthreads();
  const double t = ((sm.part[0][lane] + sm.part[1][lane]) + sm.part[2][lane]) + sm.part[3][lane];
  __syncthreads();
  return t;
}

// One forward step of k_chol_solve's forward workgroup in phase k >= 1:
// b_i -= L_{i,k-1} y_{k-1} for i = k..T-1 (y_{k-1} in sm.w; from yv at k = 1,
// where workgroup 0 formed it), then y_k = L_kk^-1 b_k into yv and sm.w.
// Thread (lane, w) takes row `lane`, columns 16w..16w+15 of each tile; four
// tiles' loads are in flight together, their partial sums meet in sm.X and
// are added as (p0 + p1) + (p2 + p3) -- the single-workgroup path's order.
__device__ void forward_step(CholSmem& sm, const double* __restrict__ Lf, int ld, int T, int k,
                             double* __restrict__ bw, double* __restrict__ yv) {
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  if (k == 1) {
    if (tid < NB) sm.w[tid] = yv[tid];
    __syncthreads();
  }
  double* part = sm.X;  // [4 tiles][4 waves][NB]
  for (int i = k; i < T; i += 4) {
    const int nt = min(4, T - i);
    double l[4][16];
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (u < nt) {
        const double* Lr = Lf + (size_t)(i + u) * NB * ld + (size_t)(k - 1) * NB + (size_t)lane * ld + 16 * w;
#pragma unroll
        for (int c = 0; c < 16; ++c) l[u][c] = Lr[c];
      }
    double b[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) b[u] = (u < nt && tid < NB) ? bw[(size_t)(i + u) * NB + tid] : 0.0;
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (u < nt) {
        double s = 0.0;
#pragma unroll
        for (int c = 0; c < 16; ++c) s += l[u][c] * sm.w[16 * w + c];
        part[(4 * u + w) * NB + lane] = s;
      }
    __syncthreads();
    if (tid < NB) {
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (u < nt) {
          const double* pu = part + 4 * u * NB;
          const double bi = b[u] - ((pu[tid] + pu[NB + tid]) + (pu[2 * NB + tid] + pu[3 * NB + tid]));
          bw[(size_t)(i + u) * NB + tid] = bi;
          if (i + u == k) sm.v[tid] = bi;
        }
    }
    __syncthreads();
  }
  load_lower(sm.P, Lf + (size_t)k * NB * ld + (size_t)k * NB, ld);
  __syncthreads();
  tile_forward(LDS(sm.P), LDS(sm.v), yv + (size_t)k * NB, LDS(sm.w), NB);  // k < T - 1: no padding
  __syncthreads();
}

__global__ __launch_bounds__(256) void k_chol_solve(double* __restrict__ A, int ld, int T, double* __restrict__ Lf,
                                                    const double* __restrict__ ea, double* __restrict__ bw,
                                                    double* __restrict__ yv, double* __restrict__ xs,
                                                    double* __restrict__ dp, int n, double* __restrict__ ok,
                                                    unsigned long long* bar, int* __restrict__ fail,
                                                    const unsigned char* __restrict__ gate) {
  __shared__ CholSmem sm;
  const int G = gridDim.x;
  // with gate (device-side LM control): a trial queued behind the one that
  // stopped the solve ends before anything is factored or any workgroup
  // arrives at a barrier (the counters, epoch and fail word untouched); the
  // gate is read now and tested once workgroup 0's first tile is loaded
  const bool run = !gate || *gate;
  // read before this workgroup's first arrival, so before workgroup 0 (which
  // passes every barrier last) advances it at the end
  const unsigned long long bar_base =
      __hip_atomic_load(bar + 2 + T, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const int tid = threadIdx.x;
  unsigned long long nbar = 0;
  auto tile = [&](double* base, int i, int j) { return base + (size_t)i * NB * ld + (size_t)j * NB; };
  // rows of tile i below n (the rest, in the last tile, are identity padding)
  auto rows_of = [&](int i) { return min(NB, n - i * NB); };
  auto nkb_of = [&](int i) { return (rows_of(i) + 15) >> 4; };
  // workgroup 0 carries the diagonal chain: when the diagonal task is its last
  // task of the phase it factors straight into sm.P (with 1/diag in sm.rd) and
  // keeps y_k in sm.w, ready for the next phase
  bool carry = false;
  // a workgroup beyond every phase's task count: the forward substitution
  const bool fwd = T >= 3 && G > (T - 1) * T / 2;

  // ---- phase -1: factor tile (0,0) and y_0; copy the right-hand side
  if (blockIdx.x == 0) {
    // (all sixteen loads of a thread in flight together)
    load_full(sm.P, A, ld);
    if (tid < NB) sm.part[0][tid] = ea[tid];
    __syncthreads();
    if (!run) return;
    const int bad = tile_potrf(*(lds_smem*)&sm, LDS(sm.P), true, nkb_of(0));
    if (bad && tid == 0) __hip_atomic_fetch_or(fail, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    store_tile_w123(sm.P, Lf, ld);
    tile_forward(LDS(sm.P), LDS(sm.part[0]), yv, LDS(sm.w), rows_of(0));
    carry = true;
  }
  if (!run) return;
  for (int q = NB + (int)blockIdx.x * (int)blockDim.x + tid; q < T * NB; q += G * (int)blockDim.x) bw[q] = ea[q];
  grid_barrier(bar, bar_base + (unsigned long long)G * (++nbar), fail);

  // ---- phases 0 .. T-2: factor and forward substitution
  for (int k = 0; k + 1 < T; ++k) {
    const int m = T - 1 - k, ntasks = m * (m + 1) / 2;
    if (fwd && k >= 1 && (int)blockIdx.x == G - 1) forward_step(sm, Lf, ld, T, k, bw, yv);
    if ((int)blockIdx.x < ntasks) {
      if (carry) {
        if (tid < NB) sm.v[tid] = sm.w[tid];
        __syncthreads();
      } else {
        load_lower(sm.P, tile(Lf, k, k), ld);
        if (!fwd && tid < NB) sm.v[tid] = yv[(size_t)k * NB + tid];
        __syncthreads();
        if (tid < NB) sm.rd[tid] = 1.0 / sm.P[tid * TP + tid];
        __syncthreads();
      }
    }
    carry = false;
    for (int t = blockIdx.x; t < ntasks; t += G) {
      // task t -> tile (i, j): columns j = k+1.. in order, rows i = j.. within a column
      int a = 0, cum = 0;
      while (cum + (m - a) <= t) { cum += m - a; ++a; }
      const int j = k + 1 + a, i = j + (t - cum);
      // b_i is read now so that its load overlaps the tile loads
      const int br_row = tid >> 2, br_p = tid & 3;
      const double b_old = (!fwd && j == k + 1 && br_p == 0) ? bw[(size_t)i * NB + br_row] : 0.0;
      const bool diag_next = i == j && j == k + 1;
      // the next diagonal tile: factored by this workgroup; as its last task
      // of the phase, into sm.P for the next phase
      if (diag_next) carry = (t + G >= ntasks);
      double* Dt = carry ? sm.P : sm.Y;
      dvec4 acc[2][2];
      if (diag_next) {
        // L_{k+1,k} and the update of A_{k+1,k+1} by it, in one routine
        load_full(sm.X, tile(A, i, k), ld);
        __syncthreads();
        trsm_syrk(LDS(sm.X), LDS(sm.P), LDS(sm.rd), tile(A, i, j), ld, LDS(Dt));
        __syncthreads();
      } else {
        task_load(acc, sm, tile(A, i, j), tile(A, i, k), (i != j) ? tile(A, j, k) : nullptr, ld);
        trsm_rows(LDS(sm.X), LDS(sm.P), LDS(sm.rd));
        if (i != j) trsm_rows(LDS(sm.Y), LDS(sm.P), LDS(sm.rd));
        __syncthreads();
        acc_update(acc, sm.X, (i != j) ? sm.Y : sm.X, i == j);
      }
      if (j == k + 1) {
        // store L_ik, and from the task (k+2, k+1) also L_{k+1,k} (its Y,
        // the same bits as the diagonal task's X: the same tile, L_kk and
        // reciprocals), which keeps that store off the diagonal chain;
        // b_i -= L_ik y_k (4 threads per row, fixed order)
        if (i != j) store_tile(sm.X, tile(Lf, i, k), ld);
        if (i == k + 2) store_tile(sm.Y, tile(Lf, k + 1, k), ld);
        if (!fwd) {
          double s = 0.0;
#pragma unroll
          for (int c = 0; c < 16; ++c) s += sm.X[br_row * TP + 16 * br_p + c] * sm.v[16 * br_p + c];
          s += __shfl_xor(s, 1, 64);
          s += __shfl_xor(s, 2, 64);
          if (br_p == 0) {
            const double bi = b_old - s;
            bw[(size_t)i * NB + br_row] = bi;
            if (i == j) sm.part[0][br_row] = bi;
          }
        }
      }
      if (diag_next) {
        // factor the next diagonal tile (updated into Dt above; and form
        // y_{k+1} without a forward workgroup)
        __syncthreads();
        const int bad = tile_potrf(*(lds_smem*)&sm, LDS(Dt), carry, nkb_of(i));
        if (bad && tid == 0) __hip_atomic_fetch_or(fail, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (fwd) {
          store_tile(Dt, tile(Lf, i, i), ld);
        } else {
          store_tile_w123(Dt, tile(Lf, i, i), ld);
          tile_forward(LDS(Dt), LDS(sm.part[0]), yv + (size_t)i * NB, LDS(sm.w), rows_of(i));
        }
        if (i + 1 >= T) store_tile(sm.X, tile(Lf, i, k), ld);  // no task (k+2, k+1)
      } else {
        acc_store(acc, tile(A, i, j), ld);
      }
      __syncthreads();
    }
    grid_barrier(bar, bar_base + (unsigned long long)G * (++nbar), fail);
  }

  // ---- back substitution L^T x = y: r_i = y_i - sum_{j > i} L_ji^T x_j with
  // the terms subtracted in descending j, then x_i = L_ii^-T r_i.  Workgroup 0
  // solves the chain and applies x_{i+1} to r_i itself; the other workgroups
  // ("helpers") apply x_j, j >= i+2, to the rows they own as soon as x_j is
  // published, so r_i lacks only that term one step before workgroup 0 needs
  // it.  No grid barriers: workgroup 0 publishes x_k through a flag
  // (bar[1] = epoch + T - k) and each helper publishes a completed row i
  // through its own flag (bar[2 + i] = epoch + 1).  Workgroup 0 loads the next
  // step's factor tiles and completed r while it is still working; a helper
  // keeps its first row's r in registers and loads its next tile ahead.
  const unsigned long long epoch = bar_base;
  unsigned long long* xflag = bar + 1;
  unsigned long long* rflag = bar + 2;
  if (G == 1) {
    // one workgroup: every term by workgroup 0, left-looking per block
    for (int k = T - 1; k >= 0; --k) {
      double br = yv[(size_t)k * NB + (tid & 63)];
      for (int j = T - 1; j > k; --j) {
        if (tid < NB) sm.w[tid] = xs[(size_t)j * NB + tid];
        __syncthreads();
        br -= tile_tmatvec(sm, tile(Lf, j, k), ld, sm.w);
      }
      load_lower(sm.P, tile(Lf, k, k), ld);
      __syncthreads();
      tile_backward(LDS(sm.P), br, xs + (size_t)k * NB, LDS(sm.v), rows_of(k));
      __syncthreads();
      if (tid < NB) {
        const int q = k * NB + tid;
        if (q < n) dp[q] = sm.v[tid];
      }
      __syncthreads();
    }
  } else if (blockIdx.x == 0) {
    load_lower(sm.P, tile(Lf, T - 1, T - 1), ld);
    double r_next;
    if (fwd) {
      // the last forward step: b_{T-1} holds every term but L_{T-1,T-2} y_{T-2}
      // (summed as in forward_step)
      if (tid < NB) sm.w[tid] = yv[(size_t)(T - 2) * NB + tid];
      __syncthreads();
      {
        const int lane = tid & 63, w = tid >> 6;
        const double* Lr = tile(Lf, T - 1, T - 2) + (size_t)lane * ld + 16 * w;
        double s = 0.0;
#pragma unroll
        for (int c = 0; c < 16; ++c) s += Lr[c] * sm.w[16 * w + c];
        sm.part[w][lane] = s;
      }
      __syncthreads();
      if (tid < NB)
        sm.v[tid] = bw[(size_t)(T - 1) * NB + tid] - ((sm.part[0][tid] + sm.part[1][tid]) + (sm.part[2][tid] + sm.part[3][tid]));
      __syncthreads();
      tile_forward(LDS(sm.P), LDS(sm.v), yv + (size_t)(T - 1) * NB, LDS(sm.w), rows_of(T - 1));
      __syncthreads();
      r_next = sm.w[tid & 63];
    } else {
      r_next = yv[(size_t)(T - 1) * NB + (tid & 63)];
    }
    __syncthreads();
    for (int k = T - 1; k >= 0; --k) {
      // r_{k-1} (done by the helpers but for x_k when they have terms in it):
      // its flag is read now and checked after this step's solve, so that
      // the round trip overlaps the product and the solve
      const bool need_flag = k >= 1 && k - 1 + 2 <= T - 1;
      unsigned long long fseen = 0;
      if (need_flag && tid == 0)
        fseen = __hip_atomic_load(rflag + (k - 1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      // sm.w = x_{k+1}; sm.X = L_{k+1,k}
      double br = r_next;
      if (k + 1 < T) br -= tile_tmatvec(sm, sm.X, TP, sm.w);
      __syncthreads();
      tile_backward(LDS(sm.P), br, xs + (size_t)k * NB, LDS(sm.v), rows_of(k));
      // x_k was stored by wave 0 (tile_backward): wave 0 publishes it at
      // once, while waves 1-3 load
      if (k >= 1 && tid == 0) flag_publish(xflag, epoch + (unsigned long long)(T - k));
      // Waves 1-3 return from tile_backward at once and load the next
      // step's tiles while wave 0 solves: L_{k,k-1} straight into sm.X (the
      // product above has read it), L_{k-1,k-1} into registers until wave 0
      // is done with sm.P.
      double p0[22];
      const bool loader = k >= 1 && tid >= 64;
      if (loader) {
        double p1[22];
        tile_regs_w123(p0, tile(Lf, k - 1, k - 1), ld);
        tile_regs_w123(p1, tile(Lf, k, k - 1), ld);
        regs_to_lds_w123(sm.X, p1, false);
      }
      __syncthreads();
      if (tid < NB) {
        sm.w[tid] = sm.v[tid];
        const int q = k * NB + tid;
        if (q < n) dp[q] = sm.v[tid];
      }
      if (k == 0) break;
      if (need_flag) flag_wait(rflag + (k - 1), epoch + 1, fail, fseen);
      r_next = yv[(size_t)(k - 1) * NB + (tid & 63)];
      if (loader) regs_to_lds_w123(sm.P, p0, true);
      __syncthreads();
    }
  } else {
    // helper: rows i0, i0 + stride, .. (those with i + 2 <= T - 1)
    const int i0 = (int)blockIdx.x - 1, stride = G - 1;
    if (i0 + 2 <= T - 1) {
      double r_own = (tid < NB) ? yv[(size_t)i0 * NB + tid] : 0.0;
      load_full(sm.X, tile(Lf, T - 1, i0), ld);
      __syncthreads();
      for (int j = T - 1; j >= i0 + 2; --j) {
        flag_wait(xflag, epoch + (unsigned long long)(T - j), fail);
        if (tid < NB) sm.v[tid] = xs[(size_t)j * NB + tid];
        __syncthreads();
        for (int i = i0; i + 2 <= j; i += stride) {
          if (i == i0) {
            const double t = tile_tmatvec(sm, sm.X, TP, sm.v);
            r_own -= t;
            if (i + 2 == j && tid < NB) yv[(size_t)i * NB + tid] = r_own;
          } else {
            const double t = tile_tmatvec(sm, tile(Lf, j, i), ld, sm.v);
            if (tid < NB) yv[(size_t)i * NB + tid] -= t;
          }
          if (i + 2 == j && tid == 0) flag_publish(rflag + i, epoch + 1);  // stores by wave 0
        }
        if (j - 1 >= i0 + 2) {
          load_full(sm.X, tile(Lf, j - 1, i0), ld);
          __syncthreads();
        }
      }
    }
  }
  if (blockIdx.x == 0 && tid == 0) {
    const int f = __hip_atomic_load(fail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    *ok = (f & 2) ? -1.0 : (f ? 0.0 : 1.0);
    // cleared for the next launch (every other workgroup is past its last
    // wait: workgroup 0 consumed every helper's last flag), and the epoch of
    // the next launch: this launch's T barriers of G arrivals each
    __hip_atomic_store(fail, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(bar + 2 + T, bar_base + (unsigned long long)G * T, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  }
}


// S x = b with the factor L already in Lf (one workgroup; used by the mixed
// path's FP64 refinement): forward L y = b, then back L^T x = y, tile by tile.
// b has n entries (zero-padded to T tiles here); y and x are n_pad scratch.

// LDS of k_tri_solve: the diagonal tile, and the whole solution vector (y in
// the forward pass, overwritten block by block by x in the backward pass)
struct TriSmem {
  double P[NB * TP];
  double vec[kMaxSolveN];
  double part[4][NB];
  double v[NB];
};

__global__ __launch_bounds__(256) void k_tri_solve(const double* __restrict__ Lf, int ld, int T, int n,
                                                   const double* __restrict__ b, double* __restrict__ y,
                                                   double* __restrict__ x, double* __restrict__ out) {
  __shared__ TriSmem sm;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  auto tile = [&](int i, int j) { return Lf + (size_t)i * NB * ld + (size_t)j * NB; };
  // fixed-order total of the four waves' partial sums for lane `lane`
  auto total = [&](double s) {
    sm.part[w][lane] = s;
    __syncthreads();
    const double t = ((sm.part[0][lane] + sm.part[1][lane]) + sm.part[2][lane]) + sm.part[3][lane];
    __syncthreads();
    return t;
  };
  // forward: y_i = L_ii^-1 (b_i - sum_{k<i} L_ik y_k).  Thread (lane, w) takes
  // row `lane`, columns 16w .. 16w + 15 of every tile; four tiles' loads are
  // in flight together (the solve is bound by their latency, not bandwidth).
  for (int i = 0; i < T; ++i) {
    const double* Li = tile(i, 0) + (size_t)lane * ld + 16 * w;
    double s = 0.0;
    int k = 0;
    for (; k + 4 <= i; k += 4) {
      double l[4][16];
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int c = 0; c < 16; ++c) l[u][c] = Li[(k + u) * NB + c];
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int c = 0; c < 16; ++c) s += l[u][c] * sm.vec[(k + u) * NB + 16 * w + c];
    }
    for (; k < i; ++k) {
      double l[16];
#pragma unroll
      for (int c = 0; c < 16; ++c) l[c] = Li[k * NB + c];
#pragma unroll
      for (int c = 0; c < 16; ++c) s += l[c] * sm.vec[k * NB + 16 * w + c];
    }
    const double t = total(s);
    const int q = i * NB + lane;
    load_lower(sm.P, tile(i, i), ld);
    if (tid < NB) sm.v[tid] = ((q < n) ? b[q] : 0.0) - t;
    __syncthreads();
    tile_forward(LDS(sm.P), LDS(sm.v), y + (size_t)i * NB, LDS(sm.vec) + (size_t)i * NB, NB);
    __syncthreads();
  }
  // backward: x_i = L_ii^-T (y_i - sum_{k>i} L_ki^T x_k); thread (lane, w) takes
  // column `lane`, rows w, w + 4, .. of every tile
  for (int i = T - 1; i >= 0; --i) {
    double s = 0.0;
    int k = i + 1;
    for (; k + 4 <= T; k += 4) {
      double l[4][16];
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int r = 0; r < 16; ++r) l[u][r] = tile(k + u, i)[(size_t)(w + 4 * r) * ld + lane];
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int r = 0; r < 16; ++r) s += l[u][r] * sm.vec[(k + u) * NB + w + 4 * r];
    }
    for (; k < T; ++k) {
      double l[16];
#pragma unroll
      for (int r = 0; r < 16; ++r) l[r] = tile(k, i)[(size_t)(w + 4 * r) * ld + lane];
#pragma unroll
      for (int r = 0; r < 16; ++r) s += l[r] * sm.vec[k * NB + w + 4 * r];
    }
    const double t = total(s);
    const double br = sm.vec[i * NB + lane] - t;
    load_lower(sm.P, tile(i, i), ld);
    __syncthreads();
    tile_backward(LDS(sm.P), br, x + (size_t)i * NB, LDS(sm.vec) + (size_t)i * NB, NB);
    __syncthreads();
    if (tid < NB && i * NB + tid < n) out[i * NB + tid] = sm.vec[i * NB + tid];
  }
}

int launch_tri_solve(psba_ctx* c, const double* b, double* out) {
  hipLaunchKernelGGL(k_tri_solve, dim3(1), dim3(256), 0, c->stream, c->Lf, c->n_pad, c->n_pad / NB, c->n, b, c->yv,
                     c->xsol, out);
  PSBA_LAUNCH_TRY();
  return PSBA_OK;
}

int launch_factor_solve(psba_ctx* c) {
  const int T = c->n_pad / NB;
  if (c->chol_grid <= 0) {
    int per_cu = 0, ncu = 0;
    PSBA_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_chol_solve, 256, 0));
    PSBA_HIP_TRY(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device));
    // at most one workgroup per CU (the LDS use allows no more), so that every
    // workgroup of a launch is resident at once on the stream's device
    c->chol_resident = std::min(per_cu, 1) * ncu;
    if (c->chol_resident < 1) return PSBA_ERR_UNSUPPORTED;
    // one workgroup per task of the first phase, and one more for the forward
    // substitution when there are at least three tiles
    c->chol_grid = std::max(1, std::min(c->chol_resident, (T - 1) * T / 2 + (T >= 3 ? 1 : 0)));
  }
  int G = c->chol_grid;
  double* ok = c->d_scal + SC_FACTOR_OK;
  int ld = c->n_pad, n = c->n;
  probe_begin(c, PSBA_PROBE_FACTOR);
  hipLaunchKernelGGL(k_chol_solve, dim3(G), dim3(256), 0, c->stream, c->S, ld, T, c->Lf, c->ea, c->bw, c->yv,
                     c->xsol, c->dp, n, ok, c->bar, c->chol_fail, gate_of(c, kGateActive));
  probe_end(c, PSBA_PROBE_FACTOR);
  c->L_cur = c->Lf;
  PSBA_LAUNCH_TRY();
  return PSBA_OK;
}

}  // namespace psba

// ---------------------------------------------------------------------------
// batched problems: one workgroup factors and solves one problem's reduced
// system with the same tile tasks as k_chol_solve, run in sequence (the
// diagonal task of each phase last, so its factor stays in LDS for the next
// phase).  Arithmetic per tile is the same as in the single-problem kernel.
// ---------------------------------------------------------------------------
#include "batch.hpp"

namespace psba {

__global__ __launch_bounds__(256) void k_b_chol_solve(double* __restrict__ S, double* __restrict__ Lf,
                                                      const long* __restrict__ S_off, const int* __restrict__ npad,
                                                      const int* __restrict__ vec_off, const int* __restrict__ cam_off,
                                                      int cnp, const double* __restrict__ ea,
                                                      double* __restrict__ bw, double* __restrict__ yv,
                                                      double* __restrict__ xs, double* __restrict__ dp,
                                                      double* __restrict__ fok, const unsigned char* __restrict__ run) {
  __shared__ CholSmem sm;
  const int p = blockIdx.x;
  if (!run[p]) return;
  const int tid = threadIdx.x;
  const int ld = npad[p], T = ld / NB, n = (cam_off[p + 1] - cam_off[p]) * cnp;
  double* A = S + S_off[p];
  double* L = Lf + S_off[p];
  const double* b = ea + (size_t)cam_off[p] * cnp;
  double* bwp = bw + vec_off[p];
  double* yvp = yv + vec_off[p];
  double* xsp = xs + vec_off[p];
  double* dpp = dp + (size_t)cam_off[p] * cnp;
  auto tile = [&](double* base, int i, int j) { return base + (size_t)i * NB * ld + (size_t)j * NB; };
  int bad = 0;
  for (int q = tid; q < T * NB; q += blockDim.x) bwp[q] = (q < n) ? b[q] : 0.0;
  load_full(sm.Y, A, ld);  // (all sixteen loads of a thread in flight together)
  if (tid < NB) sm.part[0][tid] = (tid < n) ? b[tid] : 0.0;
  __syncthreads();
  bad |= tile_potrf(*(lds_smem*)&sm, LDS(sm.Y), false, 4);
  store_tile(sm.Y, L, ld);
  tile_forward(LDS(sm.Y), LDS(sm.part[0]), yvp, LDS(sm.w), NB);
  __syncthreads();

  for (int k = 0; k + 1 < T; ++k) {
    for (int f = tid; f < NB * NB; f += blockDim.x) {
      const int r = f >> 6, c = f & 63;
      sm.P[r * TP + c] = sm.Y[r * TP + c];
    }
    if (tid < NB) sm.v[tid] = sm.w[tid];
    __syncthreads();
    if (tid < NB) sm.rd[tid] = 1.0 / sm.P[tid * TP + tid];
    __syncthreads();
    const int m = T - 1 - k, ntasks = m * (m + 1) / 2;
    // tasks in column order, the diagonal (k+1, k+1) moved to the end
    for (int tt = 1; tt <= ntasks; ++tt) {
      const int t = (tt == ntasks) ? 0 : tt;
      int a = 0, cum = 0;
      while (cum + (m - a) <= t) { cum += m - a; ++a; }
      const int j = k + 1 + a, i = j + (t - cum);
      dvec4 acc[2][2];
      task_load(acc, sm, tile(A, i, j), tile(A, i, k), (i != j) ? tile(A, j, k) : nullptr, ld);
      trsm_rows(LDS(sm.X), LDS(sm.P), LDS(sm.rd));
      if (i != j) trsm_rows(LDS(sm.Y), LDS(sm.P), LDS(sm.rd));
      __syncthreads();
      acc_update(acc, sm.X, (i != j) ? sm.Y : sm.X, i == j);
      if (j == k + 1) {
        store_tile(sm.X, tile(L, i, k), ld);
        const int r = tid >> 2, pp = tid & 3;
        double s = 0.0;
#pragma unroll
        for (int c = 0; c < 16; ++c) s += sm.X[r * TP + 16 * pp + c] * sm.v[16 * pp + c];
        s += __shfl_xor(s, 1, 64);
        s += __shfl_xor(s, 2, 64);
        if (pp == 0) {
          const double bi = bwp[(size_t)i * NB + r] - s;
          bwp[(size_t)i * NB + r] = bi;
          if (i == j) sm.part[0][r] = bi;
        }
      }
      if (i == j && j == k + 1) {
        __syncthreads();
        acc_store(acc, sm.Y, TP);
        __syncthreads();
        bad |= tile_potrf(*(lds_smem*)&sm, LDS(sm.Y), false, 4);
        store_tile(sm.Y, tile(L, i, i), ld);
        tile_forward(LDS(sm.Y), LDS(sm.part[0]), yvp + (size_t)i * NB, LDS(sm.w), NB);
      } else {
        acc_store(acc, tile(A, i, j), ld);
      }
      __syncthreads();
    }
  }

  // back substitution, right-looking as in k_chol_solve: step k applies
  // x_{k+1} (in sm.w) to r_k, solves x_k, then applies x_{k+1} to r_i, i < k
  for (int k = T - 1; k >= 0; --k) {
    load_lower(sm.P, tile(L, k, k), ld);
    double br = yvp[(size_t)k * NB + (tid & 63)];
    if (k + 1 < T) br -= tile_tmatvec(sm, tile(L, k + 1, k), ld, sm.w);
    __syncthreads();
    tile_backward(LDS(sm.P), br, xsp + (size_t)k * NB, LDS(sm.v), NB);
    __syncthreads();
    if (tid < NB) {
      const int q = k * NB + tid;
      if (q < n) dpp[q] = sm.v[tid];
    }
    if (k + 1 < T)
      for (int i = 0; i < k; ++i) {
        const double t = tile_tmatvec(sm, tile(L, k + 1, i), ld, sm.w);
        if (tid < NB) yvp[(size_t)i * NB + tid] -= t;
      }
    __syncthreads();
    if (tid < NB) sm.w[tid] = sm.v[tid];
    __syncthreads();
  }
  if (tid == 0) fok[p] = bad ? 0.0 : 1.0;
}

// Batches whose largest problem has several tiles run the same tile tasks with
// one launch per phase instead, every task of a phase (over all problems) on
// its own workgroup: the kernel boundary is the barrier between phases, so
// workgroups never wait for each other.  Task tables: phase k of problem p
// owns workgroups fo[k][p] .. fo[k][p+1]-1 (forward), back-substitution step s
// owns bo[s][p] .. bo[s][p+1]-1.

// p with off[p] <= t < off[p+1]
__device__ __forceinline__ int find_owner(const int* __restrict__ off, int P, int t) {
  int lo = 0, hi = P - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= t)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// phase -1: right-hand side, factor of tile (0,0), y_0; one-tile problems
// also solve x_0 here
__global__ __launch_bounds__(256) void k_b_chol_init(double* __restrict__ S, double* __restrict__ Lf,
                                                     const long* __restrict__ S_off, const int* __restrict__ npad,
                                                     const int* __restrict__ vec_off,
                                                     const int* __restrict__ cam_off, int cnp,
                                                     const double* __restrict__ ea, double* __restrict__ bw,
                                                     double* __restrict__ yv, double* __restrict__ xs,
                                                     double* __restrict__ dp, double* __restrict__ fok,
                                                     const unsigned char* __restrict__ run) {
  __shared__ CholSmem sm;
  const int p = blockIdx.x;
  if (!run[p]) return;
  const int tid = threadIdx.x;
  const int ld = npad[p], T = ld / NB, n = (cam_off[p + 1] - cam_off[p]) * cnp;
  const double* A = S + S_off[p];
  double* L = Lf + S_off[p];
  const double* b = ea + (size_t)cam_off[p] * cnp;
  double* bwp = bw + vec_off[p];
  for (int q = tid; q < T * NB; q += blockDim.x) bwp[q] = (q < n) ? b[q] : 0.0;
  load_full(sm.Y, A, ld);  // (all sixteen loads of a thread in flight together)
  if (tid < NB) sm.part[0][tid] = (tid < n) ? b[tid] : 0.0;
  __syncthreads();
  const int bad = tile_potrf(*(lds_smem*)&sm, LDS(sm.Y), false, 4);
  store_tile(sm.Y, L, ld);
  tile_forward(LDS(sm.Y), LDS(sm.part[0]), yv + vec_off[p], LDS(sm.w), NB);
  __syncthreads();
  if (T == 1) {
    const double br = sm.w[tid & 63];
    __syncthreads();
    tile_backward(LDS(sm.Y), br, xs + vec_off[p], LDS(sm.v), NB);
    __syncthreads();
    if (tid < n) dp[(size_t)cam_off[p] * cnp + tid] = sm.v[tid];
  }
  if (tid == 0) fok[p] = bad ? 0.0 : 1.0;
}

// forward phase k: one tile task (i, j), j >= k + 1, per workgroup
__global__ __launch_bounds__(256) void k_b_chol_phase(int k, const int* __restrict__ fo, int P,
                                                      double* __restrict__ S, double* __restrict__ Lf,
                                                      const long* __restrict__ S_off, const int* __restrict__ npad,
                                                      const int* __restrict__ vec_off, double* __restrict__ bw,
                                                      double* __restrict__ yv, double* __restrict__ fok,
                                                      const unsigned char* __restrict__ run) {
  __shared__ CholSmem sm;
  const int p = find_owner(fo, P, blockIdx.x);
  if (!run[p]) return;
  const int t = blockIdx.x - fo[p];
  const int tid = threadIdx.x;
  const int ld = npad[p], T = ld / NB, m = T - 1 - k;
  double* A = S + S_off[p];
  double* L = Lf + S_off[p];
  double* bwp = bw + vec_off[p];
  double* yvp = yv + vec_off[p];
  auto tile = [&](double* base, int i, int j) { return base + (size_t)i * NB * ld + (size_t)j * NB; };
  load_lower(sm.P, tile(L, k, k), ld);
  if (tid < NB) sm.v[tid] = yvp[(size_t)k * NB + tid];
  __syncthreads();
  if (tid < NB) sm.rd[tid] = 1.0 / sm.P[tid * TP + tid];
  __syncthreads();
  int a = 0, cum = 0;
  while (cum + (m - a) <= t) { cum += m - a; ++a; }
  const int j = k + 1 + a, i = j + (t - cum);
  dvec4 acc[2][2];
  task_load(acc, sm, tile(A, i, j), tile(A, i, k), (i != j) ? tile(A, j, k) : nullptr, ld);
  trsm_rows(LDS(sm.X), LDS(sm.P), LDS(sm.rd));
  if (i != j) trsm_rows(LDS(sm.Y), LDS(sm.P), LDS(sm.rd));
  __syncthreads();
  acc_update(acc, sm.X, (i != j) ? sm.Y : sm.X, i == j);
  if (j == k + 1) {
    store_tile(sm.X, tile(L, i, k), ld);
    const int r = tid >> 2, pp = tid & 3;
    double s = 0.0;
#pragma unroll
    for (int c = 0; c < 16; ++c) s += sm.X[r * TP + 16 * pp + c] * sm.v[16 * pp + c];
    s += __shfl_xor(s, 1, 64);
    s += __shfl_xor(s, 2, 64);
    if (pp == 0) {
      const double bi = bwp[(size_t)i * NB + r] - s;
      bwp[(size_t)i * NB + r] = bi;
      if (i == j) sm.part[0][r] = bi;
    }
  }
  if (i == j && j == k + 1) {
    __syncthreads();
    acc_store(acc, sm.Y, TP);
    __syncthreads();
    const int bad = tile_potrf(*(lds_smem*)&sm, LDS(sm.Y), false, 4);
    store_tile(sm.Y, tile(L, i, i), ld);
    tile_forward(LDS(sm.Y), LDS(sm.part[0]), yvp + (size_t)i * NB, LDS(sm.w), NB);
    if (bad && tid == 0) fok[p] = 0.0;
  } else {
    acc_store(acc, tile(A, i, j), ld);
  }
}

// back-substitution step s (block row k = T - 1 - s of each problem with
// T >= 2): workgroup 0 of a problem applies x_{k+1} to r_k and solves x_k,
// workgroup 1 + i applies x_{k+1} to r_i, i < k
__global__ __launch_bounds__(256) void k_b_chol_back(int s, const int* __restrict__ bo, int P,
                                                     const double* __restrict__ Lf, const long* __restrict__ S_off,
                                                     const int* __restrict__ npad, const int* __restrict__ vec_off,
                                                     const int* __restrict__ cam_off, int cnp,
                                                     double* __restrict__ yv, double* __restrict__ xs,
                                                     double* __restrict__ dp, const unsigned char* __restrict__ run) {
  __shared__ CholSmem sm;
  const int p = find_owner(bo, P, blockIdx.x);
  if (!run[p]) return;
  const int idx = blockIdx.x - bo[p];
  const int tid = threadIdx.x;
  const int ld = npad[p], T = ld / NB, k = T - 1 - s, n = (cam_off[p + 1] - cam_off[p]) * cnp;
  const double* L = Lf + S_off[p];
  double* yvp = yv + vec_off[p];
  double* xsp = xs + vec_off[p];
  auto tile = [&](const double* base, int i, int j) { return base + (size_t)i * NB * ld + (size_t)j * NB; };
  if (s > 0 && tid < NB) sm.w[tid] = xsp[(size_t)(k + 1) * NB + tid];
  __syncthreads();
  if (idx == 0) {
    load_lower(sm.P, tile(L, k, k), ld);
    double br = yvp[(size_t)k * NB + (tid & 63)];
    if (s > 0) br -= tile_tmatvec(sm, tile(L, k + 1, k), ld, sm.w);
    __syncthreads();
    tile_backward(LDS(sm.P), br, xsp + (size_t)k * NB, LDS(sm.v), NB);
    __syncthreads();
    if (tid < NB) {
      const int q = k * NB + tid;
      if (q < n) dp[(size_t)cam_off[p] * cnp + q] = sm.v[tid];
    }
  } else {
    const int i = idx - 1;
    const double t = tile_tmatvec(sm, tile(L, k + 1, i), ld, sm.w);
    if (tid < NB) yvp[(size_t)i * NB + tid] -= t;
  }
}

int launch_b_factor_solve(psba_batch* b) {
  if (!b->chol_phased) {
    hipLaunchKernelGGL(k_b_chol_solve, dim3(b->P), dim3(256), 0, b->stream, b->S, b->Lf, b->S_off, b->npad,
                       b->vec_off, b->cam_off, b->cnp, b->ea, b->bw, b->yv, b->xs, b->dp, b->fok, b->active);
    PSBA_LAUNCH_TRY();
    return PSBA_OK;
  }
  const int P = b->P;
  hipLaunchKernelGGL(k_b_chol_init, dim3(P), dim3(256), 0, b->stream, b->S, b->Lf, b->S_off, b->npad, b->vec_off,
                     b->cam_off, b->cnp, b->ea, b->bw, b->yv, b->xs, b->dp, b->fok, b->active);
  for (int k = 0; k + 1 < b->T_max; ++k)
    if (b->h_Gf[k] > 0)
      hipLaunchKernelGGL(k_b_chol_phase, dim3(b->h_Gf[k]), dim3(256), 0, b->stream, k,
                         b->chol_fo + (size_t)k * (P + 1), P, b->S, b->Lf, b->S_off, b->npad, b->vec_off, b->bw,
                         b->yv, b->fok, b->active);
  for (int s = 0; s < b->T_max; ++s)
    if (b->h_Gb[s] > 0)
      hipLaunchKernelGGL(k_b_chol_back, dim3(b->h_Gb[s]), dim3(256), 0, b->stream, s,
                         b->chol_bo + (size_t)s * (P + 1), P, b->Lf, b->S_off, b->npad, b->vec_off, b->cam_off,
                         b->cnp, b->yv, b->xs, b->dp, b->active);
  PSBA_LAUNCH_TRY();
  return PSBA_OK;
}

}  // namespace psba
