// Shared helpers for the HIP kernels: error codes, launch checks and
// deterministic (fixed-tree) block reductions for 64-wide wavefronts.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

#include "psba_capi.h"

namespace psba {

constexpr int kWave = 64;

#define PSBA_HIP_TRY(expr)                                                            \
  do {                                                                                \
    hipError_t _e = (expr);                                                           \
    if (_e != hipSuccess) {                                                           \
      psba::set_last_error(__FILE__, __LINE__, hipGetErrorString(_e));                \
      return PSBA_ERR_HIP;                                                            \
    }                                                                                 \
  } while (0)

#define PSBA_LAUNCH_TRY()                                                             \
  do {                                                                                \
    hipError_t _e = hipGetLastError();                                                \
    if (_e != hipSuccess) {                                                           \
      psba::set_last_error(__FILE__, __LINE__, hipGetErrorString(_e));                \
      return PSBA_ERR_HIP;                                                            \
    }                                                                                 \
  } while (0)

#define PSBA_TRY(expr)        \
  do {                        \
    int _r = (expr);          \
    if (_r != PSBA_OK) return _r; \
  } while (0)

void set_last_error(const char* file, int line, const char* msg);

// Sum over the 64 lanes of a wavefront with a fixed butterfly order; every lane
// receives the result.
__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

__device__ inline float wave_sum(float v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Deterministic block reduction of K values per thread.  `sh` must hold
// K * (blockDim.x / 64) doubles.  Result valid in thread 0.
template <int K>
__device__ inline void block_sum(double (&v)[K], double* sh) {
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  const int nw = blockDim.x / kWave;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    double s = wave_sum(v[k]);
    if (lane == 0) sh[k * nw + wid] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      double s = 0;
      for (int w = 0; w < nw; ++w) s += sh[k * nw + w];
      v[k] = s;
    }
  }
  __syncthreads();
}

// Two-level deterministic reduction: each block writes its K partials
// (block order); k_sum_partials, one block queued behind the kernel on the same
// stream, adds them in block order and writes out[0..K).  No inter-block
// atomics or device-scope fences: on a multi-XCD device each such fence forces
// an L2 write-back, which made the per-block ticket scheme cost tens of
// microseconds per launch.
template <int K>
__device__ inline void grid_sum_partials(double (&v)[K], double* partials, double* sh, long blk) {
  block_sum<K>(v, sh);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) partials[(size_t)blk * K + k] = v[k];
  }
}
template <int K>
__device__ inline void grid_sum_partials(double (&v)[K], double* partials, double* sh) {
  grid_sum_partials<K>(v, partials, sh, blockIdx.x);
}

// One block adds nb blocks' K partials in block order (thread t takes blocks
// t, t + blockDim, .., then block_sum) and writes out[0..K) from thread 0.
// Eight of a thread's blocks are loaded at a time (clamped to the last block,
// the additions past the end skipped), so a batch waits for one load latency;
// the additions keep the order b = t, t + blockDim, ...  With res, thread 0
// also gets the totals there.
template <int K>
__device__ inline void sum_partials_block(const double* __restrict__ partials, int nb, double* __restrict__ out,
                                          double* sh, double* res = nullptr) {
  constexpr int B = 8;
  double acc[K];
#pragma unroll
  for (int k = 0; k < K; ++k) acc[k] = 0;
  const int step = blockDim.x;
  for (int b0 = threadIdx.x; b0 < nb; b0 += B * step) {
    double v[B][K];
#pragma unroll
    for (int u = 0; u < B; ++u) {
      const int b = min(b0 + u * step, nb - 1);
#pragma unroll
      for (int k = 0; k < K; ++k) v[u][k] = partials[(size_t)b * K + k];
    }
#pragma unroll
    for (int u = 0; u < B; ++u) {
      const bool live = b0 + u * step < nb;
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const double t = acc[k] + v[u][k];
        acc[k] = live ? t : acc[k];
      }
    }
  }
  block_sum<K>(acc, sh);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      out[k] = acc[k];
      if (res) res[k] = acc[k];
    }
  }
}

template <int K>
__global__ __launch_bounds__(256) void k_sum_partials(const double* __restrict__ partials, int nb,
                                                      double* __restrict__ out) {
  __shared__ double sh[K * 4];
  sum_partials_block<K>(partials, nb, out, sh);
}

template <int K>
inline void launch_sum_partials(hipStream_t st, const double* partials, int nb, double* out) {
  hipLaunchKernelGGL(k_sum_partials<K>, dim3(1), dim3(256), 0, st, partials, nb, out);
}

inline unsigned ceil_div(long a, long b) { return (unsigned)((a + b - 1) / b); }

// Sum over the four lanes of a quad (DPP quad permutes), the same value on
// all four: (v0 + v1) + (v2 + v3) in a fixed order.  Every lane of the wave
// must take part.
__device__ __forceinline__ double quad_partner(double v, int ctrl_is_xor2) {
  const long long bits = __double_as_longlong(v);
  int lo, hi;
  if (ctrl_is_xor2) {
    lo = __builtin_amdgcn_mov_dpp((int)bits, 0x4E, 0xF, 0xF, true);
    hi = __builtin_amdgcn_mov_dpp((int)(bits >> 32), 0x4E, 0xF, 0xF, true);
  } else {
    lo = __builtin_amdgcn_mov_dpp((int)bits, 0xB1, 0xF, 0xF, true);
    hi = __builtin_amdgcn_mov_dpp((int)(bits >> 32), 0xB1, 0xF, 0xF, true);
  }
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ float quad_partner(float v, int ctrl_is_xor2) {
  return __int_as_float(ctrl_is_xor2 ? __builtin_amdgcn_mov_dpp(__float_as_int(v), 0x4E, 0xF, 0xF, true)
                                     : __builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xF, 0xF, true));
}
template <typename R>
__device__ __forceinline__ R quad_sum(R v) {
  const R a = v + quad_partner(v, 0);  // lanes (0,1) and (2,3)
  return a + quad_partner(a, 1);       // the two pair sums
}

// Launch order of the Schur pair chunks: the chunks of camera pairs in the
// same block of 24 x 24 cameras and the same range of 2048 points run
// together, so the W rows they gather come back from L2 instead of HBM (with
// the XCD mapping each XCD walks its own share of this order).  Sorting is
// stable, so ties keep the chunk order.  (Measured on the 88-camera scene:
// 8 x 8 cameras over 4096 points 254 us, 16 or 32 cameras over 1024-4096
// points 249.5-251 us, 24 over 2048 248 us.)
constexpr int kSchurTileCams = 24, kSchurPointRangeLog2 = 11;
__host__ __device__ inline unsigned long long schunk_launch_key(int k, int l, int pt) {
  const unsigned long long K = (unsigned)k / kSchurTileCams, L = (unsigned)l / kSchurTileCams;
  return ((K * (K + 1) / 2 + L) << 32) | ((unsigned)pt >> kSchurPointRangeLog2);
}

// Robust loss of an observation's squared residual norm s with scale d
// (PSBA_LOSS_*): rho(s) and rho'(s).
__device__ __forceinline__ void loss_eval(int kind, double d, double s, double& rho, double& drho) {
  const double d2 = d * d;
  if (kind == PSBA_LOSS_HUBER) {
    if (s <= d2) {
      rho = s;
      drho = 1.0;
    } else {
      const double r = sqrt(s);
      rho = 2.0 * d * r - d2;
      drho = d / r;
    }
  } else if (kind == PSBA_LOSS_CAUCHY) {
    rho = d2 * log1p(s / d2);
    drho = 1.0 / (1.0 + s / d2);
  } else {
    rho = s;
    drho = 1.0;
  }
}

// acc = 0 + load(q0) + load(q0 + 1) + .. + load(q1 - 1), added in that order,
// with the loads issued B at a time and unconditionally (indices past the end
// clamped to the last one, their additions skipped), so that each batch of B
// costs one wait rather than one per unrolled group of a runtime-bounded loop.
template <int B, typename F>
__device__ inline double ordered_sum(int q0, int q1, F load) {
  double acc = 0.0;
  for (int qb = q0; qb < q1; qb += B) {
    double v[B];
#pragma unroll
    for (int u = 0; u < B; ++u) v[u] = load(min(qb + u, q1 - 1));
#pragma unroll
    for (int u = 0; u < B; ++u) {
      const double t = acc + v[u];
      acc = (qb + u < q1) ? t : acc;
    }
  }
  return acc;
}

// A problem's gate (batch: per problem through ownr; one problem: ownr may be
// null, the gate is run[0]).
__device__ inline bool gated_off(const unsigned char* run, const int* ownr, int i) {
  return run && !run[ownr ? ownr[i] : 0];
}

// Workgroups of a launch go to the 8 XCDs round-robin (workgroup b to XCD
// b % 8), each XCD with its own L2.  A grid of xcd_grid(n) workgroups with
// item = xcd_item(b, n) gives every XCD one contiguous eighth of an n-item
// work list (neighbouring items share data); -1 past the end.
constexpr int kXcds = 8;
inline unsigned xcd_grid(long n) { return (unsigned)(kXcds * ((n + kXcds - 1) / kXcds)); }
__device__ inline int xcd_item(int b, int n) {
  const int per = (n + kXcds - 1) / kXcds;
  const int it = (b % kXcds) * per + b / kXcds;
  return it < n ? it : -1;
}

// call FN<cnp>(args...) for the supported camera models, else return
// PSBA_ERR_UNSUPPORTED from the enclosing function
#define PSBA_DISPATCH_CNP(cnp, FN, ...)   \
  switch (cnp) {                          \
    case 6: FN<6>(__VA_ARGS__); break;    \
    case 9: FN<9>(__VA_ARGS__); break;    \
    case 11: FN<11>(__VA_ARGS__); break;  \
    case 16: FN<16>(__VA_ARGS__); break;  \
    default: return PSBA_ERR_UNSUPPORTED; \
  }

}  // namespace psba
