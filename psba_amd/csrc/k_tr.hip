// Vector kernels of the dogleg / 2-D subspace step (PSBA/trust_region.cpp:
// 125-130 Cauchy direction, :520-595 compute_p_2).  Reductions are the same
// deterministic two-level sums as the rest of the path.
#include "context.hpp"

namespace psba {

// P_U = -(g * (g.g)) / (g^T B g), trust_region.cpp:128-130
__global__ __launch_bounds__(256) void k_cauchy(double* __restrict__ out, const double* __restrict__ g, double gtg,
                                                double gtBg, long n) {
  long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) out[q] = -(g[q] * gtg) / gtBg;
}

// mode 0: (P_U.g, P_B.g, P_U.P_U, P_U.P_B, P_B.P_B)            trust_region.cpp:530-534
// mode 1: sum (B-U)^2, sum (B-U)(2U-B), sum (2U-B)^2, -, -      trust_region.cpp:577-582
__global__ __launch_bounds__(256) void k_tr_dots(const double* __restrict__ PU, const double* __restrict__ PB,
                                                 const double* __restrict__ g, long n, int mode,
                                                 double* partials) {
  __shared__ double sh[20];
  double acc[5] = {0, 0, 0, 0, 0};
  long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) {
    double u = PU[q], b = PB[q];
    if (mode == 0) {
      double gg = g[q];
      acc[0] = u * gg;
      acc[1] = b * gg;
      acc[2] = u * u;
      acc[3] = u * b;
      acc[4] = b * b;
    } else {
      double A = b - u, B = 2 * u - b;
      acc[0] = A * A;
      acc[1] = A * B;
      acc[2] = B * B;
    }
  }
  grid_sum_partials<5>(acc, partials, sh);
}

// op 0: s1 a + s2 b        (2-D subspace minimiser)
// op 1: (s1 a) / s2        (scaled Cauchy step)
// op 2: a + (s1 - 1)(b - a) (dogleg point)
__global__ __launch_bounds__(256) void k_comb(double* __restrict__ out, const double* __restrict__ a,
                                              const double* __restrict__ b, double s1, double s2, int op, long n) {
  long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  double v;
  if (op == 0) v = s1 * a[q] + s2 * b[q];
  else if (op == 1) v = s1 * a[q] / s2;
  else v = a[q] + (s1 - 1) * (b[q] - a[q]);
  out[q] = v;
}

int launch_cauchy(psba_ctx* c, double* out, const double* g, double gtg, double gtBg) {
  hipLaunchKernelGGL(k_cauchy, dim3(ceil_div(c->ntot, 256)), dim3(256), 0, c->stream, out, g, gtg, gtBg,
                     (long)c->ntot);
  PSBA_LAUNCH_TRY();
  return PSBA_OK;
}

int launch_tr_dots(psba_ctx* c, const double* PU, const double* PB, const double* g, int mode) {
  const unsigned nb = ceil_div(c->ntot, 256);
  hipLaunchKernelGGL(k_tr_dots, dim3(nb), dim3(256), 0, c->stream, PU, PB, g, (long)c->ntot, mode, c->red_part);
  launch_sum_partials<5>(c->stream, c->red_part, (int)nb, c->d_scal + SC_D0);
  PSBA_LAUNCH_TRY();
  return PSBA_OK;
}

int launch_comb(psba_ctx* c, double* out, const double* a, const double* b, double s1, double s2, int op) {
  hipLaunchKernelGGL(k_comb, dim3(ceil_div(c->ntot, 256)), dim3(256), 0, c->stream, out, a, b, s1, s2, op,
                     (long)c->ntot);
  PSBA_LAUNCH_TRY();
  return PSBA_OK;
}

}  // namespace psba
