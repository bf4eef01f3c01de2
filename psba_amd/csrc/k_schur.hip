// Reduced camera system of the plain FP64 path:
//   S_kl = delta_kl (U_k + mu I) - sum_{i in P(k,l)} W_ik V_i^-1 W_il^T   (k >= l)
//   e_a,j = g_a,j - sum_i W_ij (V_i^-1 g_b,i)
// Reference: PSBA/CL_files/compute_Yblks.cl, compute_S.cl, compute_ea.cl.
// Y = W V^-1 is formed on the fly in LDS (never written to HBM); e_a uses
// z_i = V_i^-1 g_b,i from the inverse kernel.  Both sums are deterministic
// two-stage reductions over fixed chunks (item order = ascending point).
// R is the storage and chunk-accumulation type (double on the plain path,
// float on the mixed path); V^-1 is always FP64 and the chunk partials are
// summed in FP64.
#include "context.hpp"

#ifndef PSBA_SCHUR_XCD
#define PSBA_SCHUR_XCD 1
#endif

namespace psba {

// One workgroup per chunk (one camera pair, at most kPairChunk shared points).
//   1. the chunk's item indices (W_a, W_b observation rows and the point) go to
//      LDS in one round of loads;
//   2. four threads per item load W_a, W_b and V^-1 with independent
//      addresses (no dependent global round trips) and stage Y = W_a V^-1 and
//      W_b in LDS;
//   3. thread (g, blk) accumulates a 3x3 block of the CNP x CNP product over
//      items g, g + G, ... in registers (18 LDS values per 27 multiply-adds);
//   4. the G group partials are summed in group order into part[b].
// The order of every sum is fixed by the chunk, so results do not depend on
// scheduling.
// With SLOTS (one problem): the workgroup of launch slot s reads its chunk's
// item records at rec[s * kPairChunk ..] (observations, point, chunk; -1 past
// the chunk's items), an address its index gives, so the W rows and V^-1 are
// its second round of loads.
// The workgroup's work as a routine of its launch index `wg` and its LDS
// (sBuf: 2 kPairChunk 3 CNP values, sIdx: 3 kPairChunk, sMeta: 2), so that
// k_schur_ea_chunks can run it beside the e_a chunks.
template <int CNP, typename R, bool SLOTS>
__device__ __forceinline__ void schur_chunk(int wg, const int2* __restrict__ items, const int* __restrict__ sb,
                                            const int* __restrict__ se, const int* __restrict__ obs_pt,
                                            const R* __restrict__ W, const double* __restrict__ Vinv,
                                            R* __restrict__ part, const int* __restrict__ ownr,
                                            const unsigned char* __restrict__ run, int nchunk,
                                            const int* __restrict__ order, const int4* __restrict__ rec, R* sBuf,
                                            int* sIdx, int* sMeta) {
  constexpr int C2 = CNP * CNP, REC = 3 * CNP;
  constexpr int NBK = (CNP + 2) / 3;  // 3-wide row / column blocks
  constexpr int NB2 = NBK * NBK;
  constexpr int G = 256 / NB2;        // item groups
  static_assert(G * C2 <= 2 * kPairChunk * REC, "group partials must fit the staging buffer");
  R* const sY = sBuf;
  R* const sW = sBuf + kPairChunk * REC;
#if PSBA_SCHUR_XCD
  int b = xcd_item(wg, nchunk);  // one share of the launch order per XCD
  if (b < 0) return;
#else
  int b = wg;
#endif
  const int t = threadIdx.x;
  int cnt;
  if constexpr (SLOTS) {
    if (t < kPairChunk) {
      const int4 r = rec[(size_t)b * kPairChunk + t];
      sIdx[t] = r.x;
      sIdx[kPairChunk + t] = r.y;
      sIdx[2 * kPairChunk + t] = r.z;
      // the chunk's items are the records before the first -1
      const unsigned long long m = __ballot(r.x >= 0);
      if (t == 0) {
        sMeta[0] = r.w;
        sMeta[1] = __popcll(m);
      }
    }
    __syncthreads();
    b = sMeta[0];
    cnt = sMeta[1];
  } else {
    if (order) b = order[b];
    if (run && !run[ownr[b]]) return;
    const int beg = sb[b];
    cnt = se[b] - beg;
    if (t < cnt) {
      const int2 pr = items[beg + t];
      sIdx[t] = pr.x;
      sIdx[kPairChunk + t] = pr.y;
      sIdx[2 * kPairChunk + t] = obs_pt[pr.x];
    }
    __syncthreads();
  }
  {
    const int it = t >> 2, q = t & 3;
    if (it < cnt) {
      const R* wa = W + (size_t)sIdx[it] * REC;
      const R* wb = W + (size_t)sIdx[kPairChunk + it] * REC;
      const double* vi = Vinv + (size_t)sIdx[2 * kPairChunk + it] * 6;
      // every load first and unconditional (a lane's rows and elements past
      // the record are clamped to the last one and not stored), so that one
      // wait covers them all: per-lane loop bounds made each remainder a
      // round of its own
      constexpr int NR = (CNP + 3) / 4, NW = (REC + 3) / 4;
      double v[6];
      R wv[NR][3], bv[NW];
#pragma unroll
      for (int k = 0; k < 6; ++k) v[k] = vi[k];
#pragma unroll
      for (int j = 0; j < NR; ++j) {
        const int r = min(q + 4 * j, CNP - 1);
#pragma unroll
        for (int c = 0; c < 3; ++c) wv[j][c] = wa[3 * r + c];
      }
#pragma unroll
      for (int j = 0; j < NW; ++j) bv[j] = wb[min(q + 4 * j, REC - 1)];
      // (the empty asm keeps the compiler from sinking the loads of a row
      // used under a condition into that condition's block)
#pragma unroll
      for (int j = 0; j < NR; ++j)
#pragma unroll
        for (int c = 0; c < 3; ++c) __asm__ volatile("" : "+v"(wv[j][c]));
      const R v00 = (R)v[0], v01 = (R)v[1], v02 = (R)v[2], v11 = (R)v[3], v12 = (R)v[4], v22 = (R)v[5];
#pragma unroll
      for (int j = 0; j < NR; ++j) {
        const int r = q + 4 * j;
        if (r < CNP) {
          const R w0 = wv[j][0], w1 = wv[j][1], w2 = wv[j][2];
          R* y = sY + it * REC + 3 * r;
          y[0] = w0 * v00 + w1 * v01 + w2 * v02;
          y[1] = w0 * v01 + w1 * v11 + w2 * v12;
          y[2] = w0 * v02 + w1 * v12 + w2 * v22;
        }
      }
#pragma unroll
      for (int j = 0; j < NW; ++j)
        if (q + 4 * j < REC) sW[it * REC + q + 4 * j] = bv[j];
    }
  }
  __syncthreads();
  const int grp = t / NB2, blk = t - grp * NB2;
  const int rb = blk / NBK, cb = blk - rb * NBK;
  R acc[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) acc[i][j] = 0;
  if (grp < G) {
    for (int it = grp; it < cnt; it += G) {
      const R* y = sY + it * REC;
      const R* w = sW + it * REC;
      R yv[3][3], wv[3][3];
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        const int r = 3 * rb + i, cc = 3 * cb + i;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          yv[i][c] = r < CNP ? y[3 * r + c] : (R)0;
          wv[i][c] = cc < CNP ? w[3 * cc + c] : (R)0;
        }
      }
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) acc[i][j] += yv[i][0] * wv[j][0] + yv[i][1] * wv[j][1] + yv[i][2] * wv[j][2];
    }
  }
  __syncthreads();  // the staging buffer now takes the group partials
  R* const sG = sBuf;
  if (grp < G) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const int r = 3 * rb + i, cc = 3 * cb + j;
        if (r < CNP && cc < CNP) sG[grp * C2 + r * CNP + cc] = acc[i][j];
      }
  }
  __syncthreads();
  for (int e = t; e < C2; e += blockDim.x) {
    R s = 0;
    for (int g = 0; g < G; ++g) s += sG[g * C2 + e];
    part[(size_t)b * C2 + e] = s;
  }
}

template <int CNP, typename R, bool SLOTS>
__global__ __launch_bounds__(256) void k_schur_chunks(const int2* __restrict__ items, const int* __restrict__ sb,
                                                      const int* __restrict__ se, const int* __restrict__ obs_pt,
                                                      const R* __restrict__ W, const double* __restrict__ Vinv,
                                                      R* __restrict__ part, const int* __restrict__ ownr,
                                                      const unsigned char* __restrict__ run, int nchunk,
                                                      const int* __restrict__ order, const int4* __restrict__ rec) {
  __shared__ R sBuf[2 * kPairChunk * 3 * CNP];
  __shared__ int sIdx[3 * kPairChunk];
  __shared__ int sMeta[2];
  schur_chunk<CNP, R, SLOTS>(blockIdx.x, items, sb, se, obs_pt, W, Vinv, part, ownr, run, nchunk, order, rec, sBuf,
                             sIdx, sMeta);
}

// one block per lower block position (k >= l) of S, including empty ones
// Blocks [0, nlb) write the lower block positions; the remaining blocks
// rewrite the padding rows n..n_pad-1 as identity rows (a failed
// factorisation may have left anything there).
template <int CNP, typename R>
__device__ __forceinline__ void schur_finish_block(int b, const int* __restrict__ lblock_pair,
                                                   const int* __restrict__ psp, const R* __restrict__ part,
                                                   const double* __restrict__ U, const unsigned char* __restrict__ fix,
                                                   double mu, const double* __restrict__ mu_p, int ld, int nlb, int n,
                                                   double* __restrict__ S) {
  constexpr int C2 = CNP * CNP;
  if (mu_p) mu = *mu_p;
  if (b >= nlb) {
    const int r = n + (b - nlb);
    for (int q = threadIdx.x; q <= r; q += blockDim.x) S[(size_t)r * ld + q] = (q == r) ? 1.0 : 0.0;
    return;
  }
  int k = (int)((sqrt(8.0 * b + 1.0) - 1.0) * 0.5);
  while ((k + 1) * (k + 2) / 2 <= b) ++k;
  while (k * (k + 1) / 2 > b) --k;
  const int l = b - k * (k + 1) / 2;
  const int q = lblock_pair[b];
  for (int t = threadIdx.x; t < C2; t += blockDim.x) {
    const int r = t / CNP, cc = t - r * CNP;
    double acc = 0.0;
    if (q >= 0) acc = ordered_sum<16>(psp[q], psp[q + 1], [&](int h) { return (double)part[(size_t)h * C2 + t]; });
    double val;
    if (k == l) {
      double u = U[(size_t)k * C2 + t];
      if (r == cc) u = u + mu;
      val = u - acc;
      if (r == cc && fix[k * CNP + r]) val = 1.0;
    } else {
      val = -acc;
    }
    S[(size_t)(k * CNP + r) * ld + l * CNP + cc] = val;
  }
}

// e_a chunks over the camera chunks: p_o = W_o z_i, column sums
template <int CNP>
constexpr int ea_chunk_obs() { return (CNP > 11) ? 128 : kCamChunk; }

// the workgroup of chunk b, with its LDS (sP: ea_chunk_obs() CNP values,
// sO and sI: ea_chunk_obs() each) given
template <int CNP, typename R>
__device__ __forceinline__ void ea_chunk(int b, const int* __restrict__ cam_obs, const int* __restrict__ cb,
                                         const int* __restrict__ ce, const int* __restrict__ obs_pt,
                                         const R* __restrict__ W, const double* __restrict__ z, R* __restrict__ part,
                                         const int* __restrict__ ownr, const unsigned char* __restrict__ run,
                                         const int2* __restrict__ rec, R* sP, int* sO, int* sI) {
  constexpr int CH = ea_chunk_obs<CNP>();
  static_assert(CH <= 256, "one observation per thread");
  // with rec (one problem) the observation and its point come from the
  // chunk's record at an address the workgroup index gives, beside the extent
  const int2 r_rec = (rec && (int)threadIdx.x < CH) ? rec[(size_t)b * CH + threadIdx.x] : make_int2(-1, -1);
  const int beg = cb[b], cnt = ce[b] - beg;
  if (gated_off(run, ownr, b)) return;
  // the chunk's observation and point indices first (the dependent loads once
  // per observation), then the products with independent, coalesced loads
  for (int s = threadIdx.x; s < cnt; s += blockDim.x) {
    if (rec) {
      sO[s] = r_rec.x;
      sI[s] = r_rec.y;
    } else {
      const int o = cam_obs[beg + s];
      sO[s] = o;
      sI[s] = obs_pt[o];
    }
  }
  __syncthreads();
#pragma unroll 4
  for (int task = threadIdx.x; task < cnt * CNP; task += blockDim.x) {
    const int s = task / CNP, r = task - s * CNP;
    const R* w = W + (size_t)sO[s] * 3 * CNP + 3 * r;
    const double* zi = z + (size_t)sI[s] * 3;
    sP[task] = w[0] * (R)zi[0] + w[1] * (R)zi[1] + w[2] * (R)zi[2];
  }
  __syncthreads();
  // a quad per output, lane p over observations p, p + 4, .., added in a
  // fixed order
  if (threadIdx.x < 64) {
    const int t = threadIdx.x >> 2, p = threadIdx.x & 3;
    R acc = 0;
    if (t < CNP)
      for (int s = p; s < cnt; s += 4) acc += sP[s * CNP + t];
    acc = quad_sum(acc);
    if (t < CNP && p == 0) part[(size_t)b * CNP + t] = acc;
  }
}

template <int CNP, typename R>
__global__ __launch_bounds__(256) void k_ea_chunks(const int* __restrict__ cam_obs, const int* __restrict__ cb,
                                                   const int* __restrict__ ce, const int* __restrict__ obs_pt,
                                                   const R* __restrict__ W, const double* __restrict__ z,
                                                   R* __restrict__ part, const int* __restrict__ ownr,
                                                   const unsigned char* __restrict__ run,
                                                   const int2* __restrict__ rec) {
  constexpr int CH = ea_chunk_obs<CNP>();
  __shared__ R sP[CH * CNP];
  __shared__ int sO[CH], sI[CH];
  ea_chunk<CNP, R>(blockIdx.x, cam_obs, cb, ce, obs_pt, W, z, part, ownr, run, rec, sP, sO, sI);
}

// One problem's Schur chunks and e_a chunks in one launch (they share no
// data).  The launch is in groups of eight workgroups (one per XCD); of every
// period + 1 groups, the first `period` are Schur groups and the last an e_a
// group, until the e_a groups (neg) are placed, then Schur groups only.  A
// Schur group gs gives workgroup index 8 gs + blockIdx % 8, the index
// k_schur_chunks' workgroup on the same XCD has (SLOTS form, so the same XCD
// shares of the launch order); e_a group eg takes camera chunk
// 8 eg + blockIdx % 8.  The memory-streaming e_a chunks so run beside the
// latency-bound Schur chunks.  The e_a chunks take their LDS from the Schur
// staging buffer.  Each chunk's arithmetic is its own kernel's, so the
// partials are the same bits.
template <int CNP, typename R>
__global__ __launch_bounds__(256) void k_schur_ea_chunks(const R* __restrict__ W, const double* __restrict__ Vinv,
                                                         R* __restrict__ spart, int nchunk,
                                                         const int4* __restrict__ rec, int nb_schur, int neg,
                                                         int period, int ncc, const int* __restrict__ cb,
                                                         const int* __restrict__ ce, const double* __restrict__ z,
                                                         R* __restrict__ epart, const int2* __restrict__ erec) {
  constexpr int CH = ea_chunk_obs<CNP>();
  static_assert(CH * CNP * sizeof(R) + 2 * CH * sizeof(int) <= 2 * kPairChunk * 3 * CNP * sizeof(R),
                "the e_a chunk's LDS must fit the Schur staging buffer");
  __shared__ R sBuf[2 * kPairChunk * 3 * CNP];
  __shared__ int sIdx[3 * kPairChunk];
  __shared__ int sMeta[2];
  const int g = (int)blockIdx.x >> 3, l8 = (int)blockIdx.x & 7;
  const int q = g / (period + 1), rm = g - q * (period + 1);
  int gs = -1, eg = -1;
  if (q < neg) {
    if (rm == period) eg = q;
    else gs = q * period + rm;
  } else {
    gs = neg * period + (g - neg * (period + 1));
  }
  if (gs >= 0) {
    const int wg = 8 * gs + l8;
    if (wg >= nb_schur) return;
    schur_chunk<CNP, R, true>(wg, nullptr, nullptr, nullptr, nullptr, W, Vinv, spart, nullptr, nullptr, nchunk,
                              nullptr, rec, sBuf, sIdx, sMeta);
  } else {
    const int b = 8 * eg + l8;
    if (b >= ncc) return;
    int* const sO = (int*)(sBuf + CH * CNP);
    ea_chunk<CNP, R>(b, nullptr, cb, ce, nullptr, W, z, epart, nullptr, nullptr, erec, sBuf, sO, sO + CH);
  }
}

// e_a entry t (camera t / CNP, row t % CNP)
template <int CNP, typename R>
__device__ __forceinline__ void ea_finish_one(int t, const int* __restrict__ ccp, const R* __restrict__ part,
                                              const double* __restrict__ ga, int nc, double* __restrict__ ea,
                                              const int* __restrict__ ownr, const unsigned char* __restrict__ run) {
  if (t >= nc * CNP) return;
  int j = t / CNP, r = t - j * CNP;
  if (run && !run[ownr[j]]) return;
  const double acc = ordered_sum<16>(ccp[j], ccp[j + 1], [&](int q) { return (double)part[(size_t)q * CNP + r]; });
  ea[t] = ga[t] - acc;
}

template <int CNP, typename R>
__global__ __launch_bounds__(64) void k_ea_finish(const int* __restrict__ ccp, const R* __restrict__ part,
                                                  const double* __restrict__ ga, int nc, double* __restrict__ ea,
                                                  const int* __restrict__ ownr, const unsigned char* __restrict__ run) {
  ea_finish_one<CNP, R>(blockIdx.x * blockDim.x + threadIdx.x, ccp, part, ga, nc, ea, ownr, run);
}

// The Schur finish and k_ea_finish in one launch (one problem): workgroups
// [0, nfin) each write one lower block of S (schur_finish_block), the rest
// take blockDim e_a entries each (the same per-entry sums as k_ea_finish)
template <int CNP, typename R>
__global__ __launch_bounds__(256) void k_schur_ea_finish(const int* __restrict__ lblock_pair,
                                                         const int* __restrict__ psp, const R* __restrict__ spart,
                                                         const double* __restrict__ U,
                                                         const unsigned char* __restrict__ fix,
                                                         const double* __restrict__ mu_p, int ld, int nlb, int n,
                                                         double* __restrict__ S, int nfin, const int* __restrict__ ccp,
                                                         const R* __restrict__ epart, const double* __restrict__ ga,
                                                         int nc, double* __restrict__ ea, double mu) {
  if ((int)blockIdx.x < nfin)
    schur_finish_block<CNP, R>(blockIdx.x, lblock_pair, psp, spart, U, fix, mu, mu_p, ld, nlb, n, S);
  else
    ea_finish_one<CNP, R>(((int)blockIdx.x - nfin) * blockDim.x + threadIdx.x, ccp, epart, ga, nc, ea, nullptr,
                          nullptr);
}

template <int CNP, typename R>
static void run_schur_t(psba_ctx* c, double mu, const R* W, R* schpart, R* eapart) {
  // the Schur and e_a chunks in one launch (the e_a chunks after the Schur
  // chunks, so they fill the CUs that the Schur launch's tail leaves idle),
  // then both finishes in one launch
  const int nb_schur = c->n_schunk > 0 ? (int)(PSBA_SCHUR_XCD ? xcd_grid(c->n_schunk) : c->n_schunk) : 0;
  const int nsg = (nb_schur + 7) / 8, neg = (c->n_cchunk + 7) / 8;
  if (nsg + neg > 0) {
    // (with fewer Schur groups than e_a groups, period 1 and the surplus
    // Schur slots of the interleaved region return at once)
    const int period = neg > 0 ? std::max(1, nsg / neg) : 1;
    const int ngroups = neg * (period + 1) + std::max(0, nsg - neg * period);
    probe_begin(c, PSBA_PROBE_SCHUR);
    hipLaunchKernelGGL((k_schur_ea_chunks<CNP, R>), dim3(8 * ngroups), dim3(256), 0, c->stream, W, c->Vinv,
                       schpart, c->n_schunk, (const int4*)c->slot_items, nb_schur, neg, period, c->n_cchunk,
                       c->cchunk_begin, c->cchunk_end, c->z, eapart, (const int2*)c->cchunk_rec);
    probe_end(c, PSBA_PROBE_SCHUR);
  }
  const int nfin = c->n_lblocks + (c->n_pad - c->n), bs = CNP * CNP > 128 ? 256 : 128;
  hipLaunchKernelGGL((k_schur_ea_finish<CNP, R>), dim3(nfin + ceil_div((long)c->nc * CNP, bs)), dim3(bs), 0,
                     c->stream, c->lblock_pair, c->pair_schunk_ptr, schpart, c->U, c->fix, mu_of(c), c->n_pad,
                     c->n_lblocks, c->n, c->S, nfin, c->cam_cchunk_ptr, eapart, c->g, c->nc, c->ea, mu);
}

template <int CNP>
static void run_schur(psba_ctx* c, double mu) {
  if (c->mixed())
    run_schur_t<CNP, float>(c, mu, c->W32, c->schpart32, c->eapart32);
  else
    run_schur_t<CNP, double>(c, mu, c->W, c->schpart, c->eapart);
}

// e_a = ga - sum_i W_ij z_i with caller-given ga (n) and z (3 n_p) -> out (n)
template <int CNP>
static void run_ea_rhs(psba_ctx* c, const double* ga, const double* z, double* out) {
  if (c->mixed()) {
    if (c->n_cchunk > 0)
      hipLaunchKernelGGL((k_ea_chunks<CNP, float>), dim3(c->n_cchunk), dim3(256), 0, c->stream, c->cam_obs,
                         c->cchunk_begin, c->cchunk_end, c->obs_pt, c->W32, z, c->eapart32, (const int*)nullptr,
                         (const unsigned char*)nullptr, (const int2*)c->cchunk_rec);
    hipLaunchKernelGGL((k_ea_finish<CNP, float>), dim3(ceil_div((long)c->nc * CNP, 64)), dim3(64), 0, c->stream,
                       c->cam_cchunk_ptr, c->eapart32, ga, c->nc, out, (const int*)nullptr,
                       (const unsigned char*)nullptr);
  } else {
    if (c->n_cchunk > 0)
      hipLaunchKernelGGL((k_ea_chunks<CNP, double>), dim3(c->n_cchunk), dim3(256), 0, c->stream, c->cam_obs,
                         c->cchunk_begin, c->cchunk_end, c->obs_pt, c->W, z, c->eapart, (const int*)nullptr,
                         (const unsigned char*)nullptr, (const int2*)c->cchunk_rec);
    hipLaunchKernelGGL((k_ea_finish<CNP, double>), dim3(ceil_div((long)c->nc * CNP, 64)), dim3(64), 0, c->stream,
                       c->cam_cchunk_ptr, c->eapart, ga, c->nc, out, (const int*)nullptr,
                       (const unsigned char*)nullptr);
  }
}

int launch_ea_rhs(psba_ctx* c, const double* ga, const double* z, double* out) {
  PSBA_DISPATCH_CNP(c->cnp, run_ea_rhs, c, ga, z, out);
  PSBA_LAUNCH_TRY();
  return PSBA_OK;
}

int launch_schur(psba_ctx* c, double mu) {
  switch (c->cnp) {
    case 6: run_schur<6>(c, mu); break;
    case 9: run_schur<9>(c, mu); break;
    case 11: run_schur<11>(c, mu); break;
    case 16: run_schur<16>(c, mu); break;
    default: return PSBA_ERR_UNSUPPORTED;
  }
  PSBA_LAUNCH_TRY();
  return PSBA_OK;
}

}  // namespace psba

// ---------------------------------------------------------------------------
// batched problems
// ---------------------------------------------------------------------------
#include "batch.hpp"

namespace psba {

// one block per lower block (problem p, local cameras k >= l)
template <int CNP, typename R>
__global__ __launch_bounds__(256) void k_b_schur_finish(const int* __restrict__ lb_prob, const int* __restrict__ lb_k,
                                                        const int* __restrict__ lb_l, const int* __restrict__ lb_pair,
                                                        const int* __restrict__ psp, const R* __restrict__ part,
                                                        const double* __restrict__ U,
                                                        const unsigned char* __restrict__ fix,
                                                        const double* __restrict__ mu_p, const int* __restrict__ cam_off,
                                                        const int* __restrict__ npad, const long* __restrict__ S_off,
                                                        const unsigned char* __restrict__ run, double* __restrict__ S) {
  constexpr int C2 = CNP * CNP;
  const int b = blockIdx.x;
  const int p = lb_prob[b];
  if (!run[p]) return;
  const int k = lb_k[b], l = lb_l[b], q = lb_pair[b];
  const int kg = cam_off[p] + k, ld = npad[p];
  const double mu = mu_p[p];
  double* Sp = S + S_off[p];
  for (int t = threadIdx.x; t < C2; t += blockDim.x) {
    const int r = t / CNP, cc = t - r * CNP;
    double acc = 0.0;
    if (q >= 0) acc = ordered_sum<16>(psp[q], psp[q + 1], [&](int h) { return (double)part[(size_t)h * C2 + t]; });
    double val;
    if (k == l) {
      double u = U[(size_t)kg * C2 + t];
      if (r == cc) u = u + mu;
      val = u - acc;
      if (r == cc && fix[(size_t)kg * CNP + r]) val = 1.0;
    } else {
      val = -acc;
    }
    Sp[(size_t)(k * CNP + r) * ld + l * CNP + cc] = val;
  }
}

// padding rows n..n_pad-1 of every problem's S: identity
__global__ __launch_bounds__(256) void k_b_pad(const int* __restrict__ cam_off, const int* __restrict__ npad,
                                               const long* __restrict__ S_off, int cnp,
                                               const unsigned char* __restrict__ run, double* __restrict__ S) {
  const int p = blockIdx.x;
  if (!run[p]) return;
  const int n = (cam_off[p + 1] - cam_off[p]) * cnp, ld = npad[p];
  double* Sp = S + S_off[p];
  for (long f = threadIdx.x; f < (long)(ld - n) * ld; f += blockDim.x) {
    const int r = n + (int)(f / ld), q = (int)(f % ld);
    if (q <= r) Sp[(size_t)r * ld + q] = (q == r) ? 1.0 : 0.0;
  }
}

template <int CNP, typename R>
static void run_b_schur_t(psba_batch* b) {
  const R* W = (const R*)b->W;
  R *schpart = (R*)b->schpart, *eapart = (R*)b->eapart;
  if (b->n_schunk > 0)
    hipLaunchKernelGGL((k_schur_chunks<CNP, R, false>), dim3(PSBA_SCHUR_XCD ? xcd_grid(b->n_schunk) : b->n_schunk),
                       dim3(256), 0, b->stream, b->items, b->schunk_b, b->schunk_e, b->obs_pt, W, b->Vinv, schpart,
                       b->schunk_prob, b->active, b->n_schunk, (const int*)nullptr, (const int4*)nullptr);
  if (b->nlb > 0)
    hipLaunchKernelGGL((k_b_schur_finish<CNP, R>), dim3(b->nlb), dim3(CNP * CNP > 128 ? 256 : 128), 0, b->stream,
                       b->lb_prob, b->lb_k, b->lb_l, b->lb_pair, b->psp, schpart, b->U, b->fix, b->mu, b->cam_off,
                       b->npad, b->S_off, b->active, b->S);
  hipLaunchKernelGGL(k_b_pad, dim3(b->P), dim3(256), 0, b->stream, b->cam_off, b->npad, b->S_off, CNP, b->active,
                     b->S);
  if (b->n_cchunk > 0)
    hipLaunchKernelGGL((k_ea_chunks<CNP, R>), dim3(b->n_cchunk), dim3(256), 0, b->stream, b->cam_obs, b->cchunk_b,
                       b->cchunk_e, b->obs_pt, W, b->z, eapart, b->cchunk_prob, b->active, (const int2*)nullptr);
  hipLaunchKernelGGL((k_ea_finish<CNP, R>), dim3(ceil_div(b->NC * CNP, 64)), dim3(64), 0, b->stream, b->ccp, eapart,
                     b->g, (int)b->NC, b->ea, b->cam_prob, b->active);
}

template <int CNP>
static void run_b_schur(psba_batch* b) {
  if (b->mixed())
    run_b_schur_t<CNP, float>(b);
  else
    run_b_schur_t<CNP, double>(b);
}

int launch_b_schur(psba_batch* b) {
  PSBA_DISPATCH_CNP(b->cnp, run_b_schur, b);
  PSBA_LAUNCH_TRY();
  return PSBA_OK;
}

}  // namespace psba
