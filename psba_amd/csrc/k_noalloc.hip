// generate_idxs on the device without any allocation, for callers that cannot
// call hipMalloc (code under stream capture, serving loops that forbid it).
//
// The observation order of PSBA/misc.cpp:178-218 by compaction of the
// visibility mask, as psba_generate_idxs_device computes it, but with device
// pointers in and out and the only intermediate -- one count per workgroup of
// the mask -- in a file-scope device array of fixed size.  Three launches on
// the caller's stream:
//   k_na_counts   visible entries per workgroup of kNaBlock mask entries;
//   k_na_scan     exclusive scan of those counts in place (one workgroup) and
//                 the total, which goes to *d_count;
//   k_na_scatter  each visible (point, camera) gets its observation number:
//                 iidx / jidx at that number, blk_idx[point * nCams + camera]
//                 the number or -1.
// Every count is written by the first launch before it is read, so nothing
// has to be cleared between calls.  The array is shared by every call on the
// device: calls that overlap on one device race.
#include <mutex>
#include <vector>

#include "context.hpp"

namespace psba {

namespace {

constexpr int kNaThreads = 256;
constexpr int kNaPerThread = 4;
constexpr int kNaBlock = kNaThreads * kNaPerThread;  // mask entries per workgroup
constexpr int kNaMaxBlocks = 65536;                  // workgroup counts the scratch holds

// per-workgroup counts, then their total
__device__ __align__(256) int scratch[kNaMaxBlocks + 1];

// Device address of `scratch` on each device, taken once per device with the
// device current.  The host-side name of the array is not a device address;
// only these pointers are passed to kernels.
std::mutex g_scratch_mu;
std::vector<int*> g_scratch_ptr;

int scratch_on(int device, int** out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if ((int)g_scratch_ptr.size() <= device) g_scratch_ptr.resize((size_t)device + 1, nullptr);
  if (!g_scratch_ptr[device]) {
    int* p = nullptr;
    const hipError_t e = hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_last_error(__FILE__, __LINE__, "hipGetSymbolAddress of the index scratch failed");
      return PSBA_ERR_HIP;
    }
    g_scratch_ptr[device] = p;
  }
  *out = g_scratch_ptr[device];
  return PSBA_OK;
}

// inclusive prefix sum over the 64 lanes of a wavefront
__device__ __forceinline__ int wave_scan_incl(int v, int lane) {
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int u = __shfl_up(v, off, 64);
    if (lane >= off) v += u;
  }
  return v;
}

// exclusive prefix of v over the workgroup (256 threads); *total gets the sum
__device__ __forceinline__ int block_scan_excl(int v, int* sh, int* total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int incl = wave_scan_incl(v, lane);
  if (lane == 63) sh[w] = incl;
  __syncthreads();
  int before = 0;
  for (int k = 0; k < w; ++k) before += sh[k];
  if (total) *total = sh[0] + sh[1] + sh[2] + sh[3];
  __syncthreads();
  return before + incl - v;
}

}  // namespace

__global__ __launch_bounds__(kNaThreads) void k_na_counts(const char* __restrict__ vmask, long n,
                                                          int* __restrict__ cnt) {
  __shared__ int sh[4];
  const long base = (long)blockIdx.x * kNaBlock + (long)threadIdx.x * kNaPerThread;
  int c = 0;
#pragma unroll
  for (int u = 0; u < kNaPerThread; ++u) c += (base + u < n && vmask[base + u]) ? 1 : 0;
  int total = 0;
  (void)block_scan_excl(c, sh, &total);
  if (threadIdx.x == 0) cnt[blockIdx.x] = total;
}

// one workgroup: cnt[0..nb) -> exclusive prefix in place, cnt[nb] = total
__global__ __launch_bounds__(kNaThreads) void k_na_scan(int* __restrict__ cnt, int nb, int* __restrict__ d_count) {
  __shared__ int sh[4];
  const int per = (nb + kNaThreads - 1) / kNaThreads;
  const int b0 = (int)threadIdx.x * per, b1 = min(nb, b0 + per);
  int s = 0;
  for (int b = b0; b < b1; ++b) s += cnt[b];
  int total = 0;
  int run = block_scan_excl(s, sh, &total);
  for (int b = b0; b < b1; ++b) {
    const int v = cnt[b];
    cnt[b] = run;
    run += v;
  }
  if (threadIdx.x == 0) {
    cnt[nb] = total;
    if (d_count) *d_count = total;
  }
}

__global__ __launch_bounds__(kNaThreads) void k_na_scatter(const char* __restrict__ vmask, long n, int nc,
                                                           const int* __restrict__ cnt, int n_obs,
                                                           int* __restrict__ iidx, int* __restrict__ jidx,
                                                           int* __restrict__ blk) {
  __shared__ int sh[4];
  const long base = (long)blockIdx.x * kNaBlock + (long)threadIdx.x * kNaPerThread;
  bool f[kNaPerThread];
  int c = 0;
#pragma unroll
  for (int u = 0; u < kNaPerThread; ++u) {
    f[u] = base + u < n && vmask[base + u];
    c += f[u] ? 1 : 0;
  }
  int k = cnt[blockIdx.x] + block_scan_excl(c, sh, nullptr);
#pragma unroll
  for (int u = 0; u < kNaPerThread; ++u) {
    const long q = base + u;
    if (q >= n) break;
    if (f[u]) {
      if (k < n_obs) {
        iidx[k] = (int)(q / nc);
        jidx[k] = (int)(q % nc);
      }
      blk[q] = k < n_obs ? k : -1;
      ++k;
    } else {
      blk[q] = -1;
    }
  }
}

}  // namespace psba

extern "C" int psba_generate_idxs_device_noalloc(int device, void* stream, int nCams, int n3Dpts, int n2Dprojs,
                                                 const char* d_vmask, int* d_iidx, int* d_jidx, int* d_blk_idx,
                                                 int* d_count) {
  using namespace psba;
  if (nCams <= 0 || n3Dpts <= 0 || n2Dprojs < 0 || !d_vmask || !d_iidx || !d_jidx || !d_blk_idx) {
    set_last_error(__FILE__, __LINE__, "generate_idxs (no allocation): bad argument");
    return PSBA_ERR_ARG;
  }
  const long n = (long)nCams * n3Dpts;
  if (n > (long)kNaMaxBlocks * kNaBlock) {
    char buf[160];
    snprintf(buf, sizeof(buf), "generate_idxs (no allocation): nCams * n3Dpts = %ld exceeds the limit of %ld", n,
             (long)kNaMaxBlocks * kNaBlock);
    set_last_error(__FILE__, __LINE__, buf);
    return PSBA_ERR_ARG;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    set_last_error(__FILE__, __LINE__, "no HIP device visible");
    return PSBA_ERR_NODEVICE;
  }
  if (device < 0 || device >= ndev) {
    set_last_error(__FILE__, __LINE__, "device out of range");
    return PSBA_ERR_ARG;
  }
  PSBA_HIP_TRY(hipSetDevice(device));
  int* cnt = nullptr;
  PSBA_TRY(scratch_on(device, &cnt));
  hipStream_t st = (hipStream_t)stream;
  const int nb = (int)((n + kNaBlock - 1) / kNaBlock);
  hipLaunchKernelGGL(k_na_counts, dim3(nb), dim3(kNaThreads), 0, st, d_vmask, n, cnt);
  hipLaunchKernelGGL(k_na_scan, dim3(1), dim3(kNaThreads), 0, st, cnt, nb, d_count);
  hipLaunchKernelGGL(k_na_scatter, dim3(nb), dim3(kNaThreads), 0, st, d_vmask, n, nCams, cnt, n2Dprojs, d_iidx,
                     d_jidx, d_blk_idx);
  PSBA_LAUNCH_TRY();
  return PSBA_OK;
}
