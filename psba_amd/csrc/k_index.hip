// Index construction on the device: the observation order of generate_idxs
// (PSBA/misc.cpp:178-218) by compaction of the visibility mask, and the
// context's index (point CSR, camera-ordered observations, camera chunks,
// camera-pair items in ascending point order, pair chunks, lower-block map)
// from the observation order.  Every array is bit-identical to the host
// construction (psba_generate_idxs, psba_set_index): the only reorderings are
// stable radix sorts of items generated in point-major order, which keep
// ascending observation / point order within a key exactly as the host loops
// do.  Integer counts use integer atomics, whose result does not depend on
// their order.
#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_run_length_encode.hpp>
#include <rocprim/device/device_scan.hpp>

#include "context.hpp"

namespace psba {

namespace {

template <typename T>
struct DevBuf {
  T* p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  int alloc(size_t n) {
    if (hipMalloc((void**)&p, sizeof(T) * (n ? n : 1)) != hipSuccess) {
      p = nullptr;
      set_last_error(__FILE__, __LINE__, "hipMalloc failed in the index build");
      return PSBA_ERR_NOMEM;
    }
    return PSBA_OK;
  }
};

struct Scratch {
  void* p = nullptr;
  size_t bytes = 0;
  ~Scratch() {
    if (p) (void)hipFree(p);
  }
  int ensure(size_t b) {
    if (b <= bytes) return PSBA_OK;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    if (hipMalloc(&p, b) != hipSuccess) {
      p = nullptr;
      set_last_error(__FILE__, __LINE__, "hipMalloc failed in the index build");
      return PSBA_ERR_NOMEM;
    }
    bytes = b;
    return PSBA_OK;
  }
};

int bits_for(unsigned long long v) {
  int b = 1;
  while (b < 64 && (1ull << b) <= v) ++b;
  return b;
}

}  // namespace

// ptr[p] = first observation of p (observations sorted by `key`, keys in
// [0, nkeys)); keys with no observation get an empty range
__global__ void k_ptr_from_sorted(const int* __restrict__ key, int n, int nkeys, int* __restrict__ ptr) {
  const int o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= n) return;
  const int prev = (o == 0) ? -1 : key[o - 1];
  for (int p = prev + 1; p <= key[o]; ++p) ptr[p] = o;
  if (o == n - 1)
    for (int p = key[o] + 1; p <= nkeys; ++p) ptr[p] = n;
}

__global__ void k_iota(int* __restrict__ v, int n) {
  const int o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o < n) v[o] = o;
}

// per camera: number of chunks of CH observations; per point: t(t+1)/2 items
__global__ void k_chunk_counts(const int* __restrict__ ptr, int n, int CH, int* __restrict__ cnt) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) cnt[j] = (ptr[j + 1] - ptr[j] + CH - 1) / CH;
}

__global__ void k_item_counts(const int* __restrict__ pt_ptr, int np, long long* __restrict__ cnt,
                              long long* __restrict__ t2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const long long t = pt_ptr[i + 1] - pt_ptr[i];
  cnt[i] = t * (t + 1) / 2;
  if (t) atomicAdd((unsigned long long*)t2, (unsigned long long)(t * t));
}

// launch key of each pair chunk (its first item's cameras and point)
__global__ void k_schunk_keys(const int2* __restrict__ items, const int* __restrict__ sbeg, int n,
                              const int* __restrict__ obs_cam, const int* __restrict__ obs_pt,
                              unsigned long long* __restrict__ key, int* __restrict__ val) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const int2 it = items[sbeg[q]];
  key[q] = schunk_launch_key(obs_cam[it.x], obs_cam[it.y], obs_pt[it.x]);
  val[q] = q;
}

// chunk ranges of each segment: [seg_ptr[s], seg_ptr[s+1]) cut every CH
__global__ void k_fill_chunks(const int* __restrict__ seg_ptr, int nseg, const int* __restrict__ chunk_ptr, int CH,
                              int* __restrict__ cb, int* __restrict__ ce) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nseg) return;
  int q = chunk_ptr[s];
  for (int a = seg_ptr[s]; a < seg_ptr[s + 1]; a += CH, ++q) {
    cb[q] = a;
    ce[q] = min(a + CH, seg_ptr[s + 1]);
  }
}

// items of point i, in the host's loop order: a over the point's
// observations, b from its first observation up to a
__global__ void k_gen_items(const int* __restrict__ pt_ptr, int np, const int* __restrict__ jidx, int nc,
                            const long long* __restrict__ item_off, unsigned* __restrict__ key,
                            unsigned long long* __restrict__ val) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  long long q = item_off[i];
  for (int a = pt_ptr[i]; a < pt_ptr[i + 1]; ++a)
    for (int b = pt_ptr[i]; b <= a; ++b, ++q) {
      key[q] = (unsigned)jidx[a] * (unsigned)nc + (unsigned)jidx[b];
      val[q] = ((unsigned long long)(unsigned)a << 32) | (unsigned)b;
    }
}

__global__ void k_unpack_items(const unsigned long long* __restrict__ val, long n, int2* __restrict__ items) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) items[q] = make_int2((int)(val[q] >> 32), (int)(val[q] & 0xffffffffu));
}

__global__ void k_pairs(const unsigned* __restrict__ ukey, int npairs, int nc, int* __restrict__ pk,
                        int* __restrict__ pl, int* __restrict__ lbp) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= npairs) return;
  const int k = (int)(ukey[q] / (unsigned)nc), l = (int)(ukey[q] % (unsigned)nc);
  pk[q] = k;
  pl[q] = l;
  lbp[k * (k + 1) / 2 + l] = q;
}

__global__ void k_widen(const int* __restrict__ a, int n, long long* __restrict__ b) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) b[q] = a[q];
}

__global__ void k_fill_int(int* __restrict__ v, long n, int x) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) v[q] = x;
}

__global__ void k_pair_ptr(const int* __restrict__ counts, int npairs, const long long* __restrict__ start,
                           int* __restrict__ pptr) {
  // pair q owns items [start[q], start[q] + counts[q])
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < npairs) pptr[q] = (int)start[q];
  if (q == npairs - 1) pptr[npairs] = (int)(start[q] + counts[q]);
}

// generate_idxs on the device: flags of the visibility mask, their exclusive
// scan gives each visible (point, camera) its observation number
__global__ void k_mask_flags(const char* __restrict__ vmask, long n, int* __restrict__ f) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) f[q] = vmask[q] ? 1 : 0;
}

__global__ void k_mask_scatter(const char* __restrict__ vmask, const int* __restrict__ pos, long n, int nc,
                               int* __restrict__ iidx, int* __restrict__ jidx, int* __restrict__ blk) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  if (vmask[q]) {
    const int k = pos[q];
    iidx[k] = (int)(q / nc);
    jidx[k] = (int)(q % nc);
    blk[q] = k;
  } else {
    blk[q] = -1;
  }
}

// comm3DIdxCnt: shared points of every ordered camera pair (symmetric)
__global__ void k_pair_counts(const int* __restrict__ pt_ptr, int np, const int* __restrict__ jidx, int nc,
                              int* __restrict__ cnt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  for (int a = pt_ptr[i]; a < pt_ptr[i + 1]; ++a)
    for (int b = pt_ptr[i]; b <= a; ++b) {
      const int j = jidx[a], l = jidx[b];
      atomicAdd(&cnt[j * nc + l], 1);
      if (l != j) atomicAdd(&cnt[l * nc + j], 1);
    }
}

}  // namespace psba

using namespace psba;

#define IX_TRY(expr)                      \
  do {                                    \
    hipError_t _e = (expr);               \
    if (_e != hipSuccess) {               \
      set_last_error(__FILE__, __LINE__, hipGetErrorString(_e)); \
      return PSBA_ERR_HIP;                \
    }                                     \
  } while (0)

extern "C" {

int psba_set_index_device(psba_ctx* c, const int* iidx, const int* jidx) {
  if (!c || !iidx || !jidx) {
    set_last_error(__FILE__, __LINE__, "null argument");
    return PSBA_ERR_ARG;
  }
  IX_TRY(hipSetDevice(c->device));
  const int nc = c->nc, np = c->np, no = c->no, C = c->cnp;
  PSBA_TRY(check_obs_order(nc, np, no, iidx, jidx));
  hipStream_t st = c->stream;
  const int CH = cam_chunk_size(C);
  auto grid = [](long n) { return dim3((unsigned)((n + 255) / 256)); };
  Scratch tmp;
  size_t tb = 0;

  DevBuf<int> d_i, d_j, ptptr, camptr, camkey, camobs, ord, ccnt, ccp;
  PSBA_TRY(d_i.alloc(no));
  PSBA_TRY(d_j.alloc(no));
  PSBA_TRY(ptptr.alloc(np + 1));
  PSBA_TRY(camptr.alloc(nc + 1));
  PSBA_TRY(camkey.alloc(no));
  PSBA_TRY(camobs.alloc(no));
  PSBA_TRY(ord.alloc(no));
  PSBA_TRY(ccnt.alloc(nc));
  PSBA_TRY(ccp.alloc(nc + 1));
  IX_TRY(hipMemcpyAsync(d_i.p, iidx, sizeof(int) * no, hipMemcpyHostToDevice, st));
  IX_TRY(hipMemcpyAsync(d_j.p, jidx, sizeof(int) * no, hipMemcpyHostToDevice, st));
  // point CSR (observations are point-major)
  hipLaunchKernelGGL(k_ptr_from_sorted, grid(no), dim3(256), 0, st, d_i.p, no, np, ptptr.p);
  // camera order: stable sort of observation numbers by camera
  hipLaunchKernelGGL(k_iota, grid(no), dim3(256), 0, st, ord.p, no);
  const int cbits = bits_for((unsigned long long)nc);
  IX_TRY(rocprim::radix_sort_pairs(nullptr, tb, d_j.p, camkey.p, ord.p, camobs.p, (size_t)no, 0, cbits, st));
  PSBA_TRY(tmp.ensure(tb));
  IX_TRY(rocprim::radix_sort_pairs(tmp.p, tb, d_j.p, camkey.p, ord.p, camobs.p, (size_t)no, 0, cbits, st));
  hipLaunchKernelGGL(k_ptr_from_sorted, grid(no), dim3(256), 0, st, camkey.p, no, nc, camptr.p);
  // camera chunks
  hipLaunchKernelGGL(k_chunk_counts, grid(nc), dim3(256), 0, st, camptr.p, nc, CH, ccnt.p);
  IX_TRY(rocprim::exclusive_scan(nullptr, tb, ccnt.p, ccp.p, 0, (size_t)nc, rocprim::plus<int>(), st));
  PSBA_TRY(tmp.ensure(tb));
  IX_TRY(rocprim::exclusive_scan(tmp.p, tb, ccnt.p, ccp.p, 0, (size_t)nc, rocprim::plus<int>(), st));
  // items per point and their offsets; sum of t^2
  DevBuf<long long> icnt, ioff, t2;
  PSBA_TRY(icnt.alloc(np));
  PSBA_TRY(ioff.alloc(np));
  PSBA_TRY(t2.alloc(2));
  IX_TRY(hipMemsetAsync(t2.p, 0, sizeof(long long) * 2, st));
  hipLaunchKernelGGL(k_item_counts, grid(np), dim3(256), 0, st, ptptr.p, np, icnt.p, t2.p);
  IX_TRY(rocprim::exclusive_scan(nullptr, tb, icnt.p, ioff.p, 0LL, (size_t)np, rocprim::plus<long long>(), st));
  PSBA_TRY(tmp.ensure(tb));
  IX_TRY(rocprim::exclusive_scan(tmp.p, tb, icnt.p, ioff.p, 0LL, (size_t)np, rocprim::plus<long long>(), st));
  long long h_last[2] = {0, 0}, h_t2 = 0;
  int h_cc[2] = {0, 0};
  IX_TRY(hipMemcpyAsync(&h_last[0], ioff.p + (np - 1), sizeof(long long), hipMemcpyDeviceToHost, st));
  IX_TRY(hipMemcpyAsync(&h_last[1], icnt.p + (np - 1), sizeof(long long), hipMemcpyDeviceToHost, st));
  IX_TRY(hipMemcpyAsync(&h_t2, t2.p, sizeof(long long), hipMemcpyDeviceToHost, st));
  IX_TRY(hipMemcpyAsync(&h_cc[0], ccp.p + (nc - 1), sizeof(int), hipMemcpyDeviceToHost, st));
  IX_TRY(hipMemcpyAsync(&h_cc[1], ccnt.p + (nc - 1), sizeof(int), hipMemcpyDeviceToHost, st));
  IX_TRY(hipStreamSynchronize(st));
  const long long nitems = h_last[0] + h_last[1];
  const int n_cchunk = h_cc[0] + h_cc[1];
  if (nitems > 0x7fffffffLL) {
    set_last_error(__FILE__, __LINE__, "too many camera-pair items");
    return PSBA_ERR_UNSUPPORTED;
  }
  // items in point-major order, keyed by camera pair; a stable sort by key
  // leaves each pair's items in ascending point order
  DevBuf<unsigned> ikey, skey, ukey;
  DevBuf<unsigned long long> ival, sval;
  DevBuf<int> pcnt, npairs_d;
  PSBA_TRY(ikey.alloc(nitems));
  PSBA_TRY(skey.alloc(nitems));
  PSBA_TRY(ival.alloc(nitems));
  PSBA_TRY(sval.alloc(nitems));
  hipLaunchKernelGGL(k_gen_items, grid(np), dim3(256), 0, st, ptptr.p, np, d_j.p, nc, ioff.p, ikey.p, ival.p);
  const int kbits = bits_for((unsigned long long)nc * nc);
  IX_TRY(rocprim::radix_sort_pairs(nullptr, tb, ikey.p, skey.p, ival.p, sval.p, (size_t)nitems, 0, kbits, st));
  PSBA_TRY(tmp.ensure(tb));
  IX_TRY(rocprim::radix_sort_pairs(tmp.p, tb, ikey.p, skey.p, ival.p, sval.p, (size_t)nitems, 0, kbits, st));
  const int maxpairs = nc * (nc + 1) / 2;
  PSBA_TRY(ukey.alloc(maxpairs));
  PSBA_TRY(pcnt.alloc(maxpairs));
  PSBA_TRY(npairs_d.alloc(1));
  IX_TRY(rocprim::run_length_encode(nullptr, tb, skey.p, (size_t)nitems, ukey.p, pcnt.p, npairs_d.p, st));
  PSBA_TRY(tmp.ensure(tb));
  IX_TRY(rocprim::run_length_encode(tmp.p, tb, skey.p, (size_t)nitems, ukey.p, pcnt.p, npairs_d.p, st));
  int npairs = 0;
  IX_TRY(hipMemcpyAsync(&npairs, npairs_d.p, sizeof(int), hipMemcpyDeviceToHost, st));
  IX_TRY(hipStreamSynchronize(st));
  // pair starts, pair chunks
  DevBuf<long long> pstart, pcnt64;
  DevBuf<int> pptr, scnt, psp;
  PSBA_TRY(pstart.alloc(npairs));
  PSBA_TRY(pcnt64.alloc(npairs));
  PSBA_TRY(pptr.alloc(npairs + 1));
  PSBA_TRY(scnt.alloc(npairs));
  PSBA_TRY(psp.alloc(npairs + 1));
  hipLaunchKernelGGL(k_widen, grid(npairs), dim3(256), 0, st, pcnt.p, npairs, pcnt64.p);
  IX_TRY(rocprim::exclusive_scan(nullptr, tb, pcnt64.p, pstart.p, 0LL, (size_t)npairs, rocprim::plus<long long>(),
                                 st));
  PSBA_TRY(tmp.ensure(tb));
  IX_TRY(rocprim::exclusive_scan(tmp.p, tb, pcnt64.p, pstart.p, 0LL, (size_t)npairs, rocprim::plus<long long>(), st));
  hipLaunchKernelGGL(k_pair_ptr, grid(npairs), dim3(256), 0, st, pcnt.p, npairs, pstart.p, pptr.p);
  hipLaunchKernelGGL(k_chunk_counts, grid(npairs), dim3(256), 0, st, pptr.p, npairs, kPairChunk, scnt.p);
  IX_TRY(rocprim::exclusive_scan(nullptr, tb, scnt.p, psp.p, 0, (size_t)npairs, rocprim::plus<int>(), st));
  PSBA_TRY(tmp.ensure(tb));
  IX_TRY(rocprim::exclusive_scan(tmp.p, tb, scnt.p, psp.p, 0, (size_t)npairs, rocprim::plus<int>(), st));
  int h_sc[2] = {0, 0};
  IX_TRY(hipMemcpyAsync(&h_sc[0], psp.p + (npairs - 1), sizeof(int), hipMemcpyDeviceToHost, st));
  IX_TRY(hipMemcpyAsync(&h_sc[1], scnt.p + (npairs - 1), sizeof(int), hipMemcpyDeviceToHost, st));
  IX_TRY(hipStreamSynchronize(st));
  const int n_schunk = h_sc[0] + h_sc[1];

  // the context's buffers
  PSBA_TRY(alloc_index(c, (long)nitems, npairs, n_cchunk, n_schunk));
  c->sum_t2 = (double)h_t2;
  IX_TRY(hipMemcpyAsync(c->obs_pt, d_i.p, sizeof(int) * no, hipMemcpyDeviceToDevice, st));
  IX_TRY(hipMemcpyAsync(c->obs_cam, d_j.p, sizeof(int) * no, hipMemcpyDeviceToDevice, st));
  IX_TRY(hipMemcpyAsync(c->pt_ptr, ptptr.p, sizeof(int) * (np + 1), hipMemcpyDeviceToDevice, st));
  IX_TRY(hipMemcpyAsync(c->cam_ptr, camptr.p, sizeof(int) * (nc + 1), hipMemcpyDeviceToDevice, st));
  IX_TRY(hipMemcpyAsync(c->cam_obs, camobs.p, sizeof(int) * no, hipMemcpyDeviceToDevice, st));
  IX_TRY(hipMemcpyAsync(c->cam_cchunk_ptr, ccp.p, sizeof(int) * nc, hipMemcpyDeviceToDevice, st));
  IX_TRY(hipMemcpyAsync(c->cam_cchunk_ptr + nc, &n_cchunk, sizeof(int), hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_fill_chunks, grid(nc), dim3(256), 0, st, camptr.p, nc, ccp.p, CH, c->cchunk_begin,
                     c->cchunk_end);
  hipLaunchKernelGGL(k_unpack_items, grid((long)nitems), dim3(256), 0, st, sval.p, (long)nitems, c->pair_items);
  hipLaunchKernelGGL(k_fill_int, grid(c->n_lblocks), dim3(256), 0, st, c->lblock_pair, (long)c->n_lblocks, -1);
  hipLaunchKernelGGL(k_pairs, grid(npairs), dim3(256), 0, st, ukey.p, npairs, nc, c->pair_k, c->pair_l,
                     c->lblock_pair);
  IX_TRY(hipMemcpyAsync(c->pair_schunk_ptr, psp.p, sizeof(int) * npairs, hipMemcpyDeviceToDevice, st));
  IX_TRY(hipMemcpyAsync(c->pair_schunk_ptr + npairs, &n_schunk, sizeof(int), hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_fill_chunks, grid(npairs), dim3(256), 0, st, pptr.p, npairs, psp.p, kPairChunk,
                     c->schunk_begin, c->schunk_end);
  // launch order of the pair chunks: a stable sort by schunk_launch_key
  if (n_schunk > 0) {
    DevBuf<unsigned long long> okey, okey_s;
    DevBuf<int> oval;
    PSBA_TRY(okey.alloc(n_schunk));
    PSBA_TRY(okey_s.alloc(n_schunk));
    PSBA_TRY(oval.alloc(n_schunk));
    hipLaunchKernelGGL(k_schunk_keys, grid(n_schunk), dim3(256), 0, st, c->pair_items, c->schunk_begin, n_schunk,
                       c->obs_cam, c->obs_pt, okey.p, oval.p);
    IX_TRY(rocprim::radix_sort_pairs(nullptr, tb, okey.p, okey_s.p, oval.p, c->schunk_order, (size_t)n_schunk, 0,
                                     64, st));
    PSBA_TRY(tmp.ensure(tb));
    IX_TRY(rocprim::radix_sort_pairs(tmp.p, tb, okey.p, okey_s.p, oval.p, c->schunk_order, (size_t)n_schunk, 0, 64,
                                     st));
  }
  IX_TRY(hipGetLastError());
  IX_TRY(hipStreamSynchronize(st));  // the temporaries are freed on return
  PSBA_TRY(build_point_order(c));
  PSBA_TRY(build_schur_slots(c));
  c->indexed = true;
  return PSBA_OK;
}

}  // extern "C"

namespace psba {

__global__ void k_point_order_keys(const int* __restrict__ pt_ptr, int np, unsigned* __restrict__ key,
                                   int* __restrict__ val) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  key[i] = ~(unsigned)(pt_ptr[i + 1] - pt_ptr[i]);  // ascending key = descending length
  val[i] = i;
}

__global__ void k_slot_items(const int* __restrict__ order, const int* __restrict__ sb, const int* __restrict__ se,
                             const int2* __restrict__ items, const int* __restrict__ obs_pt, int n,
                             int4* __restrict__ rec) {
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (long)n * kPairChunk) return;
  const int s = (int)(g / kPairChunk), t = (int)(g - (long)s * kPairChunk);
  const int b = order[s], beg = sb[b];
  if (t < se[b] - beg) {
    const int2 pr = items[beg + t];
    rec[g] = make_int4(pr.x, pr.y, obs_pt[pr.x], b);
  } else {
    rec[g] = make_int4(-1, -1, -1, b);
  }
}

__global__ void k_cchunk_recs(const int* __restrict__ cam_obs, const int* __restrict__ cb, const int* __restrict__ ce,
                              const int* __restrict__ obs_pt, int n, int ch, int2* __restrict__ rec) {
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (long)n * ch) return;
  const int b = (int)(g / ch), s = (int)(g - (long)b * ch);
  const int beg = cb[b];
  if (s < ce[b] - beg) {
    const int o = cam_obs[beg + s];
    rec[g] = make_int2(o, obs_pt[o]);
  } else {
    rec[g] = make_int2(-1, -1);
  }
}

// The Schur chunks' item records in launch order and the camera chunks'
// records, from the finished index (host or device build alike).
int build_schur_slots(psba_ctx* c) {
  hipStream_t st = c->stream;
  const int ch = cam_chunk_size(c->cnp);
  if (c->n_cchunk > 0)
    hipLaunchKernelGGL(k_cchunk_recs, dim3((unsigned)(((long)c->n_cchunk * ch + 255) / 256)), dim3(256), 0, st,
                       c->cam_obs, c->cchunk_begin, c->cchunk_end, c->obs_pt, c->n_cchunk, ch, c->cchunk_rec);
  if (c->n_schunk > 0)
    hipLaunchKernelGGL(k_slot_items, dim3((unsigned)(((long)c->n_schunk * kPairChunk + 255) / 256)), dim3(256), 0,
                       st, c->schunk_order, c->schunk_begin, c->schunk_end, c->pair_items, c->obs_pt, c->n_schunk,
                       c->slot_items);
  PSBA_HIP_TRY(hipGetLastError());
  PSBA_HIP_TRY(hipStreamSynchronize(st));
  return PSBA_OK;
}

// Points by descending track length, ties in ascending id (a stable radix
// sort).  The back substitution takes its points in this order, so that the
// points one wave handles in lockstep have tracks of similar length.
int build_point_order(psba_ctx* c) {
  const int np = c->np;
  hipStream_t st = c->stream;
  DevBuf<unsigned> key, key_s;
  DevBuf<int> val;
  PSBA_TRY(key.alloc(np));
  PSBA_TRY(key_s.alloc(np));
  PSBA_TRY(val.alloc(np));
  hipLaunchKernelGGL(k_point_order_keys, dim3((np + 255) / 256), dim3(256), 0, st, c->pt_ptr, np, key.p, val.p);
  size_t tb = 0;
  IX_TRY(rocprim::radix_sort_pairs(nullptr, tb, key.p, key_s.p, val.p, c->pt_order, (size_t)np, 0, 32, st));
  Scratch tmp;
  PSBA_TRY(tmp.ensure(tb));
  IX_TRY(rocprim::radix_sort_pairs(tmp.p, tb, key.p, key_s.p, val.p, c->pt_order, (size_t)np, 0, 32, st));
  IX_TRY(hipGetLastError());
  IX_TRY(hipStreamSynchronize(st));  // the temporaries are freed on return
  return PSBA_OK;
}

}  // namespace psba

extern "C" {

int psba_generate_idxs_device(int device, int nCams, int n3Dpts, int n2Dprojs, const char* vmask,
                              int* comm3DIdxCnt, int* iidx, int* jidx, int* blk_idx) {
  if (nCams <= 0 || n3Dpts <= 0 || !vmask || !iidx || !jidx || !blk_idx) {
    set_last_error(__FILE__, __LINE__, "generate_idxs: bad argument");
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
  IX_TRY(hipSetDevice(device));
  const long n = (long)nCams * n3Dpts;
  auto grid = [](long m) { return dim3((unsigned)((m + 255) / 256)); };
  hipStream_t st = nullptr;
  IX_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() { (void)hipStreamDestroy(s); }
  } guard{st};
  DevBuf<char> vm;
  DevBuf<int> flags, pos, di, dj, blk;
  PSBA_TRY(vm.alloc(n));
  PSBA_TRY(flags.alloc(n));
  PSBA_TRY(pos.alloc(n));
  PSBA_TRY(blk.alloc(n));
  IX_TRY(hipMemcpyAsync(vm.p, vmask, n, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_mask_flags, grid(n), dim3(256), 0, st, vm.p, n, flags.p);
  Scratch tmp;
  size_t tb = 0;
  IX_TRY(rocprim::exclusive_scan(nullptr, tb, flags.p, pos.p, 0, (size_t)n, rocprim::plus<int>(), st));
  PSBA_TRY(tmp.ensure(tb));
  IX_TRY(rocprim::exclusive_scan(tmp.p, tb, flags.p, pos.p, 0, (size_t)n, rocprim::plus<int>(), st));
  int h[2] = {0, 0};
  IX_TRY(hipMemcpyAsync(&h[0], pos.p + (n - 1), sizeof(int), hipMemcpyDeviceToHost, st));
  IX_TRY(hipMemcpyAsync(&h[1], flags.p + (n - 1), sizeof(int), hipMemcpyDeviceToHost, st));
  IX_TRY(hipStreamSynchronize(st));
  const int count = h[0] + h[1];
  if (count != n2Dprojs) {
    char buf[120];
    snprintf(buf, sizeof(buf), "generate_idxs: vmask has %d entries, expected %d", count, n2Dprojs);
    set_last_error(__FILE__, __LINE__, buf);
    return PSBA_ERR_ARG;
  }
  PSBA_TRY(di.alloc(count));
  PSBA_TRY(dj.alloc(count));
  hipLaunchKernelGGL(k_mask_scatter, grid(n), dim3(256), 0, st, vm.p, pos.p, n, nCams, di.p, dj.p, blk.p);
  if (comm3DIdxCnt) {
    DevBuf<int> ptptr, cnt;
    PSBA_TRY(ptptr.alloc(n3Dpts + 1));
    PSBA_TRY(cnt.alloc((size_t)nCams * nCams));
    IX_TRY(hipMemsetAsync(cnt.p, 0, sizeof(int) * (size_t)nCams * nCams, st));
    hipLaunchKernelGGL(k_ptr_from_sorted, grid(count), dim3(256), 0, st, di.p, count, n3Dpts, ptptr.p);
    hipLaunchKernelGGL(k_pair_counts, grid(n3Dpts), dim3(256), 0, st, ptptr.p, n3Dpts, dj.p, nCams, cnt.p);
    IX_TRY(hipMemcpyAsync(comm3DIdxCnt, cnt.p, sizeof(int) * (size_t)nCams * nCams, hipMemcpyDeviceToHost, st));
    IX_TRY(hipStreamSynchronize(st));
  }
  IX_TRY(hipMemcpyAsync(iidx, di.p, sizeof(int) * count, hipMemcpyDeviceToHost, st));
  IX_TRY(hipMemcpyAsync(jidx, dj.p, sizeof(int) * count, hipMemcpyDeviceToHost, st));
  IX_TRY(hipMemcpyAsync(blk_idx, blk.p, sizeof(int) * n, hipMemcpyDeviceToHost, st));
  IX_TRY(hipGetLastError());
  IX_TRY(hipStreamSynchronize(st));
  return PSBA_OK;
}

}  // extern "C"
