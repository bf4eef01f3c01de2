// Context lifecycle, upload, device index construction and the stage entry
// points of the C-ABI (include/psba_capi.h).
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "context.hpp"

namespace psba {

static thread_local std::string g_last_error;

void set_last_error(const char* file, int line, const char* msg) {
  char buf[512];
  snprintf(buf, sizeof(buf), "%s:%d: %s", file, line, msg);
  g_last_error = buf;
}

static int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

template <typename T>
static int dalloc(T** p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  hipError_t e = hipMalloc((void**)p, count * sizeof(T));
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(PSBA_ERR_NOMEM, "hipMalloc of %zu bytes failed: %s", count * sizeof(T), hipGetErrorString(e));
  }
  return PSBA_OK;
}

template <typename T>
static void dfree(T*& p) {
  if (p) (void)hipFree((void*)p);
  p = nullptr;
}

__global__ void k_pad_identity(double* S, int ld, int n, int n_pad) {
  int r = n + blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n_pad) S[(size_t)r * ld + r] = 1.0;
}

static void free_index(psba_ctx* c) {
  dfree(c->pt_ptr);
  dfree(c->cam_ptr);
  dfree(c->cam_obs);
  dfree(c->cchunk_begin);
  dfree(c->cchunk_end);
  dfree(c->cam_cchunk_ptr);
  dfree(c->pair_items);
  dfree(c->pair_k);
  dfree(c->pair_l);
  dfree(c->schunk_begin);
  dfree(c->schunk_end);
  dfree(c->pair_schunk_ptr);
  dfree(c->schunk_order);
  dfree(c->slot_items);
  dfree(c->cchunk_rec);
  dfree(c->pt_order);
  dfree(c->lblock_pair);
  dfree(c->ucpart);
  dfree(c->eapart);
  dfree(c->schpart);
  dfree(c->ucpart32);
  dfree(c->eapart32);
  dfree(c->schpart32);
  dfree(c->obs_cam);
  dfree(c->obs_pt);
  c->indexed = false;
}

static hipEvent_t probe_event(psba_ctx* c) {
  if (c->probe_pool.empty()) {
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
  }
  hipEvent_t e = c->probe_pool.back();
  c->probe_pool.pop_back();
  return e;
}

void probe_begin(psba_ctx* c, int id) {
  if (c->probe_id != id) return;
  hipEvent_t e = probe_event(c);
  if (!e) return;
  (void)hipEventRecord(e, c->stream);
  c->probe_pending.push_back({e, nullptr});
}

void probe_end(psba_ctx* c, int id) {
  if (c->probe_id != id || c->probe_pending.empty() || c->probe_pending.back().second) return;
  hipEvent_t e = probe_event(c);
  if (!e) return;
  (void)hipEventRecord(e, c->stream);
  c->probe_pending.back().second = e;
}

}  // namespace psba

using namespace psba;

extern "C" {

const char* psba_strerror(int code) {
  switch (code) {
    case PSBA_OK: return "ok";
    case PSBA_ERR_ARG: return "invalid argument";
    case PSBA_ERR_HIP: return "HIP runtime error";
    case PSBA_ERR_NOMEM: return "out of device memory";
    case PSBA_ERR_STATE: return "context not ready (upload and index first)";
    case PSBA_ERR_UNSUPPORTED: return "unsupported configuration";
    case PSBA_ERR_NODEVICE: return "no HIP device";
  }
  if (code > 0) {
    switch (code) {
      case PSBA_ITER_TURN_TO_LM: return "ITER_TURN_TO_LM";
      case PSBA_ITER_TURN_TO_TR: return "ITER_TURN_TO_TR";
      case PSBA_ITER_CONTINUE: return "ITER_CONTINUE";
      case PSBA_ITER_ERR: return "ITER_ERR";
      case PSBA_ITER_DP_NO_CHANGE: return "ITER_DP_NO_CHANGE";
      case PSBA_ITER_ERR_SMALL_ENOUGH: return "ITER_ERR_SMALL_ENOUGH";
      case PSBA_ITER_PASS: return "ITER_PASS";
    }
  }
  return "unknown status";
}

const char* psba_last_error(void) { return g_last_error.c_str(); }

int psba_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int psba_device_info(int device, char* name, int name_cap, int* cus, int* clock_mhz, long long* mem_bytes) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return fail(PSBA_ERR_NODEVICE, "no HIP device visible");
  if (device < 0 || device >= n) return fail(PSBA_ERR_ARG, "device out of range");
  hipDeviceProp_t p;
  PSBA_HIP_TRY(hipGetDeviceProperties(&p, device));
  if (name && name_cap > 0) snprintf(name, (size_t)name_cap, "%s (%s)", p.name, p.gcnArchName);
  if (cus) *cus = p.multiProcessorCount;
  if (clock_mhz) *clock_mhz = p.clockRate / 1000;
  if (mem_bytes) *mem_bytes = (long long)p.totalGlobalMem;
  return PSBA_OK;
}

int psba_create(const psba_desc* d, psba_ctx** out) {
  if (!d || !out) return fail(PSBA_ERR_ARG, "null argument");
  *out = nullptr;
  if (!(d->cnp == 6 || d->cnp == 9 || d->cnp == 11 || d->cnp == 16))
    return fail(PSBA_ERR_UNSUPPORTED, "cnp=%d not supported (6, 9, 11, 16)", d->cnp);
  if (d->pnp != 3 || d->mnp != 2) return fail(PSBA_ERR_UNSUPPORTED, "pnp must be 3 and mnp 2");
  if (d->n_cams <= 0 || d->n_pts <= 0 || d->n_obs <= 0) return fail(PSBA_ERR_ARG, "empty problem");
  if (d->precision != PSBA_PREC_FP64 && d->precision != PSBA_PREC_MIXED)
    return fail(PSBA_ERR_UNSUPPORTED, "precision %d not available", d->precision);
  {
    long n = (long)d->cnp * d->n_cams, n_pad = (n + kNB - 1) / kNB * kNB;
    if (n_pad > kMaxSolveN)
      return fail(PSBA_ERR_UNSUPPORTED, "reduced camera system of order %ld exceeds the limit %d", n, kMaxSolveN);
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(PSBA_ERR_NODEVICE, "no HIP device visible");
  if (d->device < 0 || d->device >= ndev) return fail(PSBA_ERR_ARG, "device %d out of range", d->device);
  PSBA_HIP_TRY(hipSetDevice(d->device));
  psba_ctx* c = new psba_ctx();
  c->cnp = d->cnp;
  c->nc = d->n_cams;
  c->np = d->n_pts;
  c->no = d->n_obs;
  c->n = c->cnp * c->nc;
  c->ntot = c->n + 3 * c->np;
  c->n_pad = ((c->n + kNB - 1) / kNB) * kNB;
  c->precision = d->precision;
  c->device = d->device;
  int rc = PSBA_OK;
  auto chk = [&](int r) {
    if (rc == PSBA_OK && r != PSBA_OK) rc = r;
  };
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) rc = fail(PSBA_ERR_HIP, "stream");
  const int C = c->cnp;
  chk(dalloc(&c->K, 5 * (size_t)c->nc));
  chk(dalloc(&c->q0, 4 * (size_t)c->nc));
  chk(dalloc(&c->meas, 2 * (size_t)c->no));
  chk(dalloc(&c->cams, (size_t)C * c->nc));
  chk(dalloc(&c->pts, 3 * (size_t)c->np));
  chk(dalloc(&c->fix, (size_t)C * c->nc));
  chk(dalloc(&c->e, 2 * (size_t)c->no));
  if (c->precision == PSBA_PREC_MIXED) {
    chk(dalloc(&c->e32, 2 * (size_t)c->no));
    chk(dalloc(&c->JA32, 2 * (size_t)C * c->no));
    chk(dalloc(&c->JB32, 6 * (size_t)c->no));
    chk(dalloc(&c->W32, 3 * (size_t)C * c->no));
  } else {
    chk(dalloc(&c->JA, 2 * (size_t)C * c->no));
    chk(dalloc(&c->JB, 6 * (size_t)c->no));
    chk(dalloc(&c->W, 3 * (size_t)C * c->no));
  }
  chk(dalloc(&c->U, (size_t)C * C * c->nc));
  chk(dalloc(&c->V, 6 * (size_t)c->np));
  chk(dalloc(&c->g, (size_t)c->ntot));
  chk(dalloc(&c->uvdiag, (size_t)c->ntot));
  chk(dalloc(&c->Vinv, 6 * (size_t)c->np));
  chk(dalloc(&c->z, 3 * (size_t)c->np));
  chk(dalloc(&c->S, (size_t)c->n_pad * c->n_pad));
  chk(dalloc(&c->ea, (size_t)c->n_pad));
  chk(dalloc(&c->xsol, (size_t)c->n_pad));
  chk(dalloc(&c->dp, (size_t)c->ntot));
  chk(dalloc(&c->eb, 3 * (size_t)c->np));
  chk(dalloc(&c->vecA, (size_t)c->ntot));
  chk(dalloc(&c->vecB, (size_t)c->ntot));
  chk(dalloc(&c->vecC, (size_t)c->ntot));
  chk(dalloc(&c->vecD, (size_t)c->ntot));
  chk(dalloc(&c->mc_work, 3 * (size_t)c->n_pad));
  chk(dalloc(&c->Lf, (size_t)c->n_pad * c->n_pad));
  chk(dalloc(&c->bw, (size_t)c->n_pad));
  chk(dalloc(&c->yv, (size_t)c->n_pad));
  chk(dalloc(&c->bar, (size_t)3 + (size_t)(c->n_pad / kNB)));  // barrier counter, back-substitution flags, epoch
  chk(dalloc(&c->chol_fail, (size_t)1));
  c->red_part_cap = 8 * (long)std::max(ceil_div(c->no, 256), ceil_div(c->ntot, 256)) + 64;
  chk(dalloc(&c->red_part, (size_t)c->red_part_cap));
  chk(dalloc(&c->d_scal, (size_t)SC_COUNT));
  if (rc == PSBA_OK && hipHostMalloc((void**)&c->h_scal, sizeof(double) * SC_COUNT, hipHostMallocDefault) != hipSuccess)
    rc = fail(PSBA_ERR_NOMEM, "pinned host allocation failed");
  if (rc == PSBA_OK) {
    hipError_t e = hipSuccess;
    auto z = [&](hipError_t x) {
      if (e == hipSuccess) e = x;
    };
    z(hipMemsetAsync(c->fix, 0, (size_t)C * c->nc, c->stream));
    z(hipMemsetAsync(c->S, 0, sizeof(double) * (size_t)c->n_pad * c->n_pad, c->stream));
    z(hipMemsetAsync(c->ea, 0, sizeof(double) * (size_t)c->n_pad, c->stream));
    z(hipMemsetAsync(c->dp, 0, sizeof(double) * (size_t)c->ntot, c->stream));
    z(hipMemsetAsync(c->d_scal, 0, sizeof(double) * SC_COUNT, c->stream));
    z(hipMemsetAsync(c->chol_fail, 0, sizeof(int), c->stream));
    z(hipMemsetAsync(c->bar, 0, sizeof(unsigned long long) * (3 + (size_t)(c->n_pad / kNB)), c->stream));
    z(hipMemsetAsync(c->Lf, 0, sizeof(double) * (size_t)c->n_pad * c->n_pad, c->stream));
    if (c->n_pad > c->n)
      hipLaunchKernelGGL(k_pad_identity, dim3(1), dim3(64), 0, c->stream, c->S, c->n_pad, c->n, c->n_pad);
    z(hipStreamSynchronize(c->stream));
    if (e != hipSuccess) rc = fail(PSBA_ERR_HIP, "initialisation: %s", hipGetErrorString(e));
  }
  c->h_fix.assign((size_t)C * c->nc, 0);
  if (rc != PSBA_OK) {
    psba_destroy(c);
    return rc;
  }
  *out = c;
  return PSBA_OK;
}

int psba_destroy(psba_ctx* c) {
  if (!c) return PSBA_OK;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  free_index(c);
  double* bufs[] = {c->K, c->q0, c->meas, c->cams, c->pts, c->e, c->JA, c->JB, c->W, c->U, c->V, c->g,
                    c->uvdiag, c->Vinv, c->z, c->S, c->ea, c->xsol, c->dp, c->eb, c->vecA, c->vecB,
                    c->vecC, c->vecD, c->mc_work, c->red_part, c->d_scal, c->Saux, c->snap,
                    c->Lf, c->bw, c->yv};
  for (double* p : bufs)
    if (p) (void)hipFree(p);
  dfree(c->fix);
  dfree(c->e32);
  dfree(c->JA32);
  dfree(c->JB32);
  dfree(c->W32);
  dfree(c->ref_obs);
  dfree(c->ref_part);
  dfree(c->bar);
  dfree(c->chol_fail);
  if (c->lm) (void)hipFree(c->lm);
  if (c->gate) (void)hipFree(c->gate);
  if (c->lm_trace) (void)hipFree(c->lm_trace);
  if (c->h_lm) (void)hipHostFree(c->h_lm);
  for (auto e : c->ev_lm)
    if (e) (void)hipEventDestroy(e);
  if (c->h_scal) (void)hipHostFree(c->h_scal);
  for (auto& pr : c->probe_pending) {
    if (pr.first) (void)hipEventDestroy(pr.first);
    if (pr.second) (void)hipEventDestroy(pr.second);
  }
  for (auto e : c->probe_pool) (void)hipEventDestroy(e);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return PSBA_OK;
}

int psba_upload(psba_ctx* c, const double* K, const double* impts, const double* initrot, const double* cams,
                const double* pts) {
  if (!c || !K || !impts || !initrot || !cams || !pts) return fail(PSBA_ERR_ARG, "null argument");
  PSBA_HIP_TRY(hipSetDevice(c->device));
  PSBA_HIP_TRY(hipMemcpyAsync(c->K, K, sizeof(double) * 5 * c->nc, hipMemcpyHostToDevice, c->stream));
  PSBA_HIP_TRY(hipMemcpyAsync(c->meas, impts, sizeof(double) * 2 * c->no, hipMemcpyHostToDevice, c->stream));
  PSBA_HIP_TRY(hipMemcpyAsync(c->q0, initrot, sizeof(double) * 4 * c->nc, hipMemcpyHostToDevice, c->stream));
  PSBA_HIP_TRY(hipMemcpyAsync(c->cams, cams, sizeof(double) * c->n, hipMemcpyHostToDevice, c->stream));
  PSBA_HIP_TRY(hipMemcpyAsync(c->pts, pts, sizeof(double) * 3 * c->np, hipMemcpyHostToDevice, c->stream));
  PSBA_HIP_TRY(hipStreamSynchronize(c->stream));
  c->uploaded = true;
  return PSBA_OK;
}

int psba_set_loss(psba_ctx* c, int kind, double scale) {
  if (!c) return fail(PSBA_ERR_ARG, "null context");
  if (kind != PSBA_LOSS_NONE && kind != PSBA_LOSS_HUBER && kind != PSBA_LOSS_CAUCHY)
    return fail(PSBA_ERR_ARG, "unknown loss %d", kind);
  if (kind != PSBA_LOSS_NONE && !(scale > 0 && std::isfinite(scale)))
    return fail(PSBA_ERR_ARG, "loss scale must be positive and finite");
  c->loss = kind;
  c->loss_scale = kind == PSBA_LOSS_NONE ? 1.0 : scale;
  return PSBA_OK;
}

int psba_set_fixed(psba_ctx* c, const unsigned char* mask) {
  if (!c) return fail(PSBA_ERR_ARG, "null context");
  PSBA_HIP_TRY(hipSetDevice(c->device));
  size_t cnt = (size_t)c->cnp * c->nc;
  if (mask)
    for (size_t q = 0; q < cnt; ++q) c->h_fix[q] = mask[q] ? 1 : 0;
  else
    std::fill(c->h_fix.begin(), c->h_fix.end(), 0);
  PSBA_HIP_TRY(hipMemcpyAsync(c->fix, c->h_fix.data(), cnt, hipMemcpyHostToDevice, c->stream));
  PSBA_HIP_TRY(hipStreamSynchronize(c->stream));
  return PSBA_OK;
}

int psba_read_params(psba_ctx* c, double* cams, double* pts) {
  if (!c) return fail(PSBA_ERR_ARG, "null context");
  PSBA_HIP_TRY(hipSetDevice(c->device));
  if (cams) PSBA_HIP_TRY(hipMemcpyAsync(cams, c->cams, sizeof(double) * c->n, hipMemcpyDeviceToHost, c->stream));
  if (pts) PSBA_HIP_TRY(hipMemcpyAsync(pts, c->pts, sizeof(double) * 3 * c->np, hipMemcpyDeviceToHost, c->stream));
  PSBA_HIP_TRY(hipStreamSynchronize(c->stream));
  return PSBA_OK;
}

int psba_write_params(psba_ctx* c, const double* cams, const double* pts) {
  if (!c) return fail(PSBA_ERR_ARG, "null context");
  PSBA_HIP_TRY(hipSetDevice(c->device));
  if (cams) PSBA_HIP_TRY(hipMemcpyAsync(c->cams, cams, sizeof(double) * c->n, hipMemcpyHostToDevice, c->stream));
  if (pts) PSBA_HIP_TRY(hipMemcpyAsync(c->pts, pts, sizeof(double) * 3 * c->np, hipMemcpyHostToDevice, c->stream));
  PSBA_HIP_TRY(hipStreamSynchronize(c->stream));
  return PSBA_OK;
}

int psba_snapshot_params(psba_ctx* c) {
  if (!c) return fail(PSBA_ERR_ARG, "null context");
  PSBA_HIP_TRY(hipSetDevice(c->device));
  if (!c->snap) PSBA_TRY(dalloc(&c->snap, (size_t)c->ntot));
  PSBA_HIP_TRY(hipMemcpyAsync(c->snap, c->cams, sizeof(double) * c->n, hipMemcpyDeviceToDevice, c->stream));
  PSBA_HIP_TRY(hipMemcpyAsync(c->snap + c->n, c->pts, sizeof(double) * 3 * c->np, hipMemcpyDeviceToDevice, c->stream));
  PSBA_HIP_TRY(hipStreamSynchronize(c->stream));
  return PSBA_OK;
}

int psba_restore_params(psba_ctx* c) {
  if (!c || !c->snap) return fail(PSBA_ERR_STATE, "no parameter snapshot");
  PSBA_HIP_TRY(hipSetDevice(c->device));
  PSBA_HIP_TRY(hipMemcpyAsync(c->cams, c->snap, sizeof(double) * c->n, hipMemcpyDeviceToDevice, c->stream));
  PSBA_HIP_TRY(hipMemcpyAsync(c->pts, c->snap + c->n, sizeof(double) * 3 * c->np, hipMemcpyDeviceToDevice, c->stream));
  return PSBA_OK;
}

// PSBA/misc.cpp:178-218, bit-exact restatement (shared with the drop-in layer)
int psba_generate_idxs(int nCams, int n3Dpts, int n2Dprojs, const char* vmask, int* comm3DIdx, int* comm3DIdxCnt,
                       int* iidx, int* jidx, int* blk_idx) {
  if (nCams <= 0 || n3Dpts <= 0 || !vmask || !comm3DIdxCnt || !iidx || !jidx || !blk_idx)
    return fail(PSBA_ERR_ARG, "generate_idxs: bad argument");
  const long rowsize = (long)nCams * n3Dpts;
  std::memset(comm3DIdxCnt, 0, sizeof(int) * (size_t)nCams * nCams);
  int k = 0;
  for (int i = 0; i < n3Dpts; ++i) {
    const char* vm = vmask + (long)i * nCams;
    for (int j = 0; j < nCams; ++j) {
      if (!vm[j]) {
        blk_idx[(long)i * nCams + j] = -1;
        continue;
      }
      if (k >= n2Dprojs) return fail(PSBA_ERR_ARG, "generate_idxs: vmask has more than n2Dprojs entries");
      iidx[k] = i;
      jidx[k] = j;
      blk_idx[(long)i * nCams + j] = k;
      for (int l = 0; l <= j; ++l) {
        if (!vm[l]) continue;
        int t = comm3DIdxCnt[j * nCams + l]++;
        if (comm3DIdx) comm3DIdx[j * rowsize + (long)l * n3Dpts + t] = i;
        if (l != j) {
          comm3DIdxCnt[l * nCams + j]++;
          if (comm3DIdx) comm3DIdx[l * rowsize + (long)j * n3Dpts + t] = i;
        }
      }
      ++k;
    }
  }
  if (k != n2Dprojs) return fail(PSBA_ERR_ARG, "generate_idxs: vmask has %d entries, expected %d", k, n2Dprojs);
  return PSBA_OK;
}

}  // extern "C"

namespace psba {

int check_obs_order(int nc, int np, int no, const int* iidx, const int* jidx) {
  for (int k = 0; k < no; ++k) {
    if (iidx[k] < 0 || iidx[k] >= np || jidx[k] < 0 || jidx[k] >= nc)
      return fail(PSBA_ERR_ARG, "observation %d has index out of range", k);
    if (k > 0 && (iidx[k] < iidx[k - 1] || (iidx[k] == iidx[k - 1] && jidx[k] <= jidx[k - 1])))
      return fail(PSBA_ERR_ARG, "observations must be point-major and camera-ascending (obs %d)", k);
  }
  return PSBA_OK;
}

int alloc_index(psba_ctx* c, long nitems, int npairs, int n_cchunk, int n_schunk) {
  free_index(c);
  const int nc = c->nc, np = c->np, no = c->no, C = c->cnp;
  const int nlb = nc * (nc + 1) / 2;
  c->n_cchunk = n_cchunk;
  c->npairs = npairs;
  c->n_items = nitems;
  c->n_schunk = n_schunk;
  c->n_lblocks = nlb;
  const int NE = C * (C + 1) / 2 + C;
  PSBA_TRY(dalloc(&c->obs_cam, no));
  PSBA_TRY(dalloc(&c->obs_pt, no));
  PSBA_TRY(dalloc(&c->pt_ptr, np + 1));
  PSBA_TRY(dalloc(&c->cam_ptr, nc + 1));
  PSBA_TRY(dalloc(&c->cam_obs, no));
  PSBA_TRY(dalloc(&c->cchunk_begin, n_cchunk));
  PSBA_TRY(dalloc(&c->cchunk_end, n_cchunk));
  PSBA_TRY(dalloc(&c->cam_cchunk_ptr, nc + 1));
  PSBA_TRY(dalloc(&c->pair_items, (size_t)nitems));
  PSBA_TRY(dalloc(&c->pair_k, npairs));
  PSBA_TRY(dalloc(&c->pair_l, npairs));
  PSBA_TRY(dalloc(&c->schunk_begin, n_schunk));
  PSBA_TRY(dalloc(&c->schunk_end, n_schunk));
  PSBA_TRY(dalloc(&c->pair_schunk_ptr, npairs + 1));
  PSBA_TRY(dalloc(&c->schunk_order, n_schunk));
  PSBA_TRY(dalloc(&c->slot_items, (size_t)n_schunk * kPairChunk));
  PSBA_TRY(dalloc(&c->cchunk_rec, (size_t)n_cchunk * cam_chunk_size(C)));
  PSBA_TRY(dalloc(&c->pt_order, np));
  PSBA_TRY(dalloc(&c->lblock_pair, nlb));
  if (c->mixed()) {
    PSBA_TRY(dalloc(&c->ucpart32, (size_t)n_cchunk * NE));
    PSBA_TRY(dalloc(&c->eapart32, (size_t)n_cchunk * C));
    PSBA_TRY(dalloc(&c->schpart32, (size_t)n_schunk * C * C));
  } else {
    PSBA_TRY(dalloc(&c->ucpart, (size_t)n_cchunk * NE));
    PSBA_TRY(dalloc(&c->eapart, (size_t)n_cchunk * C));
    PSBA_TRY(dalloc(&c->schpart, (size_t)n_schunk * C * C));
  }
  return PSBA_OK;
}

}  // namespace psba

extern "C" {

int psba_set_index(psba_ctx* c, const int* iidx, const int* jidx) {
  if (!c || !iidx || !jidx) return fail(PSBA_ERR_ARG, "null argument");
  PSBA_HIP_TRY(hipSetDevice(c->device));
  const int nc = c->nc, np = c->np, no = c->no, C = c->cnp;
  PSBA_TRY(check_obs_order(nc, np, no, iidx, jidx));
  free_index(c);
  std::vector<int> pt_ptr(np + 1, 0), cam_ptr(nc + 1, 0), cam_obs(no);
  for (int k = 0; k < no; ++k) {
    pt_ptr[iidx[k] + 1]++;
    cam_ptr[jidx[k] + 1]++;
  }
  for (int i = 0; i < np; ++i) pt_ptr[i + 1] += pt_ptr[i];
  for (int j = 0; j < nc; ++j) cam_ptr[j + 1] += cam_ptr[j];
  {
    std::vector<int> fill(nc, 0);
    for (int k = 0; k < no; ++k) cam_obs[cam_ptr[jidx[k]] + fill[jidx[k]]++] = k;
  }
  // camera chunks
  const int CH = cam_chunk_size(C);
  std::vector<int> cb, ce, ccp(nc + 1, 0);
  for (int j = 0; j < nc; ++j) {
    for (int s = cam_ptr[j]; s < cam_ptr[j + 1]; s += CH) {
      cb.push_back(s);
      ce.push_back(std::min(s + CH, cam_ptr[j + 1]));
    }
    ccp[j + 1] = (int)cb.size();
  }
  // camera pairs k >= l and their items in ascending point order
  std::vector<long> cnt((size_t)nc * nc, 0);
  double st2 = 0;
  for (int i = 0; i < np; ++i) {
    int t = pt_ptr[i + 1] - pt_ptr[i];
    st2 += (double)t * t;
    for (int a = pt_ptr[i]; a < pt_ptr[i + 1]; ++a)
      for (int b = pt_ptr[i]; b <= a; ++b) cnt[(size_t)jidx[a] * nc + jidx[b]]++;
  }
  std::vector<int> pid((size_t)nc * nc, -1), pk, pl;
  std::vector<long> pptr(1, 0);
  for (int k = 0; k < nc; ++k)
    for (int l = 0; l <= k; ++l) {
      long q = cnt[(size_t)k * nc + l];
      if (!q) continue;
      pid[(size_t)k * nc + l] = (int)pk.size();
      pk.push_back(k);
      pl.push_back(l);
      pptr.push_back(pptr.back() + q);
    }
  const int npairs = (int)pk.size();
  const long nitems = pptr.back();
  if (nitems > 0x7fffffffL) return fail(PSBA_ERR_UNSUPPORTED, "too many camera-pair items (%ld)", nitems);
  std::vector<int2> items((size_t)nitems);
  {
    std::vector<long> pf(npairs, 0);
    for (int i = 0; i < np; ++i)
      for (int a = pt_ptr[i]; a < pt_ptr[i + 1]; ++a)
        for (int b = pt_ptr[i]; b <= a; ++b) {
          int q = pid[(size_t)jidx[a] * nc + jidx[b]];
          items[(size_t)(pptr[q] + pf[q]++)] = make_int2(a, b);
        }
  }
  std::vector<int> sb, se, psp(npairs + 1, 0);
  for (int q = 0; q < npairs; ++q) {
    for (long s = pptr[q]; s < pptr[q + 1]; s += kPairChunk) {
      sb.push_back((int)s);
      se.push_back((int)std::min(s + (long)kPairChunk, pptr[q + 1]));
    }
    psp[q + 1] = (int)sb.size();
  }
  // launch order: by camera-block tile, then point range, then chunk (stable)
  std::vector<int> sord(sb.size());
  {
    std::vector<unsigned long long> key(sb.size());
    for (size_t q = 0; q < sb.size(); ++q) {
      const int2 it = items[(size_t)sb[q]];
      key[q] = schunk_launch_key(jidx[it.x], jidx[it.y], iidx[it.x]);
      sord[q] = (int)q;
    }
    std::stable_sort(sord.begin(), sord.end(), [&](int x, int y) { return key[x] < key[y]; });
  }
  const int nlb = nc * (nc + 1) / 2;
  std::vector<int> lbp(nlb);
  for (int k = 0; k < nc; ++k)
    for (int l = 0; l <= k; ++l) lbp[k * (k + 1) / 2 + l] = pid[(size_t)k * nc + l];

  PSBA_TRY(alloc_index(c, nitems, npairs, (int)cb.size(), (int)sb.size()));
  c->sum_t2 = st2;
  auto up = [&](void* dst, const void* src, size_t bytes) -> int {
    if (bytes == 0) return PSBA_OK;
    PSBA_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    return PSBA_OK;
  };
  PSBA_TRY(up(c->obs_cam, jidx, sizeof(int) * no));
  PSBA_TRY(up(c->obs_pt, iidx, sizeof(int) * no));
  PSBA_TRY(up(c->pt_ptr, pt_ptr.data(), sizeof(int) * (np + 1)));
  PSBA_TRY(up(c->cam_ptr, cam_ptr.data(), sizeof(int) * (nc + 1)));
  PSBA_TRY(up(c->cam_obs, cam_obs.data(), sizeof(int) * no));
  PSBA_TRY(up(c->cchunk_begin, cb.data(), sizeof(int) * cb.size()));
  PSBA_TRY(up(c->cchunk_end, ce.data(), sizeof(int) * ce.size()));
  PSBA_TRY(up(c->cam_cchunk_ptr, ccp.data(), sizeof(int) * (nc + 1)));
  PSBA_TRY(up(c->pair_items, items.data(), sizeof(int2) * (size_t)nitems));
  PSBA_TRY(up(c->pair_k, pk.data(), sizeof(int) * npairs));
  PSBA_TRY(up(c->pair_l, pl.data(), sizeof(int) * npairs));
  PSBA_TRY(up(c->schunk_begin, sb.data(), sizeof(int) * sb.size()));
  PSBA_TRY(up(c->schunk_end, se.data(), sizeof(int) * se.size()));
  PSBA_TRY(up(c->pair_schunk_ptr, psp.data(), sizeof(int) * (npairs + 1)));
  PSBA_TRY(up(c->lblock_pair, lbp.data(), sizeof(int) * nlb));
  PSBA_TRY(up(c->schunk_order, sord.data(), sizeof(int) * sord.size()));
  PSBA_TRY(build_point_order(c));
  PSBA_TRY(build_schur_slots(c));
  PSBA_HIP_TRY(hipStreamSynchronize(c->stream));
  c->indexed = true;
  return PSBA_OK;
}

static int ready(psba_ctx* c) {
  if (!c) return fail(PSBA_ERR_ARG, "null context");
  if (!c->uploaded || !c->indexed) return fail(PSBA_ERR_STATE, "upload data and set the index first");
  PSBA_HIP_TRY(hipSetDevice(c->device));
  return PSBA_OK;
}

// the dense stages also run on a reduced system loaded with psba_stage_load_S
static int ready_dense(psba_ctx* c) {
  if (c && c->s_loaded) {
    PSBA_HIP_TRY(hipSetDevice(c->device));
    return PSBA_OK;
  }
  return ready(c);
}

int psba_stage_residual(psba_ctx* c, int use_trial, double* cost) {
  PSBA_TRY(ready(c));
  PSBA_TRY(launch_residual(c, use_trial != 0));
  PSBA_HIP_TRY(hipMemcpyAsync(c->h_scal, c->d_scal, sizeof(double) * SC_COUNT, hipMemcpyDeviceToHost, c->stream));
  PSBA_HIP_TRY(hipStreamSynchronize(c->stream));
  if (cost) *cost = c->h_scal[SC_COST];
  return PSBA_OK;
}

int psba_stage_jacobian_normal(psba_ctx* c, double coeff, double coeff_g) {
  PSBA_TRY(ready(c));
  PSBA_TRY(launch_jacobian_normal(c, coeff, coeff_g));
  PSBA_HIP_TRY(hipStreamSynchronize(c->stream));
  return PSBA_OK;
}

int psba_stage_schur(psba_ctx* c, double mu) {
  PSBA_TRY(ready(c));
  PSBA_TRY(launch_vinv(c, mu));
  PSBA_TRY(launch_schur(c, mu));
  PSBA_HIP_TRY(hipStreamSynchronize(c->stream));
  return PSBA_OK;
}

int psba_stage_vinv(psba_ctx* c, const double* V9, double mu, double* out9) {
  if (!c || !V9 || !out9) return fail(PSBA_ERR_ARG, "null argument");
  PSBA_HIP_TRY(hipSetDevice(c->device));
  // symmetric 3x3 blocks in the device's packed order (xx xy xz yy yz zz)
  static const int kPack[6] = {0, 1, 2, 4, 5, 8};
  std::vector<double> packed((size_t)6 * c->np), inv((size_t)6 * c->np);
  for (int i = 0; i < c->np; ++i)
    for (int k = 0; k < 6; ++k) packed[(size_t)6 * i + k] = V9[(size_t)9 * i + kPack[k]];
  PSBA_HIP_TRY(hipMemcpyAsync(c->V, packed.data(), sizeof(double) * packed.size(), hipMemcpyHostToDevice, c->stream));
  PSBA_HIP_TRY(hipMemsetAsync(c->g + c->n, 0, sizeof(double) * 3 * (size_t)c->np, c->stream));
  PSBA_TRY(launch_vinv(c, mu));
  PSBA_HIP_TRY(hipMemcpyAsync(inv.data(), c->Vinv, sizeof(double) * inv.size(), hipMemcpyDeviceToHost, c->stream));
  PSBA_HIP_TRY(hipStreamSynchronize(c->stream));
  static const int kFull[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};
  for (int i = 0; i < c->np; ++i)
    for (int k = 0; k < 9; ++k) out9[(size_t)9 * i + k] = inv[(size_t)6 * i + kFull[k]];
  return PSBA_OK;
}

int psba_stage_factor_solve(psba_ctx* c, int* ok) {
  PSBA_TRY(ready_dense(c));
  PSBA_TRY(launch_factor_solve(c));
  PSBA_HIP_TRY(hipMemcpyAsync(c->h_scal, c->d_scal, sizeof(double) * SC_COUNT, hipMemcpyDeviceToHost, c->stream));
  PSBA_HIP_TRY(hipStreamSynchronize(c->stream));
  if (c->h_scal[SC_FACTOR_OK] < 0.0)
    return fail(PSBA_ERR_HIP, "dense solve: a grid-barrier wait timed out (workgroups not all resident)");
  if (ok) *ok = c->h_scal[SC_FACTOR_OK] != 0.0;
  return PSBA_OK;
}

int psba_stage_backsub(psba_ctx* c) {
  PSBA_TRY(ready(c));
  PSBA_TRY(launch_backsub(c));
  PSBA_HIP_TRY(hipStreamSynchronize(c->stream));
  return PSBA_OK;
}

int psba_stage_refine(psba_ctx* c, double mu, int sweeps) {
  PSBA_TRY(ready(c));
  if (!c->mixed()) return fail(PSBA_ERR_UNSUPPORTED, "refinement applies to mixed-precision contexts");
  if (sweeps < 0) return fail(PSBA_ERR_ARG, "sweeps < 0");
  for (int k = 0; k < sweeps; ++k) {
    PSBA_TRY(launch_refine(c, mu, 1));
    PSBA_TRY(launch_axpby(c, c->dp, 1.0, c->dp, 1.0, c->vecB, c->ntot));
  }
  PSBA_HIP_TRY(hipStreamSynchronize(c->stream));
  return PSBA_OK;
}

int psba_stage_modchol(psba_ctx* c, double* lambda) {
  PSBA_TRY(ready_dense(c));
  double* A = nullptr;
  PSBA_TRY(dalloc(&A, (size_t)c->n_pad * c->n_pad));
  int rc = launch_copy_S(c, A, c->S);
  if (rc == PSBA_OK) rc = launch_modchol(c, A);
  if (rc == PSBA_OK && hipMemcpyAsync(c->h_scal, c->d_scal, sizeof(double) * SC_COUNT, hipMemcpyDeviceToHost,
                                      c->stream) != hipSuccess)
    rc = fail(PSBA_ERR_HIP, "read-back");
  if (hipStreamSynchronize(c->stream) != hipSuccess && rc == PSBA_OK) rc = fail(PSBA_ERR_HIP, "sync");
  (void)hipFree(A);
  if (rc == PSBA_OK && lambda) *lambda = c->h_scal[SC_LAMBDA];
  return rc;
}

int psba_stage_load_S(psba_ctx* c, const double* S, const double* b) {
  if (!c || !S) return fail(PSBA_ERR_ARG, "null argument");
  PSBA_HIP_TRY(hipSetDevice(c->device));
  PSBA_HIP_TRY(hipMemcpy2DAsync(c->S, sizeof(double) * c->n_pad, S, sizeof(double) * c->n, sizeof(double) * c->n,
                                c->n, hipMemcpyHostToDevice, c->stream));
  if (c->n_pad > c->n) {
    // padding rows: identity (their first n columns are zero)
    std::vector<double> pad((size_t)(c->n_pad - c->n) * c->n_pad, 0.0);
    for (int r = c->n; r < c->n_pad; ++r) pad[(size_t)(r - c->n) * c->n_pad + r] = 1.0;
    PSBA_HIP_TRY(hipMemcpyAsync(c->S + (size_t)c->n * c->n_pad, pad.data(), sizeof(double) * pad.size(),
                                hipMemcpyHostToDevice, c->stream));
  }
  if (b) {
    PSBA_HIP_TRY(hipMemsetAsync(c->ea, 0, sizeof(double) * c->n_pad, c->stream));
    PSBA_HIP_TRY(hipMemcpyAsync(c->ea, b, sizeof(double) * c->n, hipMemcpyHostToDevice, c->stream));
  }
  PSBA_HIP_TRY(hipStreamSynchronize(c->stream));
  c->s_loaded = true;
  return PSBA_OK;
}

long psba_buffer_size(const psba_ctx* c, int which) {
  if (!c) return -1;
  const long C = c->cnp;
  switch (which) {
    case PSBA_BUF_EX: return 2L * c->no;
    case PSBA_BUF_JA: return 2L * C * c->no;
    case PSBA_BUF_JB: return 6L * c->no;
    case PSBA_BUF_U: return C * C * c->nc;
    case PSBA_BUF_V: return 9L * c->np;
    case PSBA_BUF_W: return 3L * C * c->no;
    case PSBA_BUF_G: return c->ntot;
    case PSBA_BUF_VINV: return 9L * c->np;
    case PSBA_BUF_S: return (long)c->n * c->n;
    case PSBA_BUF_EA: return c->n;
    case PSBA_BUF_DP: return c->ntot;
    case PSBA_BUF_EB: return 3L * c->np;
    case PSBA_BUF_L: return (long)c->n * c->n;
  }
  return -1;
}

int psba_read_buffer(psba_ctx* c, int which, double* out, long n) {
  if (!c || !out) return fail(PSBA_ERR_ARG, "null argument");
  long need = psba_buffer_size(c, which);
  if (need < 0) return fail(PSBA_ERR_ARG, "unknown buffer %d", which);
  if (n < need) return fail(PSBA_ERR_ARG, "buffer %d needs %ld doubles", which, need);
  PSBA_HIP_TRY(hipSetDevice(c->device));
  auto d2h = [&](double* dst, const double* src, size_t cnt) -> int {
    PSBA_HIP_TRY(hipMemcpyAsync(dst, src, sizeof(double) * cnt, hipMemcpyDeviceToHost, c->stream));
    PSBA_HIP_TRY(hipStreamSynchronize(c->stream));
    return PSBA_OK;
  };
  auto f2h = [&](double* dst, const float* src, size_t cnt) -> int {
    std::vector<float> h(cnt);
    PSBA_HIP_TRY(hipMemcpyAsync(h.data(), src, sizeof(float) * cnt, hipMemcpyDeviceToHost, c->stream));
    PSBA_HIP_TRY(hipStreamSynchronize(c->stream));
    for (size_t q = 0; q < cnt; ++q) dst[q] = h[q];
    return PSBA_OK;
  };
  if (c->mixed()) {
    switch (which) {
      case PSBA_BUF_EX: return f2h(out, c->e32, need);
      case PSBA_BUF_JA: return f2h(out, c->JA32, need);
      case PSBA_BUF_JB: return f2h(out, c->JB32, need);
      case PSBA_BUF_W: return f2h(out, c->W32, need);
    }
  }
  switch (which) {
    case PSBA_BUF_EX: return d2h(out, c->e, need);
    case PSBA_BUF_JA: return d2h(out, c->JA, need);
    case PSBA_BUF_JB: return d2h(out, c->JB, need);
    case PSBA_BUF_U: return d2h(out, c->U, need);
    case PSBA_BUF_W: return d2h(out, c->W, need);
    case PSBA_BUF_G: return d2h(out, c->g, need);
    case PSBA_BUF_EA: return d2h(out, c->ea, need);
    case PSBA_BUF_DP: return d2h(out, c->dp, need);
    case PSBA_BUF_EB: return d2h(out, c->eb, need);
    case PSBA_BUF_V:
    case PSBA_BUF_VINV: {
      std::vector<double> h6(6 * (size_t)c->np);
      PSBA_TRY(d2h(h6.data(), which == PSBA_BUF_V ? c->V : c->Vinv, h6.size()));
      for (int i = 0; i < c->np; ++i) {
        const double* s = &h6[6 * (size_t)i];
        double* d = out + 9 * (size_t)i;
        d[0] = s[0]; d[1] = s[1]; d[2] = s[2];
        d[3] = s[1]; d[4] = s[3]; d[5] = s[4];
        d[6] = s[2]; d[7] = s[4]; d[8] = s[5];
      }
      return PSBA_OK;
    }
    case PSBA_BUF_S:
    case PSBA_BUF_L: {
      std::vector<double> h((size_t)c->n_pad * c->n_pad);
      PSBA_TRY(d2h(h.data(), (which == PSBA_BUF_L && c->L_cur) ? c->L_cur : c->S, h.size()));
      for (int r = 0; r < c->n; ++r)
        for (int q = 0; q < c->n; ++q) out[(size_t)r * c->n + q] = h[(size_t)r * c->n_pad + q];
      return PSBA_OK;
    }
  }
  return fail(PSBA_ERR_ARG, "unknown buffer %d", which);
}

static const int* index_ptr(const psba_ctx* c, int which, long* n) {
  switch (which) {
    case PSBA_IDX_OBS_PT: *n = c->no; return c->obs_pt;
    case PSBA_IDX_OBS_CAM: *n = c->no; return c->obs_cam;
    case PSBA_IDX_PT_PTR: *n = c->np + 1; return c->pt_ptr;
    case PSBA_IDX_CAM_PTR: *n = c->nc + 1; return c->cam_ptr;
    case PSBA_IDX_CAM_OBS: *n = c->no; return c->cam_obs;
    case PSBA_IDX_CCHUNK_BEGIN: *n = c->n_cchunk; return c->cchunk_begin;
    case PSBA_IDX_CCHUNK_END: *n = c->n_cchunk; return c->cchunk_end;
    case PSBA_IDX_CAM_CCHUNK_PTR: *n = c->nc + 1; return c->cam_cchunk_ptr;
    case PSBA_IDX_PAIR_ITEMS: *n = 2 * c->n_items; return (const int*)c->pair_items;
    case PSBA_IDX_PAIR_K: *n = c->npairs; return c->pair_k;
    case PSBA_IDX_PAIR_L: *n = c->npairs; return c->pair_l;
    case PSBA_IDX_SCHUNK_BEGIN: *n = c->n_schunk; return c->schunk_begin;
    case PSBA_IDX_SCHUNK_END: *n = c->n_schunk; return c->schunk_end;
    case PSBA_IDX_PAIR_SCHUNK_PTR: *n = c->npairs + 1; return c->pair_schunk_ptr;
    case PSBA_IDX_LBLOCK_PAIR: *n = c->n_lblocks; return c->lblock_pair;
    case PSBA_IDX_SCHUNK_ORDER: *n = c->n_schunk; return c->schunk_order;
    case PSBA_IDX_PT_ORDER: *n = c->np; return c->pt_order;
    default: *n = -1; return nullptr;
  }
}

long psba_index_size(const psba_ctx* c, int which) {
  if (!c || !c->indexed) return -1;
  long n = -1;
  (void)index_ptr(c, which, &n);
  return n;
}

int psba_read_index(psba_ctx* c, int which, int* out, long n) {
  if (!c || !out) return fail(PSBA_ERR_ARG, "null argument");
  if (!c->indexed) return fail(PSBA_ERR_STATE, "no index");
  long size = -1;
  const int* src = index_ptr(c, which, &size);
  if (size < 0) return fail(PSBA_ERR_ARG, "unknown index array %d", which);
  if (n < size) return fail(PSBA_ERR_ARG, "output holds %ld ints, %ld needed", n, size);
  PSBA_HIP_TRY(hipSetDevice(c->device));
  if (size > 0) PSBA_HIP_TRY(hipMemcpy(out, src, sizeof(int) * size, hipMemcpyDeviceToHost));
  return PSBA_OK;
}

int psba_get_info(psba_ctx* c, psba_info* o) {
  if (!c || !o) return fail(PSBA_ERR_ARG, "null argument");
  o->cnp = c->cnp;
  o->n_cams = c->nc;
  o->n_pts = c->np;
  o->n_obs = c->no;
  o->n = c->n;
  o->n_pad = c->n_pad;
  o->n_pairs = c->npairs;
  o->n_pair_items = c->n_items;
  o->sum_t2 = c->sum_t2;
  return PSBA_OK;
}

int psba_probe(psba_ctx* c, int which) {
  if (!c) return fail(PSBA_ERR_ARG, "null context");
  c->probe_id = which;
  return PSBA_OK;
}

int psba_probe_read(psba_ctx* c, double* total_ms, int* count) {
  if (!c) return fail(PSBA_ERR_ARG, "null context");
  PSBA_HIP_TRY(hipSetDevice(c->device));
  PSBA_HIP_TRY(hipStreamSynchronize(c->stream));
  double tot = 0;
  int n = 0;
  for (auto& pr : c->probe_pending) {
    if (pr.first && pr.second) {
      float ms = 0;
      if (hipEventElapsedTime(&ms, pr.first, pr.second) == hipSuccess) {
        tot += ms;
        ++n;
      }
    }
    if (pr.first) c->probe_pool.push_back(pr.first);
    if (pr.second) c->probe_pool.push_back(pr.second);
  }
  c->probe_pending.clear();
  if (total_ms) *total_ms = tot;
  if (count) *count = n;
  return PSBA_OK;
}

int psba_enable_timing(psba_ctx* c, int on) {
  if (!c) return fail(PSBA_ERR_ARG, "null context");
  c->timing_on = on != 0;
  if (c->timing_on && !c->ev0) {
    PSBA_HIP_TRY(hipSetDevice(c->device));
    PSBA_HIP_TRY(hipEventCreate(&c->ev0));
    PSBA_HIP_TRY(hipEventCreate(&c->ev1));
  }
  return PSBA_OK;
}

int psba_get_timing(psba_ctx* c, psba_timing* t) {
  if (!c || !t) return fail(PSBA_ERR_ARG, "null argument");
  t->jacobian_normal = c->timing.jac;
  t->schur = c->timing.schur;
  t->factor = c->timing.factor;
  t->solve = c->timing.solve;
  t->backsub = c->timing.backsub;
  t->residual = c->timing.resid;
  t->other = c->timing.other;
  t->trials = c->timing.trials;
  t->jacobians = c->timing.jacobians;
  return PSBA_OK;
}

}  // extern "C"
