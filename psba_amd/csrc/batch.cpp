// Batches of independent problems (include/psba_batch.h): host staging, the
// concatenated index, the arena, graph-captured LM rounds and the polling
// driver.  No counterpart in the reference (one problem per process,
// PSBA/main.cpp:193-208); each problem follows psba_levmar's rules.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "batch.hpp"
#include "context.hpp"

using namespace psba;

namespace {

int bfail(int code, const char* fmt, ...) {
  char buf[400];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  set_last_error(__FILE__, __LINE__, buf);
  return code;
}

long pad64(long n) { return (n + kNB - 1) / kNB * kNB; }

// index of one problem in its own numbering (the same rules as psba_set_index)
int build_local(const psba_batch::Staged& s, int C, LocalIndex& L) {
  const int nc = s.nc, np = s.np, no = s.no;
  const int* iidx = s.iidx.data();
  const int* jidx = s.jidx.data();
  L.pt_ptr.assign(np + 1, 0);
  std::vector<int> cam_ptr(nc + 1, 0);
  for (int k = 0; k < no; ++k) {
    L.pt_ptr[iidx[k] + 1]++;
    cam_ptr[jidx[k] + 1]++;
  }
  for (int i = 0; i < np; ++i) L.pt_ptr[i + 1] += L.pt_ptr[i];
  for (int j = 0; j < nc; ++j) cam_ptr[j + 1] += cam_ptr[j];
  L.cam_obs.resize(no);
  {
    std::vector<int> fill(nc, 0);
    for (int k = 0; k < no; ++k) L.cam_obs[cam_ptr[jidx[k]] + fill[jidx[k]]++] = k;
  }
  const int CH = cam_chunk_size(C);
  L.ccp.assign(nc + 1, 0);
  for (int j = 0; j < nc; ++j) {
    for (int a = cam_ptr[j]; a < cam_ptr[j + 1]; a += CH) {
      L.cb.push_back(a);
      L.ce.push_back(std::min(a + CH, cam_ptr[j + 1]));
    }
    L.ccp[j + 1] = (int)L.cb.size();
  }
  std::vector<long> cnt((size_t)nc * nc, 0);
  for (int i = 0; i < np; ++i)
    for (int a = L.pt_ptr[i]; a < L.pt_ptr[i + 1]; ++a)
      for (int b = L.pt_ptr[i]; b <= a; ++b) cnt[(size_t)jidx[a] * nc + jidx[b]]++;
  std::vector<int> pid((size_t)nc * nc, -1);
  std::vector<long> pptr(1, 0);
  for (int k = 0; k < nc; ++k)
    for (int l = 0; l <= k; ++l) {
      const long q = cnt[(size_t)k * nc + l];
      if (!q) continue;
      pid[(size_t)k * nc + l] = L.npairs++;
      pptr.push_back(pptr.back() + q);
    }
  const long nitems = pptr.back();
  if (nitems > 0x3fffffffL) return bfail(PSBA_ERR_UNSUPPORTED, "too many camera-pair items (%ld)", nitems);
  L.items.resize((size_t)nitems);
  {
    std::vector<long> pf(L.npairs, 0);
    for (int i = 0; i < np; ++i)
      for (int a = L.pt_ptr[i]; a < L.pt_ptr[i + 1]; ++a)
        for (int b = L.pt_ptr[i]; b <= a; ++b) {
          const int q = pid[(size_t)jidx[a] * nc + jidx[b]];
          L.items[(size_t)(pptr[q] + pf[q]++)] = make_int2(a, b);
        }
  }
  L.psp.assign(L.npairs + 1, 0);
  for (int q = 0; q < L.npairs; ++q) {
    for (long a = pptr[q]; a < pptr[q + 1]; a += kPairChunk) {
      L.sb.push_back((int)a);
      L.se.push_back((int)std::min(a + (long)kPairChunk, pptr[q + 1]));
    }
    L.psp[q + 1] = (int)L.sb.size();
  }
  L.lbpair.resize((size_t)nc * (nc + 1) / 2);
  for (int k = 0; k < nc; ++k)
    for (int l = 0; l <= k; ++l) L.lbpair[(size_t)k * (k + 1) / 2 + l] = pid[(size_t)k * nc + l];
  return PSBA_OK;
}

// sizes that fix the arena layout
struct Sizes {
  long P = 0, NC = 0, NP = 0, NO = 0, N = 0, items = 0, n_schunk = 0, n_cchunk = 0, nlb = 0, npairs = 0,
       n_rchunk = 0, S_len = 0, V_len = 0, T_max = 0;
};

// arena layout: every view 256-byte aligned; with base == nullptr only the
// total is computed
struct Carver {
  char* base;
  size_t off = 0;
  explicit Carver(char* b) : base(b) {}
  template <typename T>
  void take(T** p, long count) {
    off = (off + 255) & ~(size_t)255;
    if (base) *p = (T*)(base + off);
    off += sizeof(T) * (size_t)std::max(count, 1L);
  }
  void take_raw(void** p, long count, size_t elem) {
    off = (off + 255) & ~(size_t)255;
    if (base) *p = (void*)(base + off);
    off += elem * (size_t)std::max(count, 1L);
  }
};

// Views in this order: the inputs (uploaded from the staging image), then the
// working parameters, work arrays and per-problem state.  *input_end receives
// the size of the input part.
size_t carve(psba_batch* t, const Sizes& z, int C, size_t R, char* base, size_t* input_end = nullptr) {
  Carver cv(base);
  const long NE = (long)C * (C + 1) / 2 + C;
  const long ntot = z.N + 3 * z.NP;
  cv.take(&t->K, 5 * z.NC);
  cv.take(&t->q0, 4 * z.NC);
  cv.take(&t->meas, 2 * z.NO);
  cv.take(&t->cams0, z.N);
  cv.take(&t->pts0, 3 * z.NP);
  cv.take(&t->fix, z.N);
  cv.take(&t->obs_cam, z.NO);
  cv.take(&t->obs_pt, z.NO);
  cv.take(&t->obs_prob, z.NO);
  cv.take(&t->pt_ptr, z.NP + 1);
  cv.take(&t->pt_prob, z.NP);
  cv.take(&t->cam_obs, z.NO);
  cv.take(&t->cam_prob, z.NC);
  cv.take(&t->cchunk_b, z.n_cchunk);
  cv.take(&t->cchunk_e, z.n_cchunk);
  cv.take(&t->ccp, z.NC + 1);
  cv.take(&t->cchunk_prob, z.n_cchunk);
  cv.take(&t->items, z.items);
  cv.take(&t->schunk_b, z.n_schunk);
  cv.take(&t->schunk_e, z.n_schunk);
  cv.take(&t->psp, z.npairs + 1);
  cv.take(&t->schunk_prob, z.n_schunk);
  cv.take(&t->lb_prob, z.nlb);
  cv.take(&t->lb_k, z.nlb);
  cv.take(&t->lb_l, z.nlb);
  cv.take(&t->lb_pair, z.nlb);
  cv.take(&t->rchunk_b, z.n_rchunk);
  cv.take(&t->rchunk_e, z.n_rchunk);
  cv.take(&t->rchunk_prob, z.n_rchunk);
  cv.take(&t->rcp, z.P + 1);
  cv.take(&t->cam_off, z.P + 1);
  cv.take(&t->pt_off, z.P + 1);
  cv.take(&t->obs_off, z.P + 1);
  cv.take(&t->npad, z.P);
  cv.take(&t->vec_off, z.P + 1);
  cv.take(&t->S_off, z.P + 1);
  cv.take(&t->all_on, z.P);
  cv.take(&t->chol_fo, std::max(0L, z.T_max - 1) * (z.P + 1));
  cv.take(&t->chol_bo, z.T_max * (z.P + 1));
  if (input_end) *input_end = cv.off;
  cv.take(&t->cams, z.N);
  cv.take(&t->pts, 3 * z.NP);
  cv.take_raw(&t->e, 2 * z.NO, R);
  cv.take_raw(&t->JA, 2 * C * z.NO, R);
  cv.take_raw(&t->JB, 6 * z.NO, R);
  cv.take_raw(&t->W, 3 * C * z.NO, R);
  cv.take_raw(&t->ucpart, z.n_cchunk * NE, R);
  cv.take_raw(&t->schpart, z.n_schunk * C * C, R);
  cv.take_raw(&t->eapart, z.n_cchunk * C, R);
  cv.take(&t->U, (long)C * C * z.NC);
  cv.take(&t->V, 6 * z.NP);
  cv.take(&t->g, ntot);
  cv.take(&t->uvdiag, ntot);
  cv.take(&t->Vinv, 6 * z.NP);
  cv.take(&t->z, 3 * z.NP);
  cv.take(&t->S, z.S_len);
  cv.take(&t->Lf, z.S_len);
  cv.take(&t->ea, z.N);
  cv.take(&t->bw, z.V_len);
  cv.take(&t->yv, z.V_len);
  cv.take(&t->xs, z.V_len);
  cv.take(&t->dp, ntot);
  cv.take(&t->eb, 3 * z.NP);
  cv.take(&t->rpart, z.n_rchunk);
  cv.take(&t->cost, z.P);
  cv.take(&t->tcost, z.P);
  cv.take(&t->mu, z.P);
  cv.take(&t->p_L2, z.P);
  cv.take(&t->dp2, z.P);
  cv.take(&t->rhoden, z.P);
  cv.take(&t->init_err, z.P);
  cv.take(&t->fok, z.P);
  cv.take(&t->nu, z.P);
  cv.take(&t->itno, z.P);
  cv.take(&t->code, z.P);
  cv.take(&t->active, z.P);
  cv.take(&t->need_jac, z.P);
  cv.take(&t->accept, z.P);
  cv.take(&t->first, z.P);
  cv.take(&t->stat, z.P + 1);
  return cv.off;
}

Sizes problem_sizes(int C, int nc, int np, int no, long pair_items, long npairs) {
  Sizes z;
  const int CH = cam_chunk_size(C);
  z.P = 1;
  z.NC = nc;
  z.NP = np;
  z.NO = no;
  z.N = (long)C * nc;
  z.items = pair_items;
  z.npairs = npairs;
  z.n_schunk = pair_items / kPairChunk + npairs;
  z.n_cchunk = no / CH + nc;
  z.nlb = (long)nc * (nc + 1) / 2;
  z.n_rchunk = (no + kResidChunk - 1) / kResidChunk;
  const long np_ = pad64(z.N);
  z.S_len = np_ * np_;
  z.V_len = np_;
  z.T_max = np_ / kNB;
  return z;
}

// Destroys the captured rounds.  A finished solve may still have rounds
// queued behind its last check (they run as no-ops), so the stream is drained
// first: an executable graph is never destroyed while a launch of it runs.
int free_graph(psba_batch* b) {
  if ((b->graph_exec[0] || b->graph_exec[1]) && b->stream) PSBA_HIP_TRY(hipStreamSynchronize(b->stream));
  for (int k = 0; k < 2; ++k) {
    if (b->graph_exec[k]) (void)hipGraphExecDestroy(b->graph_exec[k]);
    if (b->graph[k]) (void)hipGraphDestroy(b->graph[k]);
    b->graph_exec[k] = nullptr;
    b->graph[k] = nullptr;
  }
  b->graph_iter = -1;
  return PSBA_OK;
}

// one LM round: every problem still running makes one trial; the round ends
// with the per-problem status copied to pinned status buffer `parity`
int enqueue_round(psba_batch* b, int parity) {
  PSBA_TRY(launch_b_jacobian(b));
  PSBA_TRY(launch_b_lm(b, BLM_MU0));
  PSBA_TRY(launch_b_vinv(b));
  PSBA_TRY(launch_b_schur(b));
  PSBA_TRY(launch_b_factor_solve(b));
  PSBA_TRY(launch_b_backsub(b));
  PSBA_TRY(launch_b_lm(b, BLM_STEP_STATS));
  PSBA_TRY(launch_b_lm(b, BLM_TRIAL_COST));
  PSBA_TRY(launch_b_lm(b, BLM_CONTROL));
  PSBA_TRY(launch_b_lm(b, BLM_COMMIT));
  PSBA_HIP_TRY(hipMemcpyAsync(b->h_stat + (size_t)parity * (b->P + 1), b->stat, sizeof(BatchStat) * (size_t)(b->P + 1),
                              hipMemcpyDeviceToHost, b->stream));
  return PSBA_OK;
}

int capture_round(psba_batch* b) {
  if (b->eager) return PSBA_OK;
  if (b->graph_exec[0] && b->graph_iter == b->max_iter && b->graph_tol == b->rel_tol) return PSBA_OK;
  PSBA_TRY(free_graph(b));
  for (int k = 0; k < 2; ++k) {
    PSBA_HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_round(b, k);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (rc != PSBA_OK) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    PSBA_HIP_TRY(e);
    b->graph[k] = g;
    PSBA_HIP_TRY(hipGraphInstantiate(&b->graph_exec[k], b->graph[k], nullptr, nullptr, 0));
  }
  b->graph_iter = b->max_iter;
  b->graph_tol = b->rel_tol;
  return PSBA_OK;
}

int launch_round(psba_batch* b) {
  if (b->eager) PSBA_TRY(enqueue_round(b, b->launched % 2));
  else PSBA_HIP_TRY(hipGraphLaunch(b->graph_exec[b->launched % 2], b->stream));
  PSBA_HIP_TRY(hipEventRecord(b->ev_round[b->launched % psba_batch::kDepth], b->stream));
  b->launched++;
  return PSBA_OK;
}

// wait for the rounds queued behind the last check and read the solve's time
int settle(psba_batch* b) {
  if (!b->stop_pending) return PSBA_OK;
  PSBA_HIP_TRY(hipEventSynchronize(b->ev_stop));
  float t = 0;
  PSBA_HIP_TRY(hipEventElapsedTime(&t, b->ev_start, b->ev_stop));
  b->last_ms = t;
  b->stop_pending = false;
  return PSBA_OK;
}

// the solve has stopped; rounds still queued behind the check end as no-ops
int finish(psba_batch* b) {
  PSBA_HIP_TRY(hipEventRecord(b->ev_stop, b->stream));
  b->stop_pending = true;
  b->rounds = b->checked;
  b->running = false;
  b->finished = true;
  return PSBA_OK;
}

int threads_for(long work, int items) {
  const int hw = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  const int want = (int)std::min<long>(hw, std::max<long>(1, work / 200000));
  return std::max(1, std::min(want, items));
}

template <typename F>
void parallel_for(int n, int nt, F&& fn) {
  if (nt <= 1) {
    for (int p = 0; p < n; ++p) fn(p);
    return;
  }
  std::vector<std::thread> th;
  for (int t = 0; t < nt; ++t)
    th.emplace_back([&, t] {
      for (int p = t; p < n; p += nt) fn(p);
    });
  for (auto& x : th) x.join();
}

int check_problem(const psba_batch* b, int n_cams, int n_pts, int n_obs, const int* iidx, const int* jidx) {
  if (n_cams <= 0 || n_pts <= 0 || n_obs <= 0) return bfail(PSBA_ERR_ARG, "empty problem");
  if (b->cap_p > 0 && (int)b->staged.size() >= b->cap_p)
    return bfail(PSBA_ERR_ARG, "batch holds at most %d problems", b->cap_p);
  const long n = (long)b->cnp * n_cams;
  if (pad64(n) > kMaxSolveN)
    return bfail(PSBA_ERR_UNSUPPORTED, "reduced camera system of order %ld exceeds the limit %d", n, kMaxSolveN);
  for (int k = 0; k < n_obs; ++k) {
    if (iidx[k] < 0 || iidx[k] >= n_pts || jidx[k] < 0 || jidx[k] >= n_cams)
      return bfail(PSBA_ERR_ARG, "observation %d has index out of range", k);
    if (k > 0 && (iidx[k] < iidx[k - 1] || (iidx[k] == iidx[k - 1] && jidx[k] <= jidx[k - 1])))
      return bfail(PSBA_ERR_ARG, "observations must be point-major and camera-ascending (obs %d)", k);
  }
  return PSBA_OK;
}

}  // namespace

namespace psba {

int build_local_index(const psba_batch::Staged& s, int cnp, LocalIndex& L) { return build_local(s, cnp, L); }

int batch_add_staged(psba_batch* b, psba_batch::Staged&& s) {
  b->staged.push_back(std::move(s));
  b->committed = false;
  return (int)b->staged.size() - 1;
}

long long staged_bytes(int cnp, int precision, const psba_batch::Staged& s) {
  std::vector<int> t(s.np, 0);
  for (int k = 0; k < s.no; ++k) t[s.iidx[k]]++;
  long long items = 0;
  for (int v : t) items += (long long)v * (v + 1) / 2;
  return psba_problem_bytes(cnp, precision, s.nc, s.np, s.no, items);
}

int batch_advance(psba_batch* b, bool blocking, void (*cb)(void*, const BatchStat*), void* ctx) {
  if (!b->running) return b->finished ? 1 : 0;
  while (b->launched - b->checked < psba_batch::kDepth && b->launched < b->max_rounds) PSBA_TRY(launch_round(b));
  hipEvent_t ev = b->ev_round[b->checked % psba_batch::kDepth];
  if (blocking) {
    PSBA_HIP_TRY(hipEventSynchronize(ev));
  } else {
    hipError_t q = hipEventQuery(ev);
    if (q == hipErrorNotReady) return 0;
    PSBA_HIP_TRY(q);
  }
  // round `checked` is complete and the next launch would be round checked + 2,
  // so its status buffer is stable until then
  const BatchStat* st = b->h_stat + (size_t)(b->checked % 2) * (b->P + 1);
  b->checked++;
  if (cb) cb(ctx, st);
  if (st[0].active == 0 || b->checked >= b->max_rounds) {
    PSBA_TRY(finish(b));
    return 1;
  }
  if (b->launched < b->max_rounds) PSBA_TRY(launch_round(b));
  return 0;
}

const BatchStat* batch_final_stat(const psba_batch* b) {
  return b->h_stat + (size_t)((b->launched - 1) % 2) * (b->P + 1);
}

int batch_release(psba_batch* b) {
  if (b->running) return bfail(PSBA_ERR_STATE, "a solve is in flight");
  PSBA_HIP_TRY(hipSetDevice(b->device));
  PSBA_HIP_TRY(hipStreamSynchronize(b->stream));
  free_graph(b);
  if (b->arena) (void)hipFree(b->arena);
  b->arena = nullptr;
  b->arena_bytes = 0;
  b->committed = false;
  b->finished = false;
  return PSBA_OK;
}

}  // namespace psba

extern "C" {

long long psba_problem_bytes(int cnp, int precision, int n_cams, int n_pts, int n_obs, long long pair_items) {
  if (!(cnp == 6 || cnp == 9 || cnp == 11 || cnp == 16) || n_cams <= 0 || n_pts <= 0 || n_obs <= 0 || pair_items < 0)
    return -1;
  const long npairs = std::min<long long>((long long)n_cams * (n_cams + 1) / 2, pair_items);
  const Sizes z = problem_sizes(cnp, n_cams, n_pts, n_obs, (long)pair_items, npairs);
  const size_t R = precision == PSBA_PREC_MIXED ? sizeof(float) : sizeof(double);
  psba_batch tmp;
  return (long long)carve(&tmp, z, cnp, R, nullptr);
}

int psba_batch_create(const psba_batch_desc* d, psba_batch** out) {
  if (!d || !out) return bfail(PSBA_ERR_ARG, "null argument");
  *out = nullptr;
  if (!(d->cnp == 6 || d->cnp == 9 || d->cnp == 11 || d->cnp == 16))
    return bfail(PSBA_ERR_UNSUPPORTED, "cnp=%d not supported (6, 9, 11, 16)", d->cnp);
  if (d->precision != PSBA_PREC_FP64 && d->precision != PSBA_PREC_MIXED)
    return bfail(PSBA_ERR_UNSUPPORTED, "precision %d not available", d->precision);
  if (d->max_problems < 0) return bfail(PSBA_ERR_ARG, "max_problems < 0");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return bfail(PSBA_ERR_NODEVICE, "no HIP device visible");
  if (d->device < 0 || d->device >= ndev) return bfail(PSBA_ERR_ARG, "device %d out of range", d->device);
  PSBA_HIP_TRY(hipSetDevice(d->device));
  psba_batch* b = new psba_batch();
  b->cnp = d->cnp;
  b->precision = d->precision;
  b->device = d->device;
  b->cap_p = d->max_problems;
  const char* ev = std::getenv("PSBA_BATCH_EAGER");
  b->eager = ev && ev[0] == '1';
  hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
  for (int k = 0; k < psba_batch::kDepth && e == hipSuccess; ++k) e = hipEventCreate(&b->ev_round[k]);
  if (e == hipSuccess) e = hipEventCreate(&b->ev_start);
  if (e == hipSuccess) e = hipEventCreate(&b->ev_stop);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->ev_staged, hipEventDisableTiming);
  if (e != hipSuccess) {
    psba_batch_destroy(b);
    return bfail(PSBA_ERR_HIP, "batch creation: %s", hipGetErrorString(e));
  }
  *out = b;
  return PSBA_OK;
}

int psba_batch_destroy(psba_batch* b) {
  if (!b) return PSBA_OK;
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  free_graph(b);
  if (b->arena) (void)hipFree(b->arena);
  if (b->staging) (void)hipHostFree(b->staging);
  if (b->h_stat) (void)hipHostFree(b->h_stat);
  for (auto& ev : b->ev_round)
    if (ev) (void)hipEventDestroy(ev);
  if (b->ev_start) (void)hipEventDestroy(b->ev_start);
  if (b->ev_stop) (void)hipEventDestroy(b->ev_stop);
  if (b->ev_staged) (void)hipEventDestroy(b->ev_staged);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return PSBA_OK;
}

long long psba_batch_bytes(const psba_batch* b) { return b ? (long long)b->arena_bytes : -1; }

int psba_batch_reset(psba_batch* b) {
  if (!b) return bfail(PSBA_ERR_ARG, "null batch");
  if (b->running) return bfail(PSBA_ERR_STATE, "a solve is in flight");
  b->staged.clear();
  b->committed = false;
  b->finished = false;
  b->P = 0;
  return PSBA_OK;
}

int psba_batch_add(psba_batch* b, int n_cams, int n_pts, int n_obs, const double* K, const double* impts,
                   const double* initrot, const double* cams, const double* pts, const int* iidx, const int* jidx,
                   const unsigned char* fixmask, int* slot) {
  if (!b || !K || !impts || !initrot || !cams || !pts || !iidx || !jidx) return bfail(PSBA_ERR_ARG, "null argument");
  if (b->running) return bfail(PSBA_ERR_STATE, "a solve is in flight");
  PSBA_TRY(check_problem(b, n_cams, n_pts, n_obs, iidx, jidx));
  psba_batch::Staged s;
  s.nc = n_cams;
  s.np = n_pts;
  s.no = n_obs;
  const int C = b->cnp;
  s.K.assign(K, K + 5 * (size_t)n_cams);
  s.meas.assign(impts, impts + 2 * (size_t)n_obs);
  s.q0.assign(initrot, initrot + 4 * (size_t)n_cams);
  s.cams.assign(cams, cams + (size_t)C * n_cams);
  s.pts.assign(pts, pts + 3 * (size_t)n_pts);
  s.iidx.assign(iidx, iidx + n_obs);
  s.jidx.assign(jidx, jidx + n_obs);
  if (fixmask)
    s.fix.assign(fixmask, fixmask + (size_t)C * n_cams);
  else
    s.fix.assign((size_t)C * n_cams, 0);
  const int k = batch_add_staged(b, std::move(s));
  if (slot) *slot = k;
  return PSBA_OK;
}

int psba_batch_commit(psba_batch* b) {
  if (!b) return bfail(PSBA_ERR_ARG, "null batch");
  if (b->running) return bfail(PSBA_ERR_STATE, "a solve is in flight");
  const int P = (int)b->staged.size();
  if (P == 0) return bfail(PSBA_ERR_STATE, "no problems added");
  PSBA_HIP_TRY(hipSetDevice(b->device));
  const int C = b->cnp;
  long total_obs = 0;
  for (const auto& s : b->staged) total_obs += s.no;
  const int nt = threads_for(total_obs, P);
  // per-problem index in its own numbering
  std::vector<LocalIndex> loc(P);
  std::vector<int> rcs(P, PSBA_OK);
  parallel_for(P, nt, [&](int p) {
    if (!b->staged[p].index) rcs[p] = build_local(b->staged[p], C, loc[p]);
  });
  for (int p = 0; p < P; ++p)
    if (rcs[p] != PSBA_OK) return rcs[p];
  auto L_of = [&](int p) -> const LocalIndex& { return b->staged[p].index ? *b->staged[p].index : loc[p]; };

  // offsets
  std::vector<int> cam_off(P + 1, 0), pt_off(P + 1, 0), obs_off(P + 1, 0), npad(P), vec_off(P + 1, 0);
  std::vector<int> cch_off(P + 1, 0), pair_off(P + 1, 0), sch_off(P + 1, 0), lb_off(P + 1, 0), rch_off(P + 1, 0);
  std::vector<long> item_off(P + 1, 0), S_off(P + 1, 0);
  for (int p = 0; p < P; ++p) {
    const auto& s = b->staged[p];
    const auto& L = L_of(p);
    const long lim = 0x7fffffffL;
    if ((long)obs_off[p] + s.no > lim || item_off[p] + (long)L.items.size() > lim ||
        3L * (pt_off[p] + s.np) + (long)C * (cam_off[p] + s.nc) > lim)
      return bfail(PSBA_ERR_UNSUPPORTED, "batch exceeds 32-bit indexing at problem %d", p);
    cam_off[p + 1] = cam_off[p] + s.nc;
    pt_off[p + 1] = pt_off[p] + s.np;
    obs_off[p + 1] = obs_off[p] + s.no;
    npad[p] = (int)pad64((long)C * s.nc);
    vec_off[p + 1] = vec_off[p] + npad[p];
    cch_off[p + 1] = cch_off[p] + (int)L.cb.size();
    pair_off[p + 1] = pair_off[p] + L.npairs;
    sch_off[p + 1] = sch_off[p] + (int)L.sb.size();
    lb_off[p + 1] = lb_off[p] + (int)L.lbpair.size();
    rch_off[p + 1] = rch_off[p] + (s.no + kResidChunk - 1) / kResidChunk;
    item_off[p + 1] = item_off[p] + (long)L.items.size();
    S_off[p + 1] = S_off[p] + (long)npad[p] * npad[p];
  }
  Sizes z;
  z.P = P;
  z.NC = cam_off[P];
  z.NP = pt_off[P];
  z.NO = obs_off[P];
  z.N = (long)C * z.NC;
  z.items = item_off[P];
  z.n_schunk = sch_off[P];
  z.n_cchunk = cch_off[P];
  z.nlb = lb_off[P];
  z.npairs = pair_off[P];
  z.n_rchunk = rch_off[P];
  z.S_len = S_off[P];
  z.V_len = vec_off[P];
  for (int p = 0; p < P; ++p) z.T_max = std::max(z.T_max, (long)(npad[p] / kNB));

  // device arena (grows only) and pinned staging image of its input part
  const size_t R = b->mixed() ? sizeof(float) : sizeof(double);
  size_t in_bytes = 0;
  psba_batch hv;  // views into the staging image
  const size_t need = carve(&hv, z, C, R, nullptr, &in_bytes);
  PSBA_TRY(free_graph(b));
  PSBA_HIP_TRY(hipEventSynchronize(b->ev_staged));  // the previous upload has left the staging image
  if (need > b->arena_bytes) {
    PSBA_HIP_TRY(hipStreamSynchronize(b->stream));
    if (b->arena) (void)hipFree(b->arena);
    b->arena = nullptr;
    b->arena_bytes = 0;
    hipError_t e = hipMalloc((void**)&b->arena, need);
    if (e != hipSuccess) {
      b->arena = nullptr;
      return bfail(PSBA_ERR_NOMEM, "hipMalloc of %zu bytes failed: %s", need, hipGetErrorString(e));
    }
    b->arena_bytes = need;
  }
  if (in_bytes > b->staging_bytes) {
    if (b->staging) (void)hipHostFree(b->staging);
    b->staging = nullptr;
    b->staging_bytes = 0;
    hipError_t e = hipHostMalloc((void**)&b->staging, in_bytes, hipHostMallocDefault);
    if (e != hipSuccess) {
      b->staging = nullptr;
      return bfail(PSBA_ERR_NOMEM, "pinned allocation of %zu bytes failed: %s", in_bytes, hipGetErrorString(e));
    }
    b->staging_bytes = in_bytes;
  }
  if (2 * (size_t)(P + 1) > b->h_stat_cap) {
    PSBA_HIP_TRY(hipStreamSynchronize(b->stream));  // a trailing round may still write it
    if (b->h_stat) (void)hipHostFree(b->h_stat);
    b->h_stat = nullptr;
    b->h_stat_cap = 0;
    hipError_t e = hipHostMalloc((void**)&b->h_stat, sizeof(BatchStat) * 2 * (size_t)(P + 1), hipHostMallocDefault);
    if (e != hipSuccess) {
      b->h_stat = nullptr;
      return bfail(PSBA_ERR_NOMEM, "pinned allocation failed: %s", hipGetErrorString(e));
    }
    b->h_stat_cap = 2 * (size_t)(P + 1);
  }
  carve(&hv, z, C, R, b->staging);

  // fill the staging image with the global arrays
  parallel_for(P, nt, [&](int p) {
    const auto& s = b->staged[p];
    const auto& L = L_of(p);
    const int co = cam_off[p], po = pt_off[p], oo = obs_off[p];
    std::copy(s.K.begin(), s.K.end(), hv.K + 5L * co);
    std::copy(s.q0.begin(), s.q0.end(), hv.q0 + 4L * co);
    std::copy(s.meas.begin(), s.meas.end(), hv.meas + 2L * oo);
    std::copy(s.cams.begin(), s.cams.end(), hv.cams0 + (long)C * co);
    std::copy(s.pts.begin(), s.pts.end(), hv.pts0 + 3L * po);
    std::copy(s.fix.begin(), s.fix.end(), hv.fix + (long)C * co);
    for (int k = 0; k < s.no; ++k) {
      hv.obs_cam[oo + k] = co + s.jidx[k];
      hv.obs_pt[oo + k] = po + s.iidx[k];
      hv.obs_prob[oo + k] = p;
      hv.cam_obs[oo + k] = oo + L.cam_obs[k];
    }
    for (int i = 0; i < s.np; ++i) {
      hv.pt_ptr[po + i] = oo + L.pt_ptr[i];
      hv.pt_prob[po + i] = p;
    }
    for (int j = 0; j < s.nc; ++j) {
      hv.cam_prob[co + j] = p;
      hv.ccp[co + j] = cch_off[p] + L.ccp[j];
    }
    for (size_t q = 0; q < L.cb.size(); ++q) {
      hv.cchunk_b[cch_off[p] + q] = oo + L.cb[q];
      hv.cchunk_e[cch_off[p] + q] = oo + L.ce[q];
      hv.cchunk_prob[cch_off[p] + q] = p;
    }
    const long io = item_off[p];
    for (size_t q = 0; q < L.items.size(); ++q)
      hv.items[(size_t)io + q] = make_int2(oo + L.items[q].x, oo + L.items[q].y);
    for (size_t q = 0; q < L.sb.size(); ++q) {
      hv.schunk_b[sch_off[p] + q] = (int)(io + L.sb[q]);
      hv.schunk_e[sch_off[p] + q] = (int)(io + L.se[q]);
      hv.schunk_prob[sch_off[p] + q] = p;
    }
    for (int q = 0; q < L.npairs; ++q) hv.psp[pair_off[p] + q] = sch_off[p] + L.psp[q];
    for (int k = 0, f = lb_off[p]; k < s.nc; ++k)
      for (int l = 0; l <= k; ++l, ++f) {
        const int lq = L.lbpair[(size_t)k * (k + 1) / 2 + l];
        hv.lb_prob[f] = p;
        hv.lb_k[f] = k;
        hv.lb_l[f] = l;
        hv.lb_pair[f] = lq < 0 ? -1 : pair_off[p] + lq;
      }
    for (int q = 0, a = 0; a < s.no; ++q, a += kResidChunk) {
      hv.rchunk_b[rch_off[p] + q] = oo + a;
      hv.rchunk_e[rch_off[p] + q] = oo + std::min(a + kResidChunk, s.no);
      hv.rchunk_prob[rch_off[p] + q] = p;
    }
    hv.rcp[p] = rch_off[p];
    hv.cam_off[p] = co;
    hv.pt_off[p] = po;
    hv.obs_off[p] = oo;
    hv.npad[p] = npad[p];
    hv.vec_off[p] = vec_off[p];
    hv.S_off[p] = S_off[p];
    hv.all_on[p] = 1;
  });
  hv.pt_ptr[z.NP] = (int)z.NO;
  hv.ccp[z.NC] = (int)z.n_cchunk;
  hv.psp[z.npairs] = (int)z.n_schunk;
  hv.rcp[P] = (int)z.n_rchunk;
  hv.cam_off[P] = cam_off[P];
  hv.pt_off[P] = pt_off[P];
  hv.obs_off[P] = obs_off[P];
  hv.vec_off[P] = vec_off[P];
  hv.S_off[P] = S_off[P];
  // task tables of the phase-by-phase dense solve
  const int Tm = (int)z.T_max;
  std::vector<int> Gf(std::max(0, Tm - 1), 0), Gb(Tm, 0);
  for (int k = 0; k + 1 < Tm; ++k) {
    int* fo = hv.chol_fo + (size_t)k * (P + 1);
    fo[0] = 0;
    for (int p = 0; p < P; ++p) {
      const int m = npad[p] / kNB - 1 - k;
      fo[p + 1] = fo[p] + (m >= 1 ? m * (m + 1) / 2 : 0);
    }
    Gf[k] = fo[P];
  }
  for (int s = 0; s < Tm; ++s) {
    int* bo = hv.chol_bo + (size_t)s * (P + 1);
    bo[0] = 0;
    for (int p = 0; p < P; ++p) {
      const int T = npad[p] / kNB;
      bo[p + 1] = bo[p] + ((T >= 2 && T > s) ? (s == 0 ? 1 : T - s) : 0);
    }
    Gb[s] = bo[P];
  }

  carve(b, z, C, R, b->arena);
  b->input_bytes = in_bytes;
  b->P = P;
  b->NC = z.NC;
  b->NP = z.NP;
  b->NO = z.NO;
  b->N = z.N;
  b->n_items = z.items;
  b->S_len = z.S_len;
  b->V_len = z.V_len;
  b->n_cchunk = (int)z.n_cchunk;
  b->n_schunk = (int)z.n_schunk;
  b->nlb = (int)z.nlb;
  b->npairs = (int)z.npairs;
  b->n_rchunk = (int)z.n_rchunk;
  b->h_cam_off = cam_off;
  b->h_pt_off = pt_off;
  b->h_obs_off = obs_off;
  b->T_max = Tm;
  b->chol_phased = Tm >= kPhasedMinT;
  b->h_Gf = Gf;
  b->h_Gb = Gb;

  // one upload; the dense blocks start as zeros so that untouched upper parts
  // are finite
  PSBA_HIP_TRY(hipMemcpyAsync(b->arena, b->staging, in_bytes, hipMemcpyHostToDevice, b->stream));
  PSBA_HIP_TRY(hipEventRecord(b->ev_staged, b->stream));
  PSBA_HIP_TRY(hipMemsetAsync(b->S, 0, sizeof(double) * (size_t)z.S_len, b->stream));
  PSBA_HIP_TRY(hipMemsetAsync(b->Lf, 0, sizeof(double) * (size_t)z.S_len, b->stream));
  PSBA_HIP_TRY(hipMemsetAsync(b->dp, 0, sizeof(double) * (size_t)(z.N + 3 * z.NP), b->stream));
  b->committed = true;
  b->finished = false;
  return PSBA_OK;
}

int psba_batch_launch(psba_batch* b, const psba_opts* o) {
  if (!b) return bfail(PSBA_ERR_ARG, "null batch");
  if (!b->committed) return bfail(PSBA_ERR_STATE, "batch not committed");
  if (b->running) return bfail(PSBA_ERR_STATE, "a solve is in flight");
  PSBA_HIP_TRY(hipSetDevice(b->device));
  b->max_iter = (o && o->max_iter > 0) ? o->max_iter : 100;
  b->rel_tol = o ? o->rel_tol : 1e-8;
  // every round ends an iteration or doubles nu; nu reaches 2^30 within 30
  // doublings, so each iteration takes at most 31 rounds
  b->max_rounds = b->max_iter * 31 + 2;
  PSBA_TRY(capture_round(b));
  b->stop_pending = false;  // an unread time of the previous solve is dropped
  PSBA_HIP_TRY(hipEventRecord(b->ev_start, b->stream));
  PSBA_HIP_TRY(hipMemcpyAsync(b->cams, b->cams0, sizeof(double) * (size_t)b->N, hipMemcpyDeviceToDevice, b->stream));
  PSBA_HIP_TRY(
      hipMemcpyAsync(b->pts, b->pts0, sizeof(double) * 3 * (size_t)b->NP, hipMemcpyDeviceToDevice, b->stream));
  PSBA_TRY(launch_b_lm(b, BLM_INIT));
  b->launched = 0;
  b->checked = 0;
  b->running = true;
  b->finished = false;
  for (int k = 0; k < psba_batch::kDepth && b->launched < b->max_rounds; ++k) PSBA_TRY(launch_round(b));
  return PSBA_OK;
}

int psba_batch_poll(psba_batch* b) {
  if (!b) return bfail(PSBA_ERR_ARG, "null batch");
  if (!b->running) return b->finished ? 1 : bfail(PSBA_ERR_STATE, "no solve launched");
  PSBA_HIP_TRY(hipSetDevice(b->device));
  // drain every round that has completed
  for (;;) {
    const int r = batch_advance(b, false);
    if (r != 0) return r;
    if (b->launched == b->checked) return 0;
    if (hipEventQuery(b->ev_round[b->checked % psba_batch::kDepth]) == hipErrorNotReady) return 0;
  }
}

int psba_batch_solve(psba_batch* b, const psba_opts* o, int* rounds) {
  PSBA_TRY(psba_batch_launch(b, o));
  for (;;) {
    const int r = batch_advance(b, true);
    if (r < 0) return r;
    if (r == 1) break;
  }
  if (rounds) *rounds = b->rounds;
  return PSBA_OK;
}

int psba_batch_result(psba_batch* b, int slot, int* code, int* itno, double* init_err, double* final_err,
                      double* cams, double* pts) {
  if (!b) return bfail(PSBA_ERR_ARG, "null batch");
  if (!b->finished) return bfail(PSBA_ERR_STATE, "no finished solve");
  if (slot < 0 || slot >= b->P) return bfail(PSBA_ERR_ARG, "slot %d out of range", slot);
  PSBA_HIP_TRY(hipSetDevice(b->device));
  PSBA_TRY(settle(b));
  const BatchStat& st = batch_final_stat(b)[slot + 1];
  if (code) *code = st.code;
  if (itno) *itno = st.itno;
  if (init_err) *init_err = st.init_err;
  if (final_err) *final_err = st.cost;
  const int C = b->cnp;
  if (cams)
    PSBA_HIP_TRY(hipMemcpy(cams, b->cams + (size_t)C * b->h_cam_off[slot],
                           sizeof(double) * C * (b->h_cam_off[slot + 1] - b->h_cam_off[slot]), hipMemcpyDeviceToHost));
  if (pts)
    PSBA_HIP_TRY(hipMemcpy(pts, b->pts + 3 * (size_t)b->h_pt_off[slot],
                           sizeof(double) * 3 * (b->h_pt_off[slot + 1] - b->h_pt_off[slot]), hipMemcpyDeviceToHost));
  return PSBA_OK;
}

int psba_batch_elapsed(psba_batch* b, double* ms, int* rounds) {
  if (!b) return bfail(PSBA_ERR_ARG, "null batch");
  if (!b->finished) return bfail(PSBA_ERR_STATE, "no finished solve");
  PSBA_TRY(settle(b));
  if (ms) *ms = b->last_ms;
  if (rounds) *rounds = b->rounds;
  return PSBA_OK;
}

int psba_batch_autotune(psba_batch* b, int* obs_block, int* pt_block) {
  if (!b) return bfail(PSBA_ERR_ARG, "null batch");
  if (!b->committed) return bfail(PSBA_ERR_STATE, "batch not committed");
  if (b->running) return bfail(PSBA_ERR_STATE, "a solve is in flight");
  PSBA_HIP_TRY(hipSetDevice(b->device));
  PSBA_TRY(settle(b));
  // state as at the start of a solve: every problem needs its Jacobian
  b->max_iter = 100;
  PSBA_HIP_TRY(hipMemcpyAsync(b->cams, b->cams0, sizeof(double) * (size_t)b->N, hipMemcpyDeviceToDevice, b->stream));
  PSBA_HIP_TRY(
      hipMemcpyAsync(b->pts, b->pts0, sizeof(double) * 3 * (size_t)b->NP, hipMemcpyDeviceToDevice, b->stream));
  PSBA_TRY(launch_b_lm(b, BLM_INIT));
  PSBA_TRY(launch_b_lm(b, BLM_MU0));
  const int cand[3] = {64, 128, 256};
  auto time_it = [&](auto&& fn) -> double {
    PSBA_TRY(fn());  // warm-up
    PSBA_HIP_TRY(hipEventRecord(b->ev_start, b->stream));
    for (int r = 0; r < 3; ++r) PSBA_TRY(fn());
    PSBA_HIP_TRY(hipEventRecord(b->ev_stop, b->stream));
    PSBA_HIP_TRY(hipEventSynchronize(b->ev_stop));
    float ms = 0;
    PSBA_HIP_TRY(hipEventElapsedTime(&ms, b->ev_start, b->ev_stop));
    return ms;
  };
  double best = 1e30;
  int bo = b->obs_block;
  for (int c : cand) {
    b->obs_block = c;
    const double t = time_it([&] { return launch_b_jacobian(b); });
    if (t < 0) return (int)t;
    if (t < best) {
      best = t;
      bo = c;
    }
  }
  b->obs_block = bo;
  best = 1e30;
  int bp = b->pt_block;
  for (int c : cand) {
    b->pt_block = c;
    const double t = time_it([&] {
      int r = launch_b_vinv(b);
      return r != PSBA_OK ? r : launch_b_backsub(b);
    });
    if (t < 0) return (int)t;
    if (t < best) {
      best = t;
      bp = c;
    }
  }
  b->pt_block = bp;
  PSBA_TRY(free_graph(b));
  if (obs_block) *obs_block = bo;
  if (pt_block) *pt_block = bp;
  return PSBA_OK;
}

}  // extern "C"
