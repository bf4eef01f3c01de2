// Request loop over batches (include/psba_server.h).  No counterpart in the
// reference (one problem per process, PSBA/main.cpp:193-208).
//
// State: a FIFO of queued requests, `lanes` batches (each on its own stream,
// its arena kept between uses) and the requests admitted to each.  A step
// retires requests whose solve stopped (from the per-problem status every
// round copies to pinned memory), copies their parameters back on the lane's
// copy stream, and admits queued requests into idle lanes within the budget.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdarg>
#include <cstring>
#include <deque>
#include <functional>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "batch.hpp"
#include "context.hpp"
#include "psba_server.h"

using namespace psba;

namespace {

int sfail(int code, const char* fmt, ...) {
  char buf[400];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  set_last_error(__FILE__, __LINE__, buf);
  return code;
}

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

enum ReqState { RQ_QUEUED = 0, RQ_RUNNING, RQ_COPYING, RQ_DONE };

struct Request {
  long long id = 0;
  psba_batch::Staged data;
  std::atomic<int> ready{0};  // index built (by a worker) and bytes known
  int build_rc = PSBA_OK;
  long long bytes = 0;
  int cls = 0;
  int state = RQ_QUEUED;
  double t_submit = 0, t_done = 0;
  int code = PSBA_ITER_CONTINUE, itno = 0;
  double init_err = 0, final_err = 0;
  std::vector<double> cams, pts;
  int nc = 0, np = 0;
};

struct CopyGroup {
  hipEvent_t ev;
  std::vector<int> slots;
};

struct Lane {
  psba_batch* b = nullptr;
  hipStream_t copy_stream = nullptr;
  double* h_params = nullptr;  // pinned; cams then points of the whole batch
  size_t h_params_cap = 0;
  std::vector<long long> ids;  // slot -> request
  std::vector<char> retired;
  std::vector<CopyGroup> copies;
  int left = 0;                // admitted requests not yet complete
  bool busy = false;
  bool solve_done = false;
};

// size class of a request: dense tiles of its reduced system
int size_class(int cnp, int nc) {
  const long T = ((long)cnp * nc + kNB - 1) / kNB;
  return T <= 2 ? 0 : (T <= 8 ? 1 : 2);
}

// host workers that build each request's index while earlier requests run
struct Workers {
  std::vector<std::thread> th;
  std::mutex m;
  std::condition_variable cv;
  std::deque<std::function<void()>> jobs;
  bool stop = false;
  void start(int n) {
    for (int k = 0; k < n; ++k)
      th.emplace_back([this] {
        for (;;) {
          std::function<void()> job;
          {
            std::unique_lock<std::mutex> lk(m);
            cv.wait(lk, [this] { return stop || !jobs.empty(); });
            if (jobs.empty()) return;
            job = std::move(jobs.front());
            jobs.pop_front();
          }
          job();
        }
      });
  }
  void push(std::function<void()> f) {
    {
      std::lock_guard<std::mutex> lk(m);
      jobs.push_back(std::move(f));
    }
    cv.notify_one();
  }
  void join() {
    {
      std::lock_guard<std::mutex> lk(m);
      stop = true;
    }
    cv.notify_all();
    for (auto& t : th) t.join();
    th.clear();
  }
};

}  // namespace

struct psba_server {
  psba_server_desc d{};
  std::vector<Lane> lanes;
  std::deque<long long> queue;
  std::unordered_map<long long, Request> reqs;
  std::vector<hipEvent_t> ev_pool;
  std::deque<long long> done_ids;  // completed, not yet listed by psba_server_completed
  long long next_id = 1;
  psba_server_stats st{};
  psba_opts opts{};
  Workers workers;
};

namespace {

long long held(const psba_server* s) {
  long long h = 0;
  for (const auto& l : s->lanes) h += psba_batch_bytes(l.b);
  return h;
}

int get_event(psba_server* s, hipEvent_t* ev) {
  if (!s->ev_pool.empty()) {
    *ev = s->ev_pool.back();
    s->ev_pool.pop_back();
    return PSBA_OK;
  }
  PSBA_HIP_TRY(hipEventCreateWithFlags(ev, hipEventDisableTiming));
  return PSBA_OK;
}

struct Collect {
  psba_server* s;
  Lane* L;
  std::vector<int> now;
};

// record the problems that stopped in a checked round (its status is stable
// only during this call)
void collect_stopped(void* ctx, const BatchStat* st) {
  Collect* c = (Collect*)ctx;
  Lane& L = *c->L;
  for (int p = 0; p < (int)L.retired.size(); ++p) {
    if (L.retired[p] || st[p + 1].active != 0) continue;
    L.retired[p] = 1;
    Request& q = c->s->reqs[L.ids[p]];
    q.code = st[p + 1].code;
    q.itno = st[p + 1].itno;
    q.init_err = st[p + 1].init_err;
    q.final_err = st[p + 1].cost;
    q.state = RQ_COPYING;
    c->now.push_back(p);
  }
}

// retire the stopped problems of a running lane; returns requests completed
int service_lane(psba_server* s, Lane& L) {
  int completed = 0;
  psba_batch* b = L.b;
  if (!L.solve_done) {
    Collect c{s, &L, {}};
    for (;;) {
      const int before = b->checked;
      const int r = batch_advance(b, false, collect_stopped, &c);
      if (r < 0) return r;
      if (r == 1) {
        L.solve_done = true;
        // a solve cut at its round limit: the rest stop where they are
        const BatchStat* st = batch_final_stat(b);
        for (int p = 0; p < b->P; ++p)
          if (!L.retired[p]) {
            const BatchStat f = st[p + 1];
            L.retired[p] = 1;
            Request& q = s->reqs[L.ids[p]];
            q.code = f.code;
            q.itno = f.itno;
            q.init_err = f.init_err;
            q.final_err = f.cost;
            q.state = RQ_COPYING;
            c.now.push_back(p);
          }
        break;
      }
      if (b->checked == before) break;  // the oldest round is still running
    }
    if (!c.now.empty()) {
      CopyGroup g;
      PSBA_TRY(get_event(s, &g.ev));
      const int C = b->cnp;
      for (int p : c.now) {
        const size_t c0 = (size_t)C * b->h_cam_off[p], c1 = (size_t)C * b->h_cam_off[p + 1];
        const size_t p0 = 3 * (size_t)b->h_pt_off[p], p1 = 3 * (size_t)b->h_pt_off[p + 1];
        PSBA_HIP_TRY(hipMemcpyAsync(L.h_params + c0, b->cams + c0, sizeof(double) * (c1 - c0),
                                    hipMemcpyDeviceToHost, L.copy_stream));
        PSBA_HIP_TRY(hipMemcpyAsync(L.h_params + b->N + p0, b->pts + p0, sizeof(double) * (p1 - p0),
                                    hipMemcpyDeviceToHost, L.copy_stream));
      }
      g.slots = std::move(c.now);
      PSBA_HIP_TRY(hipEventRecord(g.ev, L.copy_stream));
      L.copies.push_back(std::move(g));
    }
  }
  // copies that have landed
  size_t k = 0;
  for (; k < L.copies.size(); ++k) {
    CopyGroup& g = L.copies[k];
    hipError_t e = hipEventQuery(g.ev);
    if (e == hipErrorNotReady) break;
    PSBA_HIP_TRY(e);
    const double t = now_s();
    const int C = b->cnp;
    for (int p : g.slots) {
      Request& q = s->reqs[L.ids[p]];
      const size_t c0 = (size_t)C * b->h_cam_off[p], c1 = (size_t)C * b->h_cam_off[p + 1];
      const size_t p0 = 3 * (size_t)b->h_pt_off[p], p1 = 3 * (size_t)b->h_pt_off[p + 1];
      q.cams.assign(L.h_params + c0, L.h_params + c1);
      q.pts.assign(L.h_params + b->N + p0, L.h_params + b->N + p1);
      q.state = RQ_DONE;
      q.t_done = t;
      s->done_ids.push_back(q.id);
      L.left--;
      completed++;
      s->st.completed++;
      s->st.running--;
    }
    s->ev_pool.push_back(g.ev);
  }
  L.copies.erase(L.copies.begin(), L.copies.begin() + (long)k);
  if (L.solve_done && L.left == 0 && L.copies.empty()) L.busy = false;
  return completed;
}

// form and launch a batch on an idle lane; returns 1 if one was launched
int admit(psba_server* s, Lane& L) {
  // the oldest request whose index is built; requests that can never run
  // (index failure, larger than the whole budget) complete with an error
  Request* head = nullptr;
  for (auto it = s->queue.begin(); it != s->queue.end();) {
    Request& q = s->reqs[*it];
    if (!q.ready.load(std::memory_order_acquire)) {
      ++it;
      continue;
    }
    if (q.build_rc != PSBA_OK || q.bytes > s->d.hbm_budget) {
      q.state = RQ_DONE;
      q.code = q.build_rc != PSBA_OK ? q.build_rc : PSBA_ERR_NOMEM;
      q.t_done = now_s();
      s->done_ids.push_back(q.id);
      s->st.rejected++;
      it = s->queue.erase(it);
      continue;
    }
    head = &q;
    break;
  }
  if (!head) return 0;
  long long avail = s->d.hbm_budget - (held(s) - psba_batch_bytes(L.b));
  const int cls = head->cls;
  if (head->bytes > avail) {
    // give back arenas of the other idle lanes, then wait for running ones
    for (auto& o : s->lanes)
      if (&o != &L && !o.busy && psba_batch_bytes(o.b) > 0) PSBA_TRY(batch_release(o.b));
    avail = s->d.hbm_budget - (held(s) - psba_batch_bytes(L.b));
    if (head->bytes > avail) return 0;
  }
  std::vector<long long> pick;
  long long bytes = 0;
  for (auto it = s->queue.begin(); it != s->queue.end() && (int)pick.size() < s->d.max_batch;) {
    Request& q = s->reqs[*it];
    if (q.ready.load(std::memory_order_acquire) && q.build_rc == PSBA_OK && q.cls == cls &&
        bytes + q.bytes <= avail) {
      pick.push_back(*it);
      bytes += q.bytes;
      it = s->queue.erase(it);
    } else {
      ++it;
    }
  }
  psba_batch* b = L.b;
  PSBA_TRY(psba_batch_reset(b));
  L.ids = pick;
  L.retired.assign(pick.size(), 0);
  L.copies.clear();
  for (long long id : pick) {
    Request& q = s->reqs[id];
    batch_add_staged(b, std::move(q.data));
    q.data = psba_batch::Staged();
    q.state = RQ_RUNNING;
  }
  PSBA_TRY(psba_batch_commit(b));
  const size_t np = (size_t)b->N + 3 * (size_t)b->NP;
  if (np > L.h_params_cap) {
    if (L.h_params) (void)hipHostFree(L.h_params);
    L.h_params = nullptr;
    L.h_params_cap = 0;
    PSBA_HIP_TRY(hipHostMalloc((void**)&L.h_params, sizeof(double) * np, hipHostMallocDefault));
    L.h_params_cap = np;
  }
  PSBA_TRY(psba_batch_launch(b, &s->opts));
  L.left = (int)pick.size();
  L.busy = true;
  L.solve_done = false;
  s->st.batches++;
  s->st.running += (int)pick.size();
  s->st.bytes_peak = std::max(s->st.bytes_peak, held(s));
  return 1;
}

}  // namespace

extern "C" {

double psba_server_clock(void) { return now_s(); }

int psba_server_create(const psba_server_desc* d, psba_server** out) {
  if (!d || !out) return sfail(PSBA_ERR_ARG, "null argument");
  *out = nullptr;
  psba_server* s = new psba_server();
  s->d = *d;
  if (s->d.lanes <= 0) s->d.lanes = 4;
  if (s->d.max_batch <= 0) s->d.max_batch = 512;
  if (s->d.hbm_budget <= 0) s->d.hbm_budget = 200000000000LL;
  if (s->d.max_iter <= 0) s->d.max_iter = 100;
  if (s->d.rel_tol < 0) s->d.rel_tol = 1e-8;
  s->opts.max_iter = s->d.max_iter;
  s->opts.rel_tol = s->d.rel_tol;
  s->lanes.resize(s->d.lanes);
  s->workers.start((int)std::max(1u, std::min(8u, std::thread::hardware_concurrency() / 2)));
  for (auto& L : s->lanes) {
    psba_batch_desc bd{d->cnp, d->precision, d->device, 0};
    int rc = psba_batch_create(&bd, &L.b);
    if (rc == PSBA_OK && hipStreamCreateWithFlags(&L.copy_stream, hipStreamNonBlocking) != hipSuccess)
      rc = sfail(PSBA_ERR_HIP, "stream creation failed");
    if (rc != PSBA_OK) {
      psba_server_destroy(s);
      return rc;
    }
  }
  *out = s;
  return PSBA_OK;
}

int psba_server_destroy(psba_server* s) {
  if (!s) return PSBA_OK;
  s->workers.join();
  for (auto& L : s->lanes) {
    if (L.b) {
      (void)hipSetDevice(L.b->device);
      (void)hipStreamSynchronize(L.b->stream);
    }
    if (L.copy_stream) {
      (void)hipStreamSynchronize(L.copy_stream);
      (void)hipStreamDestroy(L.copy_stream);
    }
    for (auto& g : L.copies) (void)hipEventDestroy(g.ev);
    if (L.h_params) (void)hipHostFree(L.h_params);
    if (L.b) {
      L.b->running = false;  // the stream is drained
      psba_batch_destroy(L.b);
    }
  }
  for (auto e : s->ev_pool) (void)hipEventDestroy(e);
  delete s;
  return PSBA_OK;
}

int psba_server_submit(psba_server* s, int n_cams, int n_pts, int n_obs, const double* K, const double* impts,
                       const double* initrot, const double* cams, const double* pts, const int* iidx,
                       const int* jidx, const unsigned char* fixmask, long long* id) {
  if (!s || !K || !impts || !initrot || !cams || !pts || !iidx || !jidx || !id)
    return sfail(PSBA_ERR_ARG, "null argument");
  const int C = s->d.cnp;
  if (n_cams <= 0 || n_pts <= 0 || n_obs <= 0) return sfail(PSBA_ERR_ARG, "empty problem");
  if ((((long)C * n_cams + kNB - 1) / kNB) * kNB > kMaxSolveN)
    return sfail(PSBA_ERR_UNSUPPORTED, "reduced camera system exceeds the limit %d", kMaxSolveN);
  for (int k = 0; k < n_obs; ++k) {
    if (iidx[k] < 0 || iidx[k] >= n_pts || jidx[k] < 0 || jidx[k] >= n_cams)
      return sfail(PSBA_ERR_ARG, "observation %d has index out of range", k);
    if (k > 0 && (iidx[k] < iidx[k - 1] || (iidx[k] == iidx[k - 1] && jidx[k] <= jidx[k - 1])))
      return sfail(PSBA_ERR_ARG, "observations must be point-major and camera-ascending (obs %d)", k);
  }
  const long long rid = s->next_id++;
  Request& q = s->reqs[rid];
  q.id = rid;
  q.t_submit = now_s();
  q.nc = n_cams;
  q.np = n_pts;
  auto& d = q.data;
  d.nc = n_cams;
  d.np = n_pts;
  d.no = n_obs;
  d.K.assign(K, K + 5 * (size_t)n_cams);
  d.meas.assign(impts, impts + 2 * (size_t)n_obs);
  d.q0.assign(initrot, initrot + 4 * (size_t)n_cams);
  d.cams.assign(cams, cams + (size_t)C * n_cams);
  d.pts.assign(pts, pts + 3 * (size_t)n_pts);
  d.iidx.assign(iidx, iidx + n_obs);
  d.jidx.assign(jidx, jidx + n_obs);
  if (fixmask)
    d.fix.assign(fixmask, fixmask + (size_t)C * n_cams);
  else
    d.fix.assign((size_t)C * n_cams, 0);
  q.cls = size_class(C, n_cams);
  const int prec = s->d.precision;
  Request* qp = &q;  // element references of the map stay valid
  s->workers.push([qp, C, prec] {
    auto idx = std::make_shared<LocalIndex>();
    qp->build_rc = build_local_index(qp->data, C, *idx);
    if (qp->build_rc == PSBA_OK) {
      qp->bytes = psba_problem_bytes(C, prec, qp->data.nc, qp->data.np, qp->data.no, (long long)idx->items.size());
      qp->data.index = std::move(idx);
    }
    qp->ready.store(1, std::memory_order_release);
  });
  *id = rid;
  s->queue.push_back(rid);
  s->st.submitted++;
  return PSBA_OK;
}

int psba_server_step(psba_server* s) {
  if (!s) return sfail(PSBA_ERR_ARG, "null server");
  int completed = 0;
  for (auto& L : s->lanes)
    if (L.busy) {
      const int r = service_lane(s, L);
      if (r < 0) return r;
      completed += r;
    }
  for (auto& L : s->lanes) {
    if (s->queue.empty()) break;
    if (L.busy) continue;
    const int r = admit(s, L);
    if (r < 0) return r;
    if (r == 0) break;  // waiting for memory
  }
  s->st.bytes_held = held(s);
  s->st.bytes_peak = std::max(s->st.bytes_peak, s->st.bytes_held);
  return completed;
}

int psba_server_drain(psba_server* s) {
  if (!s) return sfail(PSBA_ERR_ARG, "null server");
  for (;;) {
    const int r = psba_server_step(s);
    if (r < 0) return r;
    bool busy = !s->queue.empty();
    for (auto& L : s->lanes) busy = busy || L.busy;
    if (!busy) return PSBA_OK;
  }
}

int psba_server_result(psba_server* s, long long id, int* code, int* itno, double* init_err, double* final_err,
                       double* cams, double* pts, double* t_submit, double* t_done) {
  if (!s) return sfail(PSBA_ERR_ARG, "null server");
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return sfail(PSBA_ERR_ARG, "unknown request %lld", id);
  Request& q = it->second;
  if (q.state != RQ_DONE) return 0;
  if (code) *code = q.code;
  if (itno) *itno = q.itno;
  if (init_err) *init_err = q.init_err;
  if (final_err) *final_err = q.final_err;
  if (cams && !q.cams.empty()) std::memcpy(cams, q.cams.data(), sizeof(double) * q.cams.size());
  if (pts && !q.pts.empty()) std::memcpy(pts, q.pts.data(), sizeof(double) * q.pts.size());
  if (t_submit) *t_submit = q.t_submit;
  if (t_done) *t_done = q.t_done;
  s->reqs.erase(it);
  return 1;
}

int psba_server_completed(psba_server* s, long long* ids, int cap) {
  if (!s || (!ids && cap > 0)) return sfail(PSBA_ERR_ARG, "null argument");
  int n = 0;
  while (n < cap && !s->done_ids.empty()) {
    ids[n++] = s->done_ids.front();
    s->done_ids.pop_front();
  }
  return n;
}

int psba_server_get_stats(psba_server* s, psba_server_stats* out) {
  if (!s || !out) return sfail(PSBA_ERR_ARG, "null argument");
  s->st.queued = (int)s->queue.size();
  s->st.bytes_held = held(s);
  *out = s->st;
  return PSBA_OK;
}

}  // extern "C"
