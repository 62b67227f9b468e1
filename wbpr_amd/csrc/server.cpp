// Serving loop: requests are admitted by device-memory footprint into launches that run
// concurrently, each on its own context (HIP stream) driven by a worker thread, so work of
// requests that arrive while others are on the GPU overlaps on the device.  Small requests are
// grouped into one persistent batched launch (a workgroup per instance); a request too large for
// one workgroup to solve quickly runs alone on the whole grid.  Admission is first come, first
// served; finished launches are retired and their memory budget is reused.
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "mf_internal.h"
#include "solver.h"

using namespace mf;

namespace {

struct Request {
  int64_t id;
  std::vector<int32_t> off, dst, cap;
  int32_t n, s, t;
  int64_t m;
  int64_t footprint;
  bool solo = false;  // solved alone by the grid-wide solver
  int state = 0;  // 0 pending, 1 running, 2 done, 3 failed
  int64_t value = 0;
  double device_ms = 0;
  mf_csr view() const {
    mf_csr c;
    c.n = n;
    c.m = m;
    c.offsets = off.data();
    c.dst = dst.data();
    c.cap = cap.data();
    c.s = s;
    c.t = t;
    return c;
  }
};

struct Launch {
  std::vector<std::shared_ptr<Request>> reqs;
  int64_t bytes = 0;
  int ctx_slot = -1;
  std::thread worker;
  bool finished = false;
  mf_status status = MF_OK;
};

// Requests above either size run alone on the whole grid: one workgroup would take tens of
// milliseconds on them while the grid-wide solver takes a few.
constexpr int32_t kSoloVertices = 16384;
constexpr int64_t kSoloArcs = 262144;

}  // namespace

struct mf_server {
  mf_context *owner = nullptr;
  int device = 0;
  int64_t budget = 0, in_use = 0, peak = 0;
  int32_t max_inflight = 1;
  std::vector<mf_context *> ctxs;
  std::vector<bool> ctx_busy;
  std::deque<std::shared_ptr<Request>> pending;
  std::map<int64_t, std::shared_ptr<Request>> all;
  std::vector<std::unique_ptr<Launch>> inflight;
  int64_t next_id = 1;
  mf_server_stats st{};
  std::mutex mu;
  std::condition_variable cv;
};

extern "C" int64_t mf_footprint_bytes(int32_t n, int64_t m) {
  // BCSR: offsets 4(n+1), {dst,residual} 8S, reverse 4S, initial residual 4S with S = 2m;
  // heights, excesses 8n, four worklists 16n, degree class n, terminals and control 256
  return 4 * ((int64_t)n + 1) + 16 * (2 * m) + 25 * (int64_t)n + 256;
}

static void run_launch(mf_server *srv, Launch *L) {
  std::vector<mf_csr> views;
  views.reserve(L->reqs.size());
  for (auto &r : L->reqs) views.push_back(r->view());
  std::vector<int32_t> off, dst, cap, src, snk, vbase;
  mf_csr packed;
  std::vector<int64_t> values(L->reqs.size());
  mf_result res{};
  mf_status stt =
      pack_instances(views.data(), (int32_t)views.size(), off, dst, cap, src, snk, vbase, packed);
  if (stt == MF_OK) {
    mf_options o;
    mf_options_default(&o);
    const bool solo = L->reqs.size() == 1 && L->reqs[0]->solo;
    // batched: thread-centric cycles inside each workgroup, relabel every cycle-time ~ BFS time;
    // solo: the single-instance grid-wide solver, vertex-centric tiles (any degree mix)
    o.batch_mode = solo ? 1 : 0;
    o.relabel_time_ratio = solo ? 0.0f : 1.0f;
    stt = batch_solve_packed(srv->ctxs[(size_t)L->ctx_slot], &packed, src.data(), snk.data(),
                             vbase.data(), (int32_t)views.size(),
                             solo ? MF_ALGO_VC : MF_ALGO_TC, MF_LAYOUT_BCSR, o, values.data(),
                             &res);
  }
  std::lock_guard<std::mutex> lk(srv->mu);
  for (size_t i = 0; i < L->reqs.size(); ++i) {
    L->reqs[i]->state = stt == MF_OK ? 2 : 3;
    L->reqs[i]->value = values[i];
    L->reqs[i]->device_ms = res.solve_ms;
  }
  L->status = stt;
  L->finished = true;
  srv->cv.notify_all();
}

// Joins finished launches and releases their budget (lock held).
static void retire(mf_server *srv) {
  for (size_t i = 0; i < srv->inflight.size();) {
    Launch *L = srv->inflight[i].get();
    if (!L->finished) {
      ++i;
      continue;
    }
    srv->mu.unlock();
    L->worker.join();
    srv->mu.lock();
    srv->in_use -= L->bytes;
    srv->ctx_busy[(size_t)L->ctx_slot] = false;
    srv->st.completed += (int64_t)L->reqs.size();
    srv->inflight.erase(srv->inflight.begin() + (long)i);
  }
}

// Starts launches for pending requests in arrival order while budget and slots allow.
static void admit(mf_server *srv) {
  while ((int32_t)srv->inflight.size() < srv->max_inflight && !srv->pending.empty()) {
    int slot = -1;
    for (size_t c = 0; c < srv->ctx_busy.size(); ++c)
      if (!srv->ctx_busy[c]) {
        slot = (int)c;
        break;
      }
    if (slot < 0) break;
    std::unique_ptr<Launch> L(new Launch());
    int64_t free_bytes = srv->budget - srv->in_use;
    // in arrival order: a large request alone, or a run of small ones up to the next large one
    while (!srv->pending.empty() && srv->pending.front()->footprint <= free_bytes) {
      auto r = srv->pending.front();
      if (r->solo && !L->reqs.empty()) break;
      srv->pending.pop_front();
      free_bytes -= r->footprint;
      L->bytes += r->footprint;
      r->state = 1;
      L->reqs.push_back(r);
      if (r->solo) break;
    }
    if (L->reqs.empty()) break;
    srv->in_use += L->bytes;
    if (srv->in_use > srv->peak) srv->peak = srv->in_use;
    srv->st.admitted += (int64_t)L->reqs.size();
    srv->st.launches++;
    srv->ctx_busy[(size_t)slot] = true;
    L->ctx_slot = slot;
    Launch *raw = L.get();
    srv->inflight.push_back(std::move(L));
    raw->worker = std::thread(run_launch, srv, raw);
  }
}

extern "C" {

mf_status mf_server_create(mf_context *ctx, int64_t hbm_budget_bytes, int32_t max_inflight_launches,
                           mf_server **out) {
  if (!ctx || !out || hbm_budget_bytes <= 0 || max_inflight_launches <= 0 ||
      max_inflight_launches > 16)
    return MF_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  int32_t ncu = 0;
  mf_device_info(ctx, &ncu, nullptr, 0);
  std::unique_ptr<mf_server> srv(new mf_server());
  srv->owner = ctx;
  srv->budget = hbm_budget_bytes;
  srv->max_inflight = max_inflight_launches;
  for (int32_t i = 0; i < max_inflight_launches; ++i) {
    mf_context *c = nullptr;
    mf_status st = mf_context_create(0, &c);
    if (st != MF_OK) {
      for (mf_context *p : srv->ctxs) mf_context_destroy(p);
      return st;
    }
    srv->ctxs.push_back(c);
    srv->ctx_busy.push_back(false);
  }
  *out = srv.release();
  return MF_OK;
}

mf_status mf_server_submit(mf_server *srv, const mf_csr *g, int64_t *request_id) {
  if (!srv || !g || !request_id) return MF_ERR_INVALID_ARGUMENT;
  mf_status st = validate_csr(g);
  if (st != MF_OK) return st;
  auto r = std::make_shared<Request>();
  r->n = g->n;
  r->m = g->m;
  r->s = g->s;
  r->t = g->t;
  r->off.assign(g->offsets, g->offsets + g->n + 1);
  r->dst.assign(g->dst, g->dst + g->m);
  r->cap.resize((size_t)g->m);
  for (int64_t i = 0; i < g->m; ++i) r->cap[(size_t)i] = g->cap ? g->cap[i] : 1;
  r->footprint = mf_footprint_bytes(g->n, g->m);
  r->solo = g->n > kSoloVertices || g->m > kSoloArcs;
  std::lock_guard<std::mutex> lk(srv->mu);
  srv->st.submitted++;
  if (r->footprint > srv->budget) {
    srv->st.rejected++;
    return MF_ERR_BUDGET;
  }
  r->id = srv->next_id++;
  *request_id = r->id;
  srv->all[r->id] = r;
  srv->pending.push_back(r);
  return MF_OK;
}

mf_status mf_server_progress(mf_server *srv, int32_t wait) {
  if (!srv) return MF_ERR_INVALID_ARGUMENT;
  std::unique_lock<std::mutex> lk(srv->mu);
  retire(srv);
  admit(srv);
  if (wait && !srv->inflight.empty()) {
    const int64_t done0 = srv->st.completed;
    for (;;) {
      bool any = false;
      for (auto &L : srv->inflight) any |= L->finished;
      if (any) break;
      srv->cv.wait(lk);
    }
    retire(srv);
    admit(srv);
    (void)done0;
  }
  return MF_OK;
}

mf_status mf_server_result(mf_server *srv, int64_t request_id, int64_t *value, double *device_ms) {
  if (!srv) return MF_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lk(srv->mu);
  auto it = srv->all.find(request_id);
  if (it == srv->all.end()) return MF_ERR_INVALID_ARGUMENT;
  const Request &r = *it->second;
  if (r.state == 3) return MF_ERR_INTERNAL;
  if (r.state != 2) return MF_ERR_NOT_READY;
  if (value) *value = r.value;
  if (device_ms) *device_ms = r.device_ms;
  return MF_OK;
}

mf_status mf_server_stats_get(mf_server *srv, mf_server_stats *out) {
  if (!srv || !out) return MF_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lk(srv->mu);
  *out = srv->st;
  out->arena_bytes = srv->in_use;
  out->arena_peak_bytes = srv->peak;
  return MF_OK;
}

mf_status mf_server_destroy(mf_server *srv) {
  if (!srv) return MF_ERR_INVALID_ARGUMENT;
  {
    std::unique_lock<std::mutex> lk(srv->mu);
    srv->pending.clear();
    while (!srv->inflight.empty()) {
      bool any = false;
      for (auto &L : srv->inflight) any |= L->finished;
      if (!any) srv->cv.wait(lk);
      retire(srv);
    }
  }
  for (mf_context *c : srv->ctxs) mf_context_destroy(c);
  delete srv;
  return MF_OK;
}

}  // extern "C"
