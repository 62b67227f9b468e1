// Serving loop with continuous batching.
//
// Requests are copied at submit and take one of two paths:
//   - batched (at most kSoloVertices vertices and kSoloArcs arcs, and small enough for the
//     arena): the request's residual slots go into ranges of one device arena, allocated once
//     when the server is created, and the request is published into the queue of a persistent
//     serving kernel (k_serve_wg) that keeps running while requests arrive.  A workgroup of the
//     running kernel takes it as soon as one is free, so requests that arrive while others are on
//     the GPU join the running batch; their ranges are freed and reused as they finish.
//   - alone (larger requests): one workgroup would take tens of milliseconds on them, so each
//     runs on the whole grid (the single-instance solver) from a worker thread with its own
//     context and stream, overlapping on the device with the serving kernel and other large
//     requests.  Its device memory is allocated per request and charged to the budget.
// The budget (hbm_budget_bytes) is split: half for the arena, half for the large requests in
// flight.  Admission goes through the pending requests in arrival order and admits every one
// that fits now, so a request waiting for memory does not hold up smaller ones behind it.
//
// The mf_server_* calls may come from several threads: every call takes the server's lock, and
// worker threads are joined only after their launch has been taken out of the in-flight list.
#include <algorithm>
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

// Requests above either size run alone on the whole grid.
constexpr int32_t kSoloVertices = 16384;
constexpr int64_t kSoloArcs = 262144;
constexpr int32_t kQueueCap = 4096;    // entries of the serving kernel's queue ring
constexpr int64_t kSlotsPerVertex = 10;  // arena slot capacity per vertex (S = 2m ~ 10n typical)

enum State { PENDING = 0, COPYING, PUBLISHED, RUNNING_ALONE, DONE, FAILED };

double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

struct Request {
  int64_t id = 0;
  State state = PENDING;
  bool batched = false;
  int64_t footprint = 0;  // budget bytes charged while admitted
  // batched path
  ServeInstance inst;
  int32_t vb = -1;
  int64_t sb = -1;
  void *copy_done = nullptr;
  int64_t q = -1;  // queue entry
  // alone path: the request's own CSR
  std::vector<int32_t> off, dst, cap;
  int32_t n = 0, s = 0, t = 0;
  int64_t m = 0;
  // result
  int64_t value = 0;
  double device_ms = 0;
  double t_submit = 0, t_admit = 0, t_start = 0, t_finish = 0;
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

// A large request solved alone on the whole grid by a worker thread.
struct Launch {
  std::shared_ptr<Request> req;
  int ctx_slot = -1;
  std::thread worker;
  bool finished = false;
  mf_status status = MF_OK;
};

// First-fit allocator of [0, cap) ranges, coalescing on free.
struct Ranges {
  std::map<int64_t, int64_t> free_;  // begin -> length
  explicit Ranges(int64_t cap = 0) {
    if (cap > 0) free_[0] = cap;
  }
  int64_t alloc(int64_t len) {
    if (len <= 0) len = 1;
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < len) continue;
      const int64_t b = it->first, l = it->second;
      free_.erase(it);
      if (l > len) free_[b + len] = l - len;
      return b;
    }
    return -1;
  }
  void release(int64_t b, int64_t len) {
    if (len <= 0) len = 1;
    auto it = free_.emplace(b, len).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
};

}  // namespace

struct mf_server {
  mf_context *owner = nullptr;
  int device = 0;
  double t0 = 0;
  int64_t budget = 0;
  int64_t arena_budget = 0, alone_budget = 0, alone_in_use = 0;
  int64_t in_use = 0, peak = 0;
  int32_t max_inflight = 1;
  // batched path
  ServeArena *arena = nullptr;
  int64_t arena_nv = 0, arena_ns = 0;
  Ranges vranges, sranges;
  int64_t next_q = 0;  // next queue entry
  std::deque<std::shared_ptr<Request>> copying, published;
  // alone path
  std::vector<mf_context *> ctxs;
  std::vector<bool> ctx_busy;
  std::vector<std::unique_ptr<Launch>> inflight;
  // all requests
  std::deque<std::shared_ptr<Request>> pending;
  std::map<int64_t, std::shared_ptr<Request>> all;
  std::vector<double> latencies;  // finish - submit of completed requests, ms
  std::deque<int64_t> completed;   // ids not yet handed out by mf_server_completed
  double first_submit = -1, last_finish = 0;
  int64_t next_id = 1;
  mf_server_stats st{};
  std::mutex mu;
  std::condition_variable cv;
};

extern "C" int64_t mf_footprint_bytes(int32_t n, int64_t m) {
  // Device memory of one graph solved alone (BCSR, S = 2m slots), the larger of
  //   the upload's peak: the slot layout ({dst, residual} 8S, reverse 4S, capacities 4S, forward
  //   slot of every arc 4m, offsets 4n) plus the raw CSR copy (8m + 4n) and the device build's
  //   arrays (five arc arrays 20m, in-offsets 4n, radix-sort buffers 8m): 72m + 12n
  // and what stays resident: the layout 36m + 4n, plus heights, excesses, stamps, four worklists
  // and the degree class, 29n; the bottom-up levels' two lists of unreached vertices, 8n; and the
  // map of changed residual lines of graphs of 2^24 slots or more, m/4;
  // plus control words, statistics and terminals (64 KiB).
  const int64_t peak = 72 * m + 12 * (int64_t)n;
  const int64_t resident = 36 * m + m / 4 + 41 * (int64_t)n;
  return std::max(peak, resident) + (64 << 10);
}

static int64_t arena_footprint(int32_t n, int64_t S) { return serve_arena_bytes((int64_t)n + 1, S); }

static void finish(mf_server *srv, Request &r, State s) {
  r.state = s;
  r.t_finish = now_ms() - srv->t0;
  srv->st.completed++;
  srv->completed.push_back(r.id);
  if (s == DONE) {
    srv->latencies.push_back(r.t_finish - r.t_submit);
    srv->last_finish = std::max(srv->last_finish, r.t_finish);
    if (r.batched) srv->st.batched++;
    else srv->st.solo++;
  }
}

static void run_alone(mf_server *srv, Launch *L) {
  Request &r = *L->req;
  const mf_csr c = r.view();
  mf_options o;
  mf_options_default(&o);
  mf_result res{};
  int64_t value = 0;
  const int32_t src = r.s, snk = r.t, vbase[2] = {0, r.n};
  o.batch_mode = 1;  // the grid-wide single-instance solver
  const mf_status stt = batch_solve_packed(srv->ctxs[(size_t)L->ctx_slot], &c, &src, &snk, vbase,
                                           1, MF_ALGO_VC, MF_LAYOUT_BCSR, o, &value, &res);
  std::lock_guard<std::mutex> lk(srv->mu);
  r.value = value;
  r.device_ms = res.solve_ms;
  L->status = stt;
  L->finished = true;
  srv->cv.notify_all();
}

// Takes finished alone-launches out of the in-flight list (lock held), joins them without the
// lock, then releases their budget.
static void retire_alone(mf_server *srv, std::unique_lock<std::mutex> &lk) {
  std::vector<std::unique_ptr<Launch>> done;
  for (size_t i = 0; i < srv->inflight.size();) {
    if (srv->inflight[i]->finished) {
      done.push_back(std::move(srv->inflight[i]));
      srv->inflight.erase(srv->inflight.begin() + (long)i);
    } else {
      ++i;
    }
  }
  if (done.empty()) return;
  lk.unlock();
  for (auto &L : done) L->worker.join();
  lk.lock();
  for (auto &L : done) {
    Request &r = *L->req;
    srv->alone_in_use -= r.footprint;
    srv->in_use -= r.footprint;
    srv->ctx_busy[(size_t)L->ctx_slot] = false;
    finish(srv, r, L->status == MF_OK ? DONE : FAILED);
  }
}

// Reports of the serving kernel (lock held): finished requests free their arena ranges.
static void poll_batched(mf_server *srv) {
  while (!srv->published.empty()) {
    // entries are reported in any order; scan the published ones
    bool any = false;
    for (size_t i = 0; i < srv->published.size();) {
      Request &r = *srv->published[i];
      int64_t value = 0;
      int32_t status = 0;
      double dms = 0;
      if (!serve_poll(srv->arena, r.q, &value, &status, &dms)) {
        ++i;
        continue;
      }
      r.value = value;
      r.device_ms = dms;
      srv->vranges.release(r.vb, (int64_t)r.inst.n + 1);
      srv->sranges.release(r.sb, r.inst.S);
      srv->in_use -= r.footprint;
      finish(srv, r, status == 0 ? DONE : FAILED);
      r.inst = ServeInstance();  // the host copy of its layout is no longer needed
      srv->published.erase(srv->published.begin() + (long)i);
      any = true;
    }
    if (!any) break;
  }
}

// Copies that finished are published into the running batch (lock held).
static void publish_copied(mf_server *srv) {
  while (!srv->copying.empty()) {
    std::shared_ptr<Request> rp = srv->copying.front();
    Request &r = *rp;
    const int cs = serve_copy_state(r.copy_done);
    if (cs == 0) break;  // copies finish in order on the arena's copy stream
    serve_copy_release(r.copy_done);
    r.copy_done = nullptr;
    r.inst.a_off.clear();
    r.inst.a_dc.clear();
    r.inst.a_rev.clear();
    srv->copying.pop_front();
    if (cs < 0) {
      srv->vranges.release(r.vb, (int64_t)r.inst.n + 1);
      srv->sranges.release(r.sb, r.inst.S);
      srv->in_use -= r.footprint;
      finish(srv, r, FAILED);
      continue;
    }
    r.q = srv->next_q++;
    r.state = PUBLISHED;
    r.t_start = now_ms() - srv->t0;
    serve_publish(srv->arena, r.q, r.inst, r.vb);
    srv->published.push_back(rp);
  }
}

// Admits pending requests in arrival order, every one that fits now (lock held).
static void admit(mf_server *srv) {
  for (size_t i = 0; i < srv->pending.size();) {
    std::shared_ptr<Request> rp = srv->pending[i];
    Request &r = *rp;
    bool ok = false;
    if (r.batched) {
      const int64_t unreported = (int64_t)(srv->copying.size() + srv->published.size());
      if (unreported < kQueueCap) {
        const int64_t vb = srv->vranges.alloc((int64_t)r.inst.n + 1);
        const int64_t sb = vb < 0 ? -1 : srv->sranges.alloc(r.inst.S);
        if (vb >= 0 && sb < 0) srv->vranges.release(vb, (int64_t)r.inst.n + 1);
        if (vb >= 0 && sb >= 0) {
          r.vb = (int32_t)vb;
          r.sb = sb;
          if (serve_copy(srv->arena, r.inst, r.vb, r.sb, &r.copy_done) != MF_OK) {
            srv->vranges.release(vb, (int64_t)r.inst.n + 1);
            srv->sranges.release(sb, r.inst.S);
            srv->pending.erase(srv->pending.begin() + (long)i);
            finish(srv, r, FAILED);
            continue;
          }
          r.state = COPYING;
          r.t_admit = now_ms() - srv->t0;
          srv->copying.push_back(rp);
          ok = true;
        }
      }
    } else if (srv->alone_in_use + r.footprint <= srv->alone_budget) {
      int slot = -1;
      for (size_t c = 0; c < srv->ctx_busy.size(); ++c)
        if (!srv->ctx_busy[c]) {
          slot = (int)c;
          break;
        }
      if (slot >= 0) {
        srv->alone_in_use += r.footprint;
        srv->ctx_busy[(size_t)slot] = true;
        std::unique_ptr<Launch> L(new Launch());
        L->req = rp;
        L->ctx_slot = slot;
        r.state = RUNNING_ALONE;
        r.t_admit = r.t_start = now_ms() - srv->t0;
        Launch *raw = L.get();
        srv->inflight.push_back(std::move(L));
        srv->st.launches++;
        raw->worker = std::thread(run_alone, srv, raw);
        ok = true;
      }
    }
    if (!ok) {
      ++i;
      continue;
    }
    srv->in_use += r.footprint;
    srv->peak = std::max(srv->peak, srv->in_use);
    srv->st.admitted++;
    srv->pending.erase(srv->pending.begin() + (long)i);
  }
}

// One turn of the loop (lock held).
static mf_status step(mf_server *srv, std::unique_lock<std::mutex> &lk) {
  retire_alone(srv, lk);
  poll_batched(srv);
  admit(srv);
  publish_copied(srv);
  // published entries that no running workgroup will take: (re)launch the serving kernel.  Once
  // the grid has left, every report it wrote is visible: collect them first, so an entry that
  // was solved while this step admitted or copied is not mistaken for one left unclaimed
  const bool running = serve_running(srv->arena);
  if (!running) poll_batched(srv);
  if (!srv->published.empty() && !running) {
    const mf_status s = serve_launch(srv->arena);
    if (s != MF_OK) return s;
    srv->st.launches++;
    srv->st.serve_kernel_launches++;
  }
  return MF_OK;
}

static bool busy(const mf_server *srv) {
  return !srv->pending.empty() || !srv->copying.empty() || !srv->published.empty() ||
         !srv->inflight.empty();
}

extern "C" {

mf_status mf_server_create(mf_context *ctx, int64_t hbm_budget_bytes, int32_t max_inflight_launches,
                           mf_server **out) {
  if (!ctx || !out || hbm_budget_bytes <= 0 || max_inflight_launches <= 0 ||
      max_inflight_launches > 16)
    return MF_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  std::unique_ptr<mf_server> srv(new mf_server());
  srv->owner = ctx;
  srv->device = context_device(ctx);
  srv->t0 = now_ms();
  srv->budget = hbm_budget_bytes;
  srv->max_inflight = max_inflight_launches;
  srv->arena_budget = hbm_budget_bytes / 2;
  srv->alone_budget = hbm_budget_bytes - srv->arena_budget;
  // arena: nv vertices and kSlotsPerVertex * nv slots within its half of the budget
  srv->arena_nv = srv->arena_budget / serve_arena_bytes(1, kSlotsPerVertex);
  srv->arena_nv = std::min<int64_t>(srv->arena_nv, (INT32_MAX / 4) - 1);
  srv->arena_ns = std::min<int64_t>(srv->arena_nv * kSlotsPerVertex, INT32_MAX - 1);
  if (srv->arena_nv < 2) return MF_ERR_BUDGET;
  mf_status s = serve_arena_create(ctx, srv->arena_nv, srv->arena_ns, kQueueCap, &srv->arena);
  if (s != MF_OK) return s;
  srv->vranges = Ranges(srv->arena_nv);
  srv->sranges = Ranges(srv->arena_ns);
  for (int32_t i = 0; i < max_inflight_launches; ++i) {
    mf_context *c = nullptr;
    s = mf_context_create(srv->device, &c);
    if (s != MF_OK) {
      for (mf_context *p : srv->ctxs) mf_context_destroy(p);
      serve_arena_destroy(srv->arena);
      return s;
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
  const double t = now_ms();
  const bool small = g->n <= kSoloVertices && g->m <= kSoloArcs &&
                     (int64_t)g->n + 1 <= srv->arena_nv && 2 * g->m <= srv->arena_ns;
  if (small) {
    // laid out here, in the caller's thread, so admission only copies it
    if ((st = serve_prepare(g, r->inst)) != MF_OK) return st;
    r->batched = true;
    r->footprint = arena_footprint(g->n, r->inst.S);
  } else {
    r->n = g->n;
    r->m = g->m;
    r->s = g->s;
    r->t = g->t;
    r->off.assign(g->offsets, g->offsets + g->n + 1);
    r->dst.assign(g->dst, g->dst + g->m);
    r->cap.resize((size_t)g->m);
    for (int64_t i = 0; i < g->m; ++i) r->cap[(size_t)i] = g->cap ? g->cap[i] : 1;
    r->footprint = mf_footprint_bytes(g->n, g->m);
  }
  std::lock_guard<std::mutex> lk(srv->mu);
  srv->st.submitted++;
  if (!small && r->footprint > srv->alone_budget) {
    srv->st.rejected++;
    return MF_ERR_BUDGET;
  }
  r->id = srv->next_id++;
  r->t_submit = t - srv->t0;
  if (srv->first_submit < 0) srv->first_submit = r->t_submit;
  *request_id = r->id;
  srv->all[r->id] = r;
  srv->pending.push_back(r);
  return MF_OK;
}

mf_status mf_server_progress(mf_server *srv, int32_t wait) {
  if (!srv) return MF_ERR_INVALID_ARGUMENT;
  std::unique_lock<std::mutex> lk(srv->mu);
  const int64_t done0 = srv->st.completed;
  for (;;) {
    const mf_status s = step(srv, lk);
    if (s != MF_OK) return s;
    if (!wait || srv->st.completed > done0 || !busy(srv)) return MF_OK;
    // wait for a report, a finished copy or a finished alone-launch
    srv->cv.wait_for(lk, std::chrono::microseconds(100));
  }
}

mf_status mf_server_result(mf_server *srv, int64_t request_id, int64_t *value, double *device_ms) {
  if (!srv) return MF_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lk(srv->mu);
  auto it = srv->all.find(request_id);
  if (it == srv->all.end()) return MF_ERR_INVALID_ARGUMENT;
  const Request &r = *it->second;
  if (r.state == FAILED) return MF_ERR_INTERNAL;
  if (r.state != DONE) return MF_ERR_NOT_READY;
  if (value) *value = r.value;
  if (device_ms) *device_ms = r.device_ms;
  return MF_OK;
}

mf_status mf_server_completed(mf_server *srv, int64_t *ids, int32_t cap, int32_t *count) {
  if (!srv || !count || cap < 0 || (cap > 0 && !ids)) return MF_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lk(srv->mu);
  int32_t k = 0;
  while (k < cap && !srv->completed.empty()) {
    ids[k++] = srv->completed.front();
    srv->completed.pop_front();
  }
  *count = k;
  return MF_OK;
}

mf_status mf_server_request_times(mf_server *srv, int64_t request_id, mf_server_times *out) {
  if (!srv || !out) return MF_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lk(srv->mu);
  auto it = srv->all.find(request_id);
  if (it == srv->all.end()) return MF_ERR_INVALID_ARGUMENT;
  const Request &r = *it->second;
  if (r.state == FAILED) return MF_ERR_INTERNAL;
  if (r.state != DONE) return MF_ERR_NOT_READY;
  out->submit_ms = r.t_submit;
  out->admit_ms = r.t_admit;
  out->start_ms = r.t_start;
  out->finish_ms = r.t_finish;
  out->device_ms = r.device_ms;
  out->batched = r.batched ? 1 : 0;
  out->reserved = 0;
  return MF_OK;
}

mf_status mf_server_latency(mf_server *srv, double *p50_ms, double *p99_ms, double *solves_per_s) {
  if (!srv) return MF_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lk(srv->mu);
  std::vector<double> v = srv->latencies;
  if (v.empty()) return MF_ERR_NOT_READY;
  std::sort(v.begin(), v.end());
  // nearest rank above the percentile, interpolated linearly like numpy's default
  auto pct = [&](double p) {
    const double x = p / 100.0 * (double)(v.size() - 1);
    const size_t lo = (size_t)x;
    const size_t hi = std::min(lo + 1, v.size() - 1);
    return v[lo] + (v[hi] - v[lo]) * (x - (double)lo);
  };
  if (p50_ms) *p50_ms = pct(50);
  if (p99_ms) *p99_ms = pct(99);
  if (solves_per_s) {
    const double span = srv->last_finish - srv->first_submit;
    *solves_per_s = span > 0 ? 1e3 * (double)v.size() / span : 0;
  }
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
    // finish what is on the GPU: copies in flight are published, published entries solved
    while (!srv->inflight.empty() || !srv->copying.empty() || !srv->published.empty()) {
      if (step(srv, lk) != MF_OK) break;
      srv->cv.wait_for(lk, std::chrono::microseconds(200));
    }
    retire_alone(srv, lk);
  }
  serve_arena_destroy(srv->arena);
  for (mf_context *c : srv->ctxs) mf_context_destroy(c);
  delete srv;
  return MF_OK;
}

}  // extern "C"
