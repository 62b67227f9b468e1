// Host-side declarations shared by the solver sources.
#pragma once

#include <cstdint>
#include <vector>

#include "../../include/mf.h"

namespace mf {

// Degree classes of the worklists: class k holds vertices with slot degree <= maxdeg[k] and is
// processed with lanes[k] lanes per vertex (1 = thread-centric, 256 = whole block, the last VC
// class = giants split over several blocks).  mf_options.class_lanes sets the first kTunedClasses.
constexpr int kClasses = 6;
constexpr int kTunedClasses = 5;
struct ClassConfig {
  int nclass = 1;
  int lanes[kClasses] = {1, 0, 0, 0, 0, 0};
  int64_t maxdeg[kClasses] = {0, 0, 0, 0, 0, 0};
};

ClassConfig class_config(mf_algo algo, int tile_lanes, const uint16_t *class_lanes = nullptr);
int class_of_degree(const ClassConfig &c, int64_t deg);

int64_t algorithmic_bytes_cycle(mf_layout layout, int64_t active, int64_t slots, int64_t pushes,
                                int64_t relabels);
int64_t algorithmic_bytes_bfs(mf_layout layout, int32_t n, int64_t rounds, int64_t verts,
                              int64_t slots, int64_t bu_verts, int64_t bu_slots);

const char *last_device_error();

}  // namespace mf

struct mf_context;
namespace mf {
// Solves k instances packed into one CSR (vertex ranges disjoint, src/snk global ids).
// vbase[k+1]: first vertex of every instance.
mf_status batch_solve_packed(mf_context *ctx, const mf_csr *packed, const int32_t *src,
                             const int32_t *snk, const int32_t *vbase, int32_t k, mf_algo algo,
                             mf_layout layout, const mf_options &o, int64_t *values,
                             mf_result *res);

// A packed batch on the host: the concatenated CSR (in the context's pinned staging buffer, or
// in own), the global terminals and the first vertex of every instance (vbase[k] = total).
struct PackedBatch {
  int32_t *off = nullptr, *dst = nullptr, *cap = nullptr;
  std::vector<int32_t> src, snk, vbase, own;
  mf_csr view{};
};
// Concatenates instances: vertex v of instance i becomes vbase[i] + v (batch.cpp).  The staging
// buffer is the context's: the packed arrays stay valid until the next call on that context that
// stages data.
mf_status pack_instances(mf_context *ctx, const mf_csr *inst, int32_t k, PackedBatch &out);
// Pinned host memory of at least `bytes`, owned by the context and reused by later calls (grown
// when needed); nullptr if it cannot be allocated.
void *context_staging(mf_context *ctx, size_t bytes);

// A packed batch uploaded once and solved again and again, with other options each time (the
// batched autotuner's trials): every solve restarts from the capacities on the device.
// batch_mode is fixed by the upload; the algorithm comes from each solve's batch_algo.
struct BatchResident;
mf_status batch_upload(mf_context *ctx, const mf_csr *packed, const int32_t *src,
                       const int32_t *snk, const int32_t *vbase, int32_t k,
                       const mf_options &o, BatchResident **out);
mf_status batch_solve_resident(mf_context *ctx, BatchResident *r, const mf_options &o,
                               int64_t *values, mf_result *res);
void batch_release(BatchResident *r);

// ---- serving arena: a device arena of vertex and slot ranges, and a persistent serving kernel
// that solves the instances the host publishes into a queue while it runs (k_serve_wg) ----
struct ServeArena;
// A request laid out for the arena on the host: BCSR slots with local ids (offsets n+1,
// {dst, capacity} pairs, reverse slots), vertex BFS classes and the class populations.  The
// a_* arrays hold the same data in arena ids while its copy is in flight.
struct ServeInstance {
  int32_t n = 0, s = 0, t = 0;
  int64_t S = 0;
  std::vector<int32_t> off, dc, rev, a_off, a_dc, a_rev;
  std::vector<unsigned char> vcls;
  int32_t pop[kClasses], bpop[kClasses];
};
mf_status serve_prepare(const mf_csr *g, ServeInstance &inst);
// Device bytes of an arena of nv vertices and ns slots (an instance takes n + 1 vertices and
// 2m slots).
int64_t serve_arena_bytes(int64_t nv, int64_t ns);
mf_status serve_arena_create(mf_context *ctx, int64_t nv, int64_t ns, int32_t qcap,
                             ServeArena **out);
// Leaves idle workgroups, waits for the kernel and the copies, frees everything.
void serve_arena_destroy(ServeArena *a);
// Queues the copies of inst into vertices [vb, vb + n] and slots [sb, sb + S) of the arena;
// *done is a ticket for serve_copy_state / serve_copy_release.  inst must stay alive (and its
// a_* arrays unchanged) until the copy has finished.
mf_status serve_copy(ServeArena *a, ServeInstance &inst, int32_t vb, int64_t sb, void **done);
int serve_copy_state(void *done);
void serve_copy_release(void *done);
// Publishes queue entry q (q = 0, 1, 2, ... in order; at most qcap unreported entries).
void serve_publish(ServeArena *a, int64_t q, const ServeInstance &inst, int32_t vb);
// True once entry q is reported: its value, status (0 ok, 1 round limit, 2 invariant broken)
// and the device time of its solve.
bool serve_poll(ServeArena *a, int64_t q, int64_t *value, int32_t *status, double *device_ms);
// Whether the serving kernel is still running; launching it (again) when it is not.
bool serve_running(ServeArena *a);
mf_status serve_launch(ServeArena *a);
// The device a context runs on.
int context_device(const mf_context *ctx);
int64_t serve_launches(const ServeArena *a);
}  // namespace mf
