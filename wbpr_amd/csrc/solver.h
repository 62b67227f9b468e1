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

// Concatenates instances: vertex v of instance i becomes vbase[i] + v (batch.cpp).
mf_status pack_instances(const mf_csr *inst, int32_t k, std::vector<int32_t> &off,
                         std::vector<int32_t> &dst, std::vector<int32_t> &cap,
                         std::vector<int32_t> &src, std::vector<int32_t> &snk,
                         std::vector<int32_t> &vbase, mf_csr &packed);

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
}  // namespace mf
