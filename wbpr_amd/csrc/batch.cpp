// Batched solve: k independent instances packed into one concatenated CSR with per-instance
// vertex offsets.  By default a persistent grid of workgroups pulls whole instances from a
// device queue (one workgroup solves one instance at a time); batch_mode 1 instead runs the
// grid-wide kernels over all instances at once.
#include <climits>
#include <vector>

#include "mf_internal.h"
#include "solver.h"

namespace mf {

// Concatenates instances: vertex v of instance i becomes vbase[i] + v.
mf_status pack_instances(const mf_csr *inst, int32_t k, std::vector<int32_t> &off,
                         std::vector<int32_t> &dst, std::vector<int32_t> &cap,
                         std::vector<int32_t> &src, std::vector<int32_t> &snk,
                         std::vector<int32_t> &vbase, mf_csr &packed) {
  int64_t N = 0, M = 0;
  for (int32_t i = 0; i < k; ++i) {
    mf_status st = validate_csr(&inst[i]);
    if (st != MF_OK) return st;
    N += inst[i].n;
    M += inst[i].m;
  }
  if (N + 2 >= INT32_MAX || 2 * M > INT32_MAX) return MF_ERR_OVERFLOW;
  off.resize((size_t)N + 1);
  dst.resize((size_t)M);
  cap.resize((size_t)M);
  src.resize((size_t)k);
  snk.resize((size_t)k);
  vbase.resize((size_t)k + 1);
  int64_t vb = 0, ab = 0;
  off[0] = 0;
  for (int32_t i = 0; i < k; ++i) {
    const mf_csr &g = inst[i];
    for (int32_t v = 0; v < g.n; ++v) off[(size_t)(vb + v + 1)] = (int32_t)(ab + g.offsets[v + 1]);
    for (int64_t a = 0; a < g.m; ++a) {
      dst[(size_t)(ab + a)] = (int32_t)(vb + g.dst[a]);
      cap[(size_t)(ab + a)] = g.cap ? g.cap[a] : 1;
    }
    vbase[(size_t)i] = (int32_t)vb;
    src[(size_t)i] = (int32_t)(vb + g.s);
    snk[(size_t)i] = (int32_t)(vb + g.t);
    vb += g.n;
    ab += g.m;
  }
  vbase[(size_t)k] = (int32_t)vb;
  packed.n = (int32_t)N;
  packed.m = M;
  packed.offsets = off.data();
  packed.dst = dst.data();
  packed.cap = cap.data();
  packed.s = src.empty() ? 0 : src[0];
  packed.t = snk.empty() ? 0 : snk[0];
  return MF_OK;
}

}  // namespace mf

using namespace mf;

extern "C" mf_status mf_batch_solve(mf_context *ctx, const mf_csr *inst, int32_t k,
                                    const mf_options *opts, int64_t *values, mf_result *total) {
  if (!ctx || !inst || k <= 0 || !values) return MF_ERR_INVALID_ARGUMENT;
  mf_options o;
  mf_options_default(&o);
  if (opts) o = *opts;
  std::vector<int32_t> off, dst, cap, src, snk, vbase;
  mf_csr packed;
  mf_status st = pack_instances(inst, k, off, dst, cap, src, snk, vbase, packed);
  if (st != MF_OK) return st;
  const mf_algo algo = o.batch_algo == 1 ? MF_ALGO_VC : MF_ALGO_TC;
  return batch_solve_packed(ctx, &packed, src.data(), snk.data(), vbase.data(), k, algo,
                            MF_LAYOUT_BCSR, o, values, total);
}
