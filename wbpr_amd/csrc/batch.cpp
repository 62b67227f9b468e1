// Batched solve: k independent instances packed into one concatenated CSR with per-instance
// vertex offsets.  By default a persistent grid of workgroups pulls whole instances from a
// device queue (one workgroup solves one instance at a time); batch_mode 1 instead runs the
// grid-wide kernels over all instances at once.
#include <algorithm>
#include <climits>
#include <cstring>
#include <vector>

#include "mf_internal.h"
#include "solver.h"

namespace mf {

// Concatenates instances: vertex v of instance i becomes vbase[i] + v, arc a of instance i arc
// abase[i] + a.  The arrays go into the context's pinned staging buffer when it can provide one
// (so the upload is one DMA per array), else into out.own; instances are copied in parallel.
mf_status pack_instances(mf_context *ctx, const mf_csr *inst, int32_t k, PackedBatch &out) {
  int64_t N = 0, M = 0;
  std::vector<int64_t> abase((size_t)k + 1);
  out.vbase.resize((size_t)k + 1);
  out.src.resize((size_t)k);
  out.snk.resize((size_t)k);
  std::vector<mf_status> valid((size_t)k, MF_OK);
  parallel_for(
      k,
      [&](int64_t b, int64_t e) {
        for (int64_t i = b; i < e; ++i) valid[(size_t)i] = validate_csr(&inst[i]);
      },
      1);
  for (int32_t i = 0; i < k; ++i) {
    if (valid[(size_t)i] != MF_OK) return valid[(size_t)i];
    if (N + inst[i].n + 2 >= INT32_MAX || 2 * (M + inst[i].m) > INT32_MAX) return MF_ERR_OVERFLOW;
    out.vbase[(size_t)i] = (int32_t)N;
    abase[(size_t)i] = M;
    out.src[(size_t)i] = (int32_t)(N + inst[i].s);
    out.snk[(size_t)i] = (int32_t)(N + inst[i].t);
    N += inst[i].n;
    M += inst[i].m;
  }
  out.vbase[(size_t)k] = (int32_t)N;
  abase[(size_t)k] = M;
  const size_t words = (size_t)N + 1 + 2 * (size_t)M;
  int32_t *buf = ctx ? static_cast<int32_t *>(context_staging(ctx, sizeof(int32_t) * words)) : nullptr;
  if (!buf) {
    out.own.resize(words);
    buf = out.own.data();
  }
  out.off = buf;
  out.dst = buf + N + 1;
  out.cap = buf + N + 1 + M;
  parallel_for(
      k,
      [&](int64_t b, int64_t e) {
        for (int64_t i = b; i < e; ++i) {
          const mf_csr &g = inst[i];
          const int32_t vb = out.vbase[(size_t)i];
          const int64_t ab = abase[(size_t)i];
          for (int32_t v = 0; v < g.n; ++v) out.off[vb + v] = (int32_t)(ab + g.offsets[v]);
          int32_t *dd = out.dst + ab, *cc = out.cap + ab;
          for (int64_t a = 0; a < g.m; ++a) dd[a] = vb + g.dst[a];
          if (g.cap) memcpy(cc, g.cap, sizeof(int32_t) * (size_t)g.m);
          else std::fill(cc, cc + g.m, 1);
        }
      },
      1);
  out.off[N] = (int32_t)M;
  out.view.n = (int32_t)N;
  out.view.m = M;
  out.view.offsets = out.off;
  out.view.dst = out.dst;
  out.view.cap = out.cap;
  out.view.s = k ? out.src[0] : 0;
  out.view.t = k ? out.snk[0] : 0;
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
  PackedBatch packed;
  mf_status st = pack_instances(ctx, inst, k, packed);
  if (st != MF_OK) return st;
  const mf_algo algo = o.batch_algo == 1 ? MF_ALGO_VC : MF_ALGO_TC;
  return batch_solve_packed(ctx, &packed.view, packed.src.data(), packed.snk.data(),
                            packed.vbase.data(), k, algo, MF_LAYOUT_BCSR, o, values, total);
}
