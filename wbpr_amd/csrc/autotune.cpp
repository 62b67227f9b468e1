// Autotuner of the solver's launch parameters: trial solves over a small candidate set, by
// coordinate descent (one parameter at a time, keeping the best so far).  The parameters
// interact (tile widths, BFS hand-over and graph replay most), so the descent sweeps them all
// at least twice and again while a sweep still changes the options, at most kPasses times.
//
//   tile widths       mf_options.class_lanes: lanes per vertex of each VC degree class.  Narrow
//                     tiles put more vertices in flight per wave, wide tiles more slots per
//                     vertex; the best split depends on the degree distribution.
//   relabel cadence   mf_options.relabel_time_ratio: cycle time per global-relabel time.  Graphs
//                     that strand excess want frequent relabels, long-path graphs fewer.
//                     mf_options.cycles_per_relabel (single instance): a fixed number of cycles
//                     per relabel instead, at 1, 1.5 and 2 times the cycles per relabel round the
//                     starting options ran.  The adaptive cadence follows measured times, so it
//                     moves with the run-to-run spread of the cycles; a fixed one does not (Genrmf
//                     64: 6% less time on average and half the spread).
//   BFS hand-over     mf_options.bfs_block_frontier: frontier size up to which all BFS levels run
//                     inside one workgroup.  mf_options.bfs_cooperative (single instance): the
//                     levels of mid-size frontiers inside one cooperative launch or one launch
//                     per level (long-path graphs gain, power-law graphs lose).
//   forwarding        mf_options.forward_hops: vertices a tile discharges per cycle by following
//                     its own pushes (long augmenting paths want many, churning graphs few).
//   graph replay      mf_options.cycle_graphs (single instance): cycle chunks replayed from
//                     captured hipGraphs, cached on the resident graph, or launched one by one.
//   batch algorithm   mf_options.batch_algo (batched solves only): TC or VC cycles.
//
// Lock-free schedules differ from run to run (solve times spread by +-15% on long-path graphs),
// so a candidate's time is the mean of three solves, the starting options' the mean of six, and
// the current best's the mean of all its solves (it is timed again before every further sweep,
// so a lucky low mean does not stop the search).  The winner is checked at the end: it
// and the starting options are timed again over six solves each, alternating between the two so
// that a drift of solve times during the run favours neither, and the faster of the two is
// returned -- tuning never hands back options slower than those it started from.  Every trial
// must give the same max-flow value (the value is unique); a mismatch is reported as
// MF_ERR_INTERNAL.
#include <algorithm>
#include <cstring>
#include <numeric>
#include <vector>

#include "../../include/mf.h"
#include "solver.h"

namespace {

constexpr int kRepeats = 3;   // solves per candidate (mean)
constexpr int kValidate = 6;  // solves per option set in the final check (alternating)
constexpr int kPasses = 3;    // sweeps over all parameters: two, then more while one changes something

// VC tile-width presets (lanes for degree classes <=8, <=16, <=32, <=1024, larger)
const uint16_t kLanePresets[][5] = {
    {8, 16, 32, 64, 256},  // default: one slot per lane per class bound, 4 in flight
    {2, 4, 8, 32, 256},    // narrow: four slots in flight per lane cover each class bound
    {4, 8, 16, 64, 256},   // middle
    {1, 2, 8, 32, 256},    // low-degree vertices thread-centric
    {4, 4, 8, 16, 256},    // narrow up to 1,024 slots (long-path graphs of low degree)
};
const float kRatios[] = {0.5f, 1.0f, 2.0f, 4.0f, 8.0f};
const int32_t kForwardHops[] = {-1, 2, 8, 32};
const int32_t kFrontiers[] = {256, 1024, 4096, 16384};
const double kCadenceFactors[] = {1.0, 1.5, 2.0};

float effective_ratio(const mf_options &o) {
  return o.relabel_time_ratio > 0 ? o.relabel_time_ratio : 2.0f;
}

int32_t effective_frontier(const mf_options &o) {
  return o.bfs_block_frontier > 0 ? o.bfs_block_frontier : 4096;
}

bool same_lanes(const uint16_t *a, const uint16_t *b) {
  static const uint16_t def[5] = {8, 16, 32, 64, 256};
  for (int k = 0; k < 5; ++k) {
    const uint16_t x = a[k] ? a[k] : def[k], y = b[k] ? b[k] : def[k];
    if (x != y) return false;
  }
  return true;
}

// One sweep over every parameter: each candidate value is tried with the others as they stand
// in cur, which consider(t) replaces by t when t is faster (cur is the caller's live value).
// cadence_base: cycles per global-relabel round measured with the starting options (0: none).
template <class Consider>
mf_status sweep(const mf_options &cur, bool tune_lanes, bool tune_batch_algo,
                double cadence_base, Consider consider) {
  mf_status st = MF_OK;
  if (tune_batch_algo) {
    mf_options t = cur;
    t.batch_algo = cur.batch_algo == 0 ? 1 : 0;
    if ((st = consider(t)) != MF_OK) return st;
  }
  if (tune_lanes && (!tune_batch_algo || cur.batch_algo == 1)) {
    for (const auto &p : kLanePresets) {
      if (same_lanes(p, cur.class_lanes)) continue;
      mf_options t = cur;
      std::memcpy(t.class_lanes, p, sizeof(t.class_lanes));
      if ((st = consider(t)) != MF_OK) return st;
    }
  }
  if (!tune_batch_algo) {  // single instance: cycle chunks as captured graphs or not
    mf_options t = cur;
    t.cycle_graphs = cur.cycle_graphs ? 0 : 1;
    if ((st = consider(t)) != MF_OK) return st;
  }
  const int32_t h0 = cur.forward_hops;
  for (int32_t h : kForwardHops) {
    if (h == h0) continue;
    mf_options t = cur;
    t.forward_hops = h;
    if ((st = consider(t)) != MF_OK) return st;
  }
  if (cur.cycles_per_relabel == 0) {  // the time ratio only steers the adaptive cadence
    const float r0 = effective_ratio(cur);
    for (float r : kRatios) {
      if (r == r0) continue;
      mf_options t = cur;
      t.relabel_time_ratio = r;
      if ((st = consider(t)) != MF_OK) return st;
    }
  }
  if (!tune_batch_algo && cadence_base > 0) {
    const int32_t k0 = cur.cycles_per_relabel;
    for (double f : kCadenceFactors) {
      const int32_t k = (int32_t)std::max(1.0, cadence_base * f + 0.5);
      if (k == k0) continue;
      mf_options t = cur;
      t.cycles_per_relabel = k;
      if ((st = consider(t)) != MF_OK) return st;
    }
    if (k0 != 0) {
      mf_options t = cur;
      t.cycles_per_relabel = 0;
      if ((st = consider(t)) != MF_OK) return st;
    }
  }
  if (!tune_batch_algo) {  // the batched kernel runs its BFS inside each workgroup anyway
    mf_options c = cur;
    c.bfs_cooperative = cur.bfs_cooperative ? 0 : 1;
    if ((st = consider(c)) != MF_OK) return st;
    const int32_t f0 = effective_frontier(cur);
    for (int32_t f : kFrontiers) {
      if (f == f0) continue;
      mf_options t = cur;
      t.bfs_block_frontier = f;
      if ((st = consider(t)) != MF_OK) return st;
    }
  }
  return MF_OK;
}

// Coordinate descent over the candidates; run(opts, ms, reps, cpr) times reps solves (mean ms)
// and, when cpr is not null, stores their mean cycles per global-relabel round there.
template <class Run>
mf_status descend(mf_options cur, bool tune_lanes, bool tune_batch_algo, Run run,
                  mf_options *best, double *best_ms) {
  // the result names every tuned value explicitly (0 would mean "library default")
  cur.relabel_time_ratio = effective_ratio(cur);
  cur.bfs_block_frontier = effective_frontier(cur);
  if (cur.forward_hops == 0) cur.forward_hops = 2;
  if (tune_lanes)
    for (int k = 0; k < 5; ++k)
      if (!cur.class_lanes[k]) cur.class_lanes[k] = kLanePresets[0][k];
  // cur_ms is the mean of every solve timed with cur so far: the options that won on a lucky
  // low mean are timed again before each further sweep, so their estimate settles back and a
  // faster candidate can still replace them
  double cur_ms = 0, cur_sum = 0;
  int cur_n = 0;
  auto time_cur = [&](int reps, double *cpr) -> mf_status {
    double ms = 0;
    const mf_status s = run(cur, ms, reps, cpr);
    if (s != MF_OK) return s;
    cur_sum += ms * reps;
    cur_n += reps;
    cur_ms = cur_sum / cur_n;
    return MF_OK;
  };
  double cadence_base = 0;
  mf_status st = time_cur(2 * kRepeats, &cadence_base);
  if (st != MF_OK) return st;
  const mf_options start = cur;
  bool changed = false;
  auto consider = [&](const mf_options &t) -> mf_status {
    double ms = 0;
    const mf_status s = run(t, ms, kRepeats, nullptr);
    if (s != MF_OK) return s;
    if (ms < cur_ms) {
      cur = t;
      cur_sum = ms * kRepeats;
      cur_n = kRepeats;
      cur_ms = ms;
      changed = true;
    }
    return MF_OK;
  };
  for (int pass = 0; pass < kPasses; ++pass) {
    changed = false;
    if (pass > 0 && (st = time_cur(kRepeats, nullptr)) != MF_OK) return st;
    // sweep() reads cur live: each parameter is tried on top of the values chosen before it
    if ((st = sweep(cur, tune_lanes, tune_batch_algo, cadence_base, consider)) != MF_OK)
      return st;
    if (!changed && pass > 0) break;  // a second sweep always runs, against a re-timed best
  }
  if (std::memcmp(&cur, &start, sizeof(cur)) != 0) {
    double a_ms = 0, b_ms = 0;
    for (int r = 0; r < kValidate; ++r) {
      double x = 0, y = 0;
      if ((st = run(start, x, 1, nullptr)) != MF_OK || (st = run(cur, y, 1, nullptr)) != MF_OK)
        return st;
      a_ms += x / kValidate;
      b_ms += y / kValidate;
    }
    if (a_ms <= b_ms) {
      cur = start;
      cur_ms = a_ms;
    } else {
      cur_ms = b_ms;
    }
  }
  *best = cur;
  if (best_ms) *best_ms = cur_ms;
  return MF_OK;
}

}  // namespace

extern "C" {

mf_status mf_autotune(mf_context *ctx, mf_graph *g, mf_algo algo, const mf_options *base,
                      mf_options *best, double *best_ms, int32_t *trials) {
  if (!ctx || !g || !best) return MF_ERR_INVALID_ARGUMENT;
  if (algo != MF_ALGO_TC && algo != MF_ALGO_VC) return MF_ERR_INVALID_ARGUMENT;
  mf_options o;
  mf_options_default(&o);
  if (base) o = *base;
  o.profile_kernels = 0;
  int64_t value = -1;
  int32_t n = 0;
  auto run = [&](const mf_options &t, double &ms, int reps, double *cpr) -> mf_status {
    ms = 0;
    double cycles = 0, rounds = 0;
    for (int r = 0; r < reps; ++r) {
      mf_result res;
      const mf_status st = mf_graph_solve(ctx, g, algo, &t, &res);
      if (st != MF_OK) return st;
      if (value >= 0 && res.value != value) return MF_ERR_INTERNAL;
      value = res.value;
      ms += res.solve_ms / reps;
      cycles += (double)res.cycles;
      rounds += (double)res.rounds;
      ++n;
    }
    if (cpr) *cpr = rounds > 0 ? cycles / rounds : 0;
    return MF_OK;
  };
  const mf_status st =
      descend(o, algo == MF_ALGO_VC && o.tile_lanes == 0, false, run, best, best_ms);
  if (trials) *trials = n;
  return st;
}

mf_status mf_batch_autotune(mf_context *ctx, const mf_csr *inst, int32_t k, int32_t sample,
                            const mf_options *base, mf_options *best, double *best_ms,
                            int32_t *trials) {
  if (!ctx || !inst || k <= 0 || !best) return MF_ERR_INVALID_ARGUMENT;
  if (sample <= 0) sample = 256;
  sample = std::min(sample, k);
  // the largest instances set the tail of a persistent launch: a quarter of the sample is the
  // largest ones, the rest an even spread over the others
  std::vector<int32_t> order((size_t)k);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
    return (int64_t)inst[a].n + inst[a].m > (int64_t)inst[b].n + inst[b].m;
  });
  const int32_t top = std::max(1, sample / 4);
  std::vector<mf_csr> sub;
  sub.reserve((size_t)sample);
  for (int32_t i = 0; i < top; ++i) sub.push_back(inst[order[(size_t)i]]);
  const int32_t rest = sample - top, pool = k - top;
  for (int32_t i = 0; i < rest && pool > 0; ++i)
    sub.push_back(inst[order[(size_t)top + (size_t)((int64_t)i * pool / rest)]]);
  mf_options o;
  mf_options_default(&o);
  if (base) o = *base;
  o.profile_kernels = 0;
  // the sample is packed and uploaded once; every trial solves it from the capacities again
  mf::PackedBatch packed;
  mf_status pst = mf::pack_instances(ctx, sub.data(), (int32_t)sub.size(), packed);
  if (pst != MF_OK) return pst;
  mf::BatchResident *resident = nullptr;
  pst = mf::batch_upload(ctx, &packed.view, packed.src.data(), packed.snk.data(),
                         packed.vbase.data(), (int32_t)sub.size(), o, &resident);
  if (pst != MF_OK) return pst;
  struct Release {
    mf::BatchResident *r;
    ~Release() { mf::batch_release(r); }
  } release{resident};
  std::vector<int64_t> ref, vals(sub.size());
  int32_t n = 0;
  auto run = [&](const mf_options &t, double &ms, int reps, double *cpr) -> mf_status {
    ms = 0;
    if (cpr) *cpr = 0;  // the batched kernel keeps its own cadence per instance
    for (int r = 0; r < reps; ++r) {
      mf_result res;
      const mf_status st = mf::batch_solve_resident(ctx, resident, t, vals.data(), &res);
      if (st != MF_OK) return st;
      if (!ref.empty() && ref != vals) return MF_ERR_INTERNAL;
      ref = vals;
      ms += res.solve_ms / reps;
      ++n;
    }
    return MF_OK;
  };
  const mf_status st = descend(o, o.tile_lanes == 0, true, run, best, best_ms);
  if (trials) *trials = n;
  return st;
}

}  // extern "C"
