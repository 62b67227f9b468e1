// Internal declarations shared by the host-side sources of libmf.
#pragma once

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/mf.h"

namespace mf {

// Runs f(begin, end) over [0, count) split into contiguous chunks, one per host thread (at most
// 16, and none for small counts: below min_chunk items per thread the caller's thread does all).
template <class F>
void parallel_for(int64_t count, F f, int64_t min_chunk = 1 << 16) {
  const int64_t hw = std::max<int64_t>(1, std::min<int64_t>(16, std::thread::hardware_concurrency()));
  const int64_t nt = std::max<int64_t>(1, std::min<int64_t>(hw, count / std::max<int64_t>(min_chunk, 1)));
  if (nt <= 1) {
    if (count > 0) f((int64_t)0, count);
    return;
  }
  std::vector<std::thread> threads;
  threads.reserve((size_t)nt - 1);
  const int64_t step = (count + nt - 1) / nt;
  for (int64_t t = 1; t < nt; ++t) {
    const int64_t b = t * step, e = std::min(count, b + step);
    if (b < e) threads.emplace_back([=, &f] { f(b, e); });
  }
  f((int64_t)0, std::min(count, step));
  for (auto &th : threads) th.join();
}

// splitmix64: seeds the per-stream generators and derives per-instance seeds.
static inline uint64_t splitmix64(uint64_t &x) {
  uint64_t z = (x += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// xoshiro256** (Blackman & Vigna), the generators' bit source.
struct Rng {
  uint64_t s[4];
  explicit Rng(uint64_t seed) {
    uint64_t x = seed;
    for (int i = 0; i < 4; ++i) s[i] = splitmix64(x);
  }
  static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    const uint64_t r = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  // uniform double in [0, 1) with 53 random bits
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
  // uniform integer in [lo, hi] (inclusive), unbiased (Lemire's multiply-shift with rejection)
  int64_t range(int64_t lo, int64_t hi) {
    uint64_t span = (uint64_t)(hi - lo) + 1ull;
    if (span == 0) return (int64_t)next();
    unsigned __int128 mprod = (unsigned __int128)next() * span;
    uint64_t l = (uint64_t)mprod;
    if (l < span) {
      uint64_t thresh = (0ull - span) % span;
      while (l < thresh) {
        mprod = (unsigned __int128)next() * span;
        l = (uint64_t)mprod;
      }
    }
    return lo + (int64_t)(mprod >> 64);
  }
};

// Arc list -> CSR, stable in arc order within each tail.
struct ArcList {
  std::vector<int32_t> u, v, c;
  void add(int32_t a, int32_t b, int32_t cap) {
    u.push_back(a);
    v.push_back(b);
    c.push_back(cap);
  }
  size_t size() const { return u.size(); }
};

// Allocates an mf_host_graph from an arc list (arcs grouped by tail, stable).
mf_status host_graph_from_arcs(int32_t n, const ArcList &arcs, int32_t s, int32_t t,
                               mf_host_graph **out);
mf_host_graph *host_graph_alloc(int32_t n, int64_t m);

// Validates a CSR view: sizes, offsets monotone, dst in range, caps >= 0, s/t valid.
mf_status validate_csr(const mf_csr *g);

// Host layout build (BCSR/BCSR2 share one slot structure; RCSR its own).
struct SlotLayout {
  int32_t n = 0;
  int64_t slots = 0;
  std::vector<int32_t> off;    // n+1
  std::vector<int32_t> dst;    // slots
  std::vector<int32_t> rev;    // slots
  std::vector<int32_t> cap0;   // slots: initial residual
  std::vector<int32_t> fslot;  // m: slot of the forward direction of arc i
};
struct RcsrLayout {
  int32_t n = 0;
  int64_t m = 0;
  std::vector<int32_t> roff;   // n+1
  std::vector<int32_t> rdst;   // m: tail of the arc
  std::vector<int32_t> fidx;   // m: arc index
};
mf_status build_slot_layout(const mf_csr *g, SlotLayout &L);
mf_status build_rcsr_layout(const mf_csr *g, RcsrLayout &L);

// Host-side flow repair: returns stranded excess to the source along flow-carrying arcs,
// cancelling flow cycles met on the way.  flow[m] is modified in place.
void return_excess_to_source(const mf_csr *g, int32_t *flow);

}  // namespace mf
