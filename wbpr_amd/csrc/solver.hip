// Push-relabel solver on the GPU: kernels and the host loop that drives them.
//
// The device graph may hold k independent instances packed into one vertex and slot space
// (k = 1 for a single solve); every kernel below works on all instances at once.  Heights use
// three reserved values: sinks 0, vertices not reached by the last global relabel H = N
// (inactive), sources H + 1.  Active vertex: excess > 0 and height < H.
//
// One round = one global relabel followed by K push-relabel cycles:
//   preflow   saturate every residual arc out of every source  (maxflow-cuda/src/preflow.cu:3-31)
//   relabel   BFS from all sinks over reverse residual arcs: h(v) = dist(v, sink of v), reached
//             vertices holding excess form the first worklist.  If no reached vertex holds
//             excess, e(t_i) is the max-flow value of every instance: all arcs from the
//             unreached side into the reached side are saturated.
//             (replaces the CPU relabel maxflow-cuda/src/global_relabel.cu:4-178 and the GPU one
//             of maxflow-bcsr2/src/global_relabel_gpu.cu:4-190)
//   cycle     every active vertex u finds its lowest residual neighbour w; if h(u) > h(w) it
//             pushes min(e(u), r(u,w)) with integer atomics, else relabels h(u) = h(w) + 1
//             (TC: maxflow-cuda/src/push_relabel_kernel.cu:14-158, one lane per vertex;
//              VC: :253-537, a tile of 8..256 lanes with a min-reduction per vertex).
//             Vertices that keep excess and vertices that receive their first excess are
//             appended to the next worklist exactly once (see finish_vertex).
// The push/relabel rule is the lock-free one of He & Hong that the reference uses: only the
// owner of u lowers e(u) and r(u,*), so stale reads of other vertices' state are safe.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <limits>
#include <mutex>
#include <vector>

#include "build.h"
#include "kernels.h"
#include "mf_internal.h"
#include "solver.h"

using namespace mfk;

namespace mf {

// ============================================================================================
// kernels
// ============================================================================================

struct CycleArgs {
  int *h;
  int *e;
  int H;  // unreached / inactive height (= number of vertices)
  const unsigned char *vcls;
  int nclass;
  ListSet in, out;
  int *clear_cnt;
  unsigned long long *stats;
  int fwd_hops;  // push forwarding: further vertices a tile may discharge in the same cycle
  int *stamp;    // per vertex: the last cycle that discharged it
  int cycle;     // this cycle's number (increasing within a solve); with cycle_base, the offset
  const int *cycle_base;  // graph replays: the chunk's first cycle number, set before each replay
  unsigned long long *work;  // optional: busy clock cycles of every wave, summed over launches
  unsigned char *work_cls;   // optional: degree class of every wave's block
  // giants: kGiantWords words of state per position in the cycle's giant list (discharge_part)
  unsigned long long *giant;
};

// This cycle's number; a launch replayed from a captured graph adds the chunk's first cycle
// number, set in device memory before the replay.
__device__ __forceinline__ int cycle_now(const CycleArgs &a) {
  return a.cycle_base ? a.cycle + *a.cycle_base : a.cycle;
}

// Per-lane scan state: the lane's lowest (height, key) over its residual slots and, for that
// slot, the neighbour and the residual read during the scan.
struct Cand {
  unsigned long long best = ~0ull;
  int w = 0, res = 0;
  __device__ __forceinline__ void consider(int hw, int key, int w_, int res_) {
    const unsigned long long k = pack_hk(hw, key);
    if (k < best) {
      best = k;
      w = w_;
      res = res_;
    }
  }
};

// Lane `first` of a group of `step` lanes scans slots begin + first, begin + first + step, ...
// (up to end) of segment seg of u, B slots in flight: the B slot records are loaded together,
// then the B neighbour heights, so a lane waits for two dependent loads per B slots instead of
// two per slot.
template <int B, class L>
__device__ __forceinline__ void scan_seg(const L &g, const int *h, int u, int seg, int begin,
                                         int end, int first, int step, Cand &c) {
  for (int j0 = begin + first; j0 < end; j0 += B * step) {
    int w[B], res[B], key[B], hw[B];
#pragma unroll
    for (int b = 0; b < B; ++b) {
      w[b] = u;
      res[b] = 0;
      key[b] = 0;
      if (j0 + b * step < end) g.scan(seg, j0 + b * step, w[b], res[b], key[b]);
    }
#pragma unroll
    for (int b = 0; b < B; ++b) hw[b] = (res[b] > 0 && w[b] != u) ? h[w[b]] : 0;
#pragma unroll
    for (int b = 0; b < B; ++b)
      if (res[b] > 0 && w[b] != u) c.consider(hw[b], key[b], w[b], res[b]);
  }
}

// All slots of u; deg = their number.  r0 = u's first segment, read by the caller.
template <int B, class L>
__device__ __forceinline__ void scan_slots(const L &g, const int *h, int u, int2 r0, int first,
                                           int step, Cand &c, int &deg) {
  deg = 0;
#pragma unroll
  for (int seg = 0; seg < L::kSeg; ++seg) {
    const int2 r = seg == 0 ? r0 : g.range(u, seg);
    deg += r.y - r.x;
    scan_seg<B>(g, h, u, seg, r.x, r.y, first, step, c);
  }
}

template <class L>
__device__ __forceinline__ int slot_degree(const L &g, int u) {
  int deg = 0;
#pragma unroll
  for (int seg = 0; seg < L::kSeg; ++seg) {
    const int2 r = g.range(u, seg);
    deg += r.y - r.x;
  }
  return deg;
}

// Slots in flight per lane: a thread-centric lane covers a whole adjacency list alone, a hub
// block lane covers deg/256 slots; narrow tiles see few slots per lane.
template <int T>
constexpr int scan_batch() { return (T == 1 || T > 64) ? 8 : 4; }

// What the group decided for u after the min-reduction.
struct Plan {
  bool dead;   // no residual arc to a reached vertex: u can never reach a sink again
  int hnew;    // height after this step (h(u), or hmin + 1 when no neighbour is lower)
  int total;   // excess pushed out of u by all lanes together
  int d;       // this lane's share
};

// This lane pushes d along its candidate arc; true when the push gave w its first excess (w
// must then be queued, unless it is a sink).
template <class L>
__device__ __forceinline__ bool push_arc(const L &g, const CycleArgs &a, const Cand &c, int d,
                                         Stats &st) {
  if (d <= 0) return false;
  const int key = (int)(unsigned)(c.best & 0xffffffffull);
  const int old_w = atomicAdd(&a.e[c.w], d);
  g.push(key, d);
  st.pushes++;
  return old_w == 0 && (int)(c.best >> 32) > 0;  // height 0 only at sinks
}

// Queue item for the next cycle (every lane of the wave calls this).
__device__ __forceinline__ void queue_next(const CycleArgs &a, Stage *S, int item, bool want) {
  const int cls = (want && a.vcls) ? a.vcls[item] : 0;
  if (S) staged_append(*S, a.out, item, want, cls, a.nclass);
  else append(a.out, item, want, cls, a.nclass);
}

// One discharge step of a narrow tile, finished: relabel (if no neighbour is lower), then the
// lane holding the lowest arc pushes its share p.d along it.  Relabelling and pushing in one step
// is the lock-free relabel followed by the push it enables, with the same reads.  Then u and the
// push target are appended to the next worklist.  Exactly-once: a vertex is in the current list iff its excess became positive
// since it was last processed.  The owner re-adds u iff its own atomicSub leaves a positive
// remainder (or it relabelled below H); a pusher adds w iff its atomicAdd found e(w) == 0.
// Whichever of two atomics on one excess comes second sees the other's effect, so exactly one
// of them appends.
//
// Push forwarding: the pusher that found e(w) == 0 holds w's only ticket into the next
// worklist, and no one else is discharging w (a vertex queued for this cycle keeps positive
// excess until its own step).  Instead of queueing w it may discharge w itself, now: the
// returned w (or -1) is what the caller continues with.  fcls: forwarding allowed into vertices
// of degree class <= fcls (-1: no forwarding), so a hub never lands on a narrow tile.
template <class L>
__device__ __forceinline__ int finish_vertex(const L &g, const CycleArgs &a, Stage *S, int u,
                                             int hu, const Plan &p, const Cand &mine,
                                             bool leader, int deg, int fcls, Stats &st) {
  bool want_u = false, want_w = false;
  // the classes the appends need, read while the atomics below are in flight
  const int cu = (leader && a.vcls) ? a.vcls[u] : 0;
  const int cw = (p.d > 0 && a.vcls) ? a.vcls[mine.w] : 0;
  if (leader) {
    st.active++;
    st.slots += deg;
  }
  if (p.dead) {
    if (leader) {
      a.h[u] = a.H;
      st.relabels++;
    }
  } else {
    if (leader && p.hnew > hu) {
      a.h[u] = p.hnew;
      st.relabels++;
    }
    // the owner's subtraction and the push touch different vertices: both are in flight
    // together (the owner is the pusher when the tile has one lane)
    int old_u = 0;
    if (leader) old_u = atomicSub(&a.e[u], p.total);
    want_w = push_arc(g, a, mine, p.d, st);
    want_u = leader && old_u - p.total > 0;
  }
  // forward only into a vertex not yet discharged in this cycle: a second discharge in one
  // launch could read residuals and excess from cache lines older than the first one's atomics
  // (too high), where a first discharge can only see values that are too low
  int fwd = -1;
  if (want_w && fcls >= 0 && cw <= fcls &&
      atomicExch(&a.stamp[mine.w], cycle_now(a)) != cycle_now(a)) {
    fwd = mine.w;
    want_w = false;
  }
  if (S) {
    staged_append(*S, a.out, u, want_u, cu, a.nclass);
    staged_append(*S, a.out, mine.w, want_w, cw, a.nclass);
  } else {
    append(a.out, u, want_u, cu, a.nclass);
    append(a.out, mine.w, want_w, cw, a.nclass);
  }
  return fwd;
}

__device__ __forceinline__ void plan_heights(unsigned long long best, int hu, int H, Plan &p) {
  const int hmin = (int)(best >> 32);
  p.dead = best == ~0ull || hmin + 1 >= H;
  p.hnew = p.dead ? H : max(hu, hmin + 1);
}

// One discharge step of u by a tile of T lanes (T < kMultiPushLanes, inside one wave): the
// lowest residual arc gets min(e(u), r) after a relabel to its height + 1 if it is not lower.
// Returns the vertex the tile continues with (push forwarding, see finish_vertex) or -1; fcls is
// the forwarding class limit (-1: none).  claim: u comes from the worklist and forwarding is on,
// so the tile first stamps it (claim_queued); the stamp's exchange is in flight together with
// the reads of u's height, excess and slot range, which a failed claim simply discards (a first
// discharge of u in this launch reads the same values before or after its claim).
template <int T, class L>
__device__ __forceinline__ int discharge_tile(const L &g, const CycleArgs &a, Stage *S, int u,
                                              int lane, int fcls, bool claim, Stats &st) {
  int old = 0;
  if (claim && lane == 0) old = atomicExch(&a.stamp[u], cycle_now(a));
  const int hu = a.h[u];
  const int e0 = lane == 0 ? a.e[u] : 0;
  const int2 r0 = g.range(u, 0);
  if (claim && tile_first<T>(old) == cycle_now(a)) return -1;  // discharged in this cycle already
  // e(u) grows while the tile runs: every lane uses lane 0's read (the value subtracted below)
  const int eu = tile_first<T>(e0);
  if (hu >= a.H) return -1;  // became inactive after it was queued (all lanes read the same value)
  Cand c;
  int deg;
  scan_slots<scan_batch<T>()>(g, a.h, u, r0, lane, T, c, deg);
  Plan p;
  const unsigned long long best = tile_min<T>(c.best);
  plan_heights(best, hu, a.H, p);
  // low-degree vertices push along the minimum arc only: fanning small excesses out to many
  // neighbours (unit-capacity bipartite networks) multiplies the pushes that follow
  const bool winner = !p.dead && c.best == best;  // one lane: keys are unique
  const int total = (int)tile_sum<T>(winner ? (long long)min(eu, c.res) : 0LL);
  p.d = winner ? total : 0;
  p.total = total;
  const int f = finish_vertex(g, a, S, u, hu, p, c, lane == 0, deg, fcls, st);
  if constexpr (T == 1) {
    return f;
  } else {
    return (int)tile_max<T>(f);  // only the winner lane can hold a vertex
  }
}

// Block-wide minimum (kBlock threads, all get the result).
__device__ __forceinline__ unsigned long long block_min(unsigned long long v) {
  __shared__ unsigned long long red[kBlock / 64];
  v = tile_min<64>(v);
  if (__lane_id() == 0) red[threadIdx.x / 64] = v;
  __syncthreads();
  for (int k = 0; k < kBlock / 64; ++k) v = red[k] < v ? red[k] : v;
  __syncthreads();
  return v;
}

// Block-wide inclusive prefix sum of v in thread order; total = the sum over the block.
__device__ __forceinline__ long long block_scan(long long v, long long &total) {
  __shared__ long long wsum[kBlock / 64];
  const int wave = threadIdx.x / 64;
  long long incl = tile_scan<64>(v);
  if (__lane_id() == 63) wsum[wave] = incl;
  __syncthreads();
  long long before = 0, sum = 0;
  for (int k = 0; k < kBlock / 64; ++k) {
    if (k < wave) before += wsum[k];
    sum += wsum[k];
  }
  __syncthreads();
  total = sum;
  return incl + before;
}

// Reductions of the group of lanes that discharges one vertex together: a tile inside one wave
// (shuffles) or the whole block (LDS).  Every lane of the group calls them.
template <int T>
struct Group {
  __device__ static unsigned long long min(unsigned long long v) { return tile_min<T>(v); }
  __device__ static long long scan(long long v, long long &total) {
    total = tile_sum<T>(v);
    return tile_scan<T>(v);
  }
  __device__ static int first(int v) { return tile_first<T>(v); }  // lane 0's value
};
template <>
struct Group<kBlock> {
  __device__ static unsigned long long min(unsigned long long v) { return block_min(v); }
  __device__ static long long scan(long long v, long long &total) { return block_scan(v, total); }
  __device__ static int first(int v) {
    __shared__ int s_v;
    if (threadIdx.x == 0) s_v = v;
    __syncthreads();
    v = s_v;
    __syncthreads();
    return v;
  }
};

// This lane pushes d along the arc (key) to w, whose height was hw; true when w must be queued
// (first excess, not a sink).
template <class L>
__device__ __forceinline__ bool push_to(const L &g, const CycleArgs &a, int w, int key, int hw,
                                        int d, Stats &st) {
  const int old_w = atomicAdd(&a.e[w], d);
  g.push(key, d);
  st.pushes++;
  return old_w == 0 && hw > 0;
}

// One discharge step of a vertex by a group of T lanes that pushes along many arcs (a 64-lane
// tile; whole-block hubs use discharge_block, below), chunk by chunk.  A chunk is one batch of B slots per lane: the lane
// loads its B slot records, then their B neighbour heights, and pushes its share of what is
// left of e(u) (a prefix split over the lanes) along every one of its arcs that is lower than
// h(u).  The scan stops as soon as the excess is placed, so the admissible arcs of the first
// chunks usually take it all and the rest of a hub's slots -- a random neighbour-height gather
// each -- are not read.  Only a full scan without any arc below h(u) relabels: u moves to the
// lowest neighbour + 1, and every lane whose lowest arc of the whole scan sits at that height
// pushes along it (relabel and push in one step).  Arcs that went lower than h(u) by a stale
// read, or arcs beyond the lane's lowest, are admissible pushes of the same kind the lock-free
// rule allows (the neighbour was lower when read).  Queueing follows finish_vertex: the owner
// re-adds u iff excess remains, pushers add w iff they gave it its first excess.
constexpr int kMultiBatch = 4;  // slots per lane per chunk (a block's chunk: 1,024 slots)

template <int T, class L>
__device__ __forceinline__ void discharge_multi(const L &g, const CycleArgs &a, Stage *S, int u,
                                                int lane, Stats &st) {
  constexpr int B = kMultiBatch;
  const int hu = a.h[u];
  // e(u) grows while the group runs: all lanes split lane 0's read (the value subtracted below)
  const int eu = Group<T>::first(lane == 0 ? a.e[u] : 0);
  if (hu >= a.H) return;  // became inactive after it was queued (uniform)
  Cand keep;              // this lane's lowest arc over the whole scan
  long long placed = 0;   // excess pushed so far (the same in every lane)
  int scanned = 0;
#pragma unroll 1
  for (int seg = 0; seg < L::kSeg; ++seg) {
    const int2 r = g.range(u, seg);
    for (int cs = r.x; cs < r.y && placed < eu; cs += T * B) {
      int w[B], res[B], key[B], hw[B];
#pragma unroll
      for (int b = 0; b < B; ++b) {
        w[b] = u;
        res[b] = 0;
        key[b] = 0;
        if (cs + lane + b * T < r.y) g.scan(seg, cs + lane + b * T, w[b], res[b], key[b]);
      }
#pragma unroll
      for (int b = 0; b < B; ++b) hw[b] = (res[b] > 0 && w[b] != u) ? a.h[w[b]] : 0;
      long long want = 0;
#pragma unroll
      for (int b = 0; b < B; ++b) {
        if (res[b] <= 0 || w[b] == u) continue;
        keep.consider(hw[b], key[b], w[b], res[b]);
        if (hw[b] < hu) want += res[b];
      }
      scanned += min(r.y - cs, T * B);
      long long sum;
      const long long incl = Group<T>::scan(want, sum);
      const long long left = eu - placed;
      long long d = max(0LL, min(want, left - (incl - want)));
#pragma unroll
      for (int b = 0; b < B; ++b) {
        bool add = false;
        if (d > 0 && res[b] > 0 && w[b] != u && hw[b] < hu) {
          const int x = (int)min(d, (long long)res[b]);
          d -= x;
          add = push_to(g, a, w[b], key[b], hw[b], x, st);
        }
        queue_next(a, S, w[b], add);
      }
      placed += min(left, sum);
    }
  }
  Plan p;
  p.dead = false;
  p.hnew = hu;
  if (placed == 0 && eu > 0) {  // a full scan found no arc below h(u): relabel
    plan_heights(Group<T>::min(keep.best), hu, a.H, p);
    if (!p.dead) {
      const long long want =
          (keep.best != ~0ull && (int)(keep.best >> 32) < p.hnew) ? keep.res : 0;
      long long sum;
      const long long incl = Group<T>::scan(want, sum);
      const int d = (int)max(0LL, min(want, (long long)eu - (incl - want)));
      queue_next(a, S, keep.w, push_arc(g, a, keep, d, st));
      placed = min((long long)eu, sum);
    }
  }
  bool want_u = false;
  if (lane == 0) {
    st.active++;
    st.slots += scanned;
    if (p.dead || p.hnew > hu) {
      a.h[u] = p.hnew;
      st.relabels++;
    }
    if (!p.dead) want_u = atomicSub(&a.e[u], (int)placed) - placed > 0;
  }
  queue_next(a, S, u, want_u);
}

// One discharge step of a hub by the whole block, without block barriers while it scans: each
// wave takes every fourth chunk of kMultiBatch x 64 slots, loads the chunk's slot records and
// neighbour heights, and reserves the residual of its arcs below h(u) on a block-shared counter
// (one LDS atomic per wave and chunk).  Its share is what the reservations before it left of
// e(u); it pushes that share along its arcs in slot order.  A wave stops once the counter shows
// e(u) placed.  Waves therefore advance at their own pace instead of the slowest wave's, chunk
// by chunk.  As in discharge_multi, only a full scan without any arc below h(u) relabels: the
// block then takes the lowest (height, key) of all lanes, and every lane whose lowest arc
// sits at the new height pushes along it.
template <class L>
__device__ __forceinline__ void discharge_block(const L &g, const CycleArgs &a, Stage *S, int u,
                                                Stats &st) {
  constexpr int B = kMultiBatch, W = kBlock / 64;
  __shared__ unsigned long long s_placed;
  __shared__ int s_eu;
  const int tid = (int)threadIdx.x, lane = tid & 63, wave = tid >> 6;
  __syncthreads();  // the block's previous hub is done with the shared words
  if (tid == 0) {
    s_eu = a.e[u];
    s_placed = 0;
  }
  const int hu = a.h[u];
  __syncthreads();
  const int eu = s_eu;
  if (hu >= a.H) return;  // became inactive after it was queued (uniform)
  Cand keep;
  int scanned = 0;
#pragma unroll 1
  for (int seg = 0; seg < L::kSeg; ++seg) {
    const int2 r = g.range(u, seg);
    for (int cs = r.x + wave * 64 * B; cs < r.y; cs += W * 64 * B) {
      if (__hip_atomic_load(&s_placed, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) >=
          (unsigned long long)eu)
        break;  // wave-uniform: one LDS word
      int w[B], res[B], key[B], hw[B];
#pragma unroll
      for (int b = 0; b < B; ++b) {
        w[b] = u;
        res[b] = 0;
        key[b] = 0;
        if (cs + lane + b * 64 < r.y) g.scan(seg, cs + lane + b * 64, w[b], res[b], key[b]);
      }
#pragma unroll
      for (int b = 0; b < B; ++b) hw[b] = (res[b] > 0 && w[b] != u) ? a.h[w[b]] : 0;
      long long want = 0;
#pragma unroll
      for (int b = 0; b < B; ++b) {
        if (res[b] <= 0 || w[b] == u) continue;
        keep.consider(hw[b], key[b], w[b], res[b]);
        if (hw[b] < hu) want += res[b];
      }
      scanned += min(r.y - cs, 64 * B);
      const long long sum = tile_sum<64>(want);
      long long d = 0;
      if (sum > 0) {
        unsigned long long got = 0;
        if (lane == 0) got = atomicAdd(&s_placed, (unsigned long long)sum);
        got = (unsigned long long)tile_first<64>((int)min(got, (unsigned long long)eu));
        const long long share = max(0LL, min(sum, (long long)eu - (long long)got));
        const long long incl = tile_scan<64>(want);
        d = max(0LL, min(want, share - (incl - want)));
      }
#pragma unroll
      for (int b = 0; b < B; ++b) {
        bool add = false;
        if (d > 0 && res[b] > 0 && w[b] != u && hw[b] < hu) {
          const int x = (int)min(d, (long long)res[b]);
          d -= x;
          add = push_to(g, a, w[b], key[b], hw[b], x, st);
        }
        queue_next(a, S, w[b], add);
      }
    }
  }
  if (lane == 0) st.slots += scanned;
  __syncthreads();  // every wave's reservations are in
  long long placed = (long long)min(s_placed, (unsigned long long)eu);
  Plan p;
  p.dead = false;
  p.hnew = hu;
  if (placed == 0 && eu > 0) {  // a full scan found no arc below h(u): relabel
    plan_heights(block_min(keep.best), hu, a.H, p);
    if (!p.dead) {
      const long long want =
          (keep.best != ~0ull && (int)(keep.best >> 32) < p.hnew) ? keep.res : 0;
      long long sum;
      const long long incl = block_scan(want, sum);
      const int d = (int)max(0LL, min(want, (long long)eu - (incl - want)));
      queue_next(a, S, keep.w, push_arc(g, a, keep, d, st));
      placed = min((long long)eu, sum);
    }
  }
  bool want_u = false;
  if (tid == 0) {
    st.active++;
    if (p.dead || p.hnew > hu) {
      a.h[u] = p.hnew;
      st.relabels++;
    }
    if (!p.dead) want_u = atomicSub(&a.e[u], (int)placed) - placed > 0;
  }
  queue_next(a, S, u, want_u);
}

// One discharge step of u by T lanes; returns the vertex to continue with (forwarding) or -1.
// The tile that takes a queued vertex stamps it with the cycle before its step, so no forwarder
// discharges it a second time in the same cycle (finish_vertex).  The stamp's old value is
// consumed, which orders the exchange before the step's own atomics.
template <int T>
__device__ __forceinline__ bool claim_queued(const CycleArgs &a, int u, int lane) {
  int old = 0;
  const int cyc = cycle_now(a);
  if (lane == 0) old = atomicExch(&a.stamp[u], cyc);
  return tile_first<T>(old) != cyc;
}

// claim: u was taken from the worklist while forwarding is on (stamp it first, claim_queued).
template <int T, class L>
__device__ __forceinline__ int discharge(const L &g, const CycleArgs &a, Stage *S, int u,
                                         int lane, int fcls, bool claim, Stats &st) {
  if constexpr (T >= kMultiPushLanes) {
    if (claim && !claim_queued<T>(a, u, lane)) return -1;
    discharge_multi<T>(g, a, S, u, lane, st);
    return -1;  // several lanes push: no single vertex to continue with
  } else {
    return discharge_tile<T>(g, a, S, u, lane, fcls, claim, st);
  }
}

// Words of per-giant state, per position in the cycle's giant list: [0] the lowest (height, key)
// any part found, [1] parts done (high 32 bits) and the excess they pushed (low 32 bits), [2]
// admissible residual reserved so far.  The last part of a giant resets all three.
constexpr int kGiantWords = 4;

// x now depends on dep: the wave waits for dep (the result of an atomic) before anything that
// uses x is issued.  An empty instruction; only the register dependency matters.
__device__ __forceinline__ void order_after(unsigned long long &x, unsigned long long dep) {
  asm volatile("" : "+v"(x) : "v"(dep));
}

// One discharge step of part p (of P) of giant u, by the whole block; i = u's position in the
// cycle's giant list.  The parts of u run in any order, on any blocks, without waiting for each
// other and without fences:
//   - every part reads h(u) (only the last part writes it, after all parts have read it) and
//     e(u).  Only the last part subtracts from e(u), so every part's read is a lower bound of
//     the excess it subtracts from.  The part scans its slots chunk by chunk; for the arcs of a
//     chunk lower than h(u) it reserves their residual on word [2] (one atomicAdd) and pushes
//     min(want, e_read - reserved before it).  Whichever part reserves last with a positive
//     share stays within its own read, so all parts together push at most e(u), and the
//     excess present when the parts started is placed if the arcs admit it.  A part stops
//     scanning when its share falls short of its chunk's want (the excess is placed);
//   - it then folds its lowest (height, key) into word [0] and reports itself and its pushed
//     total on word [1].  Both are returning atomics, performed where device-wide atomics are
//     ordered before they return; the report is issued only after the minimum returned;
//   - the part whose report comes last sees every part's minimum and total: it relabels u if
//     no part found a lower neighbour (then none pushed) and pushes along its own lowest arcs at
//     the new height, subtracts everything pushed from e(u) in one atomic and re-adds u iff
//     excess remains -- the exactly-once rule of finish_vertex: until that subtraction e(u)
//     has not dropped, so no pusher has queued u.
template <class L>
__device__ void discharge_part(const L &g, const CycleArgs &a, Stage *S, int u, int p, int P,
                               int i, Stats &st) {
  constexpr int B = kMultiBatch, W = kBlock / 64;
  __shared__ unsigned long long s_word;
  __shared__ unsigned long long s_pushed;  // the part's pushes, summed over its waves
  __shared__ int s_stop;                   // a wave's share fell short: e(u) is placed
  unsigned long long *gs = a.giant + (size_t)i * kGiantWords;
  const int hu = a.h[u];
  if (hu >= a.H) return;  // uniform over the parts: inactive since it was queued
  const int lane = (int)threadIdx.x, lane64 = lane & 63, wave = lane >> 6;
  if (lane == 0) {
    s_pushed = 0;
    s_stop = 0;
  }
  // (the barriers of Group::first order the two stores above before any wave reads them)
  const int ep = Group<kBlock>::first(lane == 0 ? ld_acq(&a.e[u]) : 0);
  const int deg = slot_degree(g, u);
  const int lo = p * kPartSlots, hi = min(deg, lo + kPartSlots);
  Cand keep;               // this lane's lowest arc over the part's scanned slots
  int scanned = 0, acc = 0;
  // each wave scans every W-th chunk of 64 x B slots on its own (no block barrier per chunk):
  // it reserves its chunk's admissible residual on word [2] and pushes its share
#pragma unroll 1
  for (int seg = 0; seg < L::kSeg; ++seg) {
    const int2 r = g.range(u, seg);
    const int len = r.y - r.x;
    const int s0 = r.x + max(lo - acc, 0), s1 = r.x + min(hi - acc, len);
    acc += len;
    for (int cs = s0 + wave * 64 * B; cs < s1; cs += W * 64 * B) {
      if (__hip_atomic_load(&s_stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) break;
      int w[B], res[B], key[B], hw[B];
#pragma unroll
      for (int b = 0; b < B; ++b) {
        w[b] = u;
        res[b] = 0;
        key[b] = 0;
        if (cs + lane64 + b * 64 < s1) g.scan(seg, cs + lane64 + b * 64, w[b], res[b], key[b]);
      }
#pragma unroll
      for (int b = 0; b < B; ++b) hw[b] = (res[b] > 0 && w[b] != u) ? a.h[w[b]] : 0;
      long long want = 0;
#pragma unroll
      for (int b = 0; b < B; ++b) {
        if (res[b] <= 0 || w[b] == u) continue;
        keep.consider(hw[b], key[b], w[b], res[b]);
        if (hw[b] < hu) want += res[b];
      }
      scanned += min(s1 - cs, 64 * B);
      const long long sum = tile_sum<64>(want);
      long long d = 0;
      if (sum > 0) {
        unsigned long long got = 0;
        if (lane64 == 0) got = atomicAdd(&gs[2], (unsigned long long)sum);
        got = (unsigned long long)tile_first<64>((int)min(got, (unsigned long long)ep));
        const long long share = max(0LL, min(sum, (long long)ep - (long long)got));
        const long long incl = tile_scan<64>(want);
        d = max(0LL, min(want, share - (incl - want)));
        if (lane64 == 0) {
          atomicAdd(&s_pushed, (unsigned long long)share);
          if (share < sum) s_stop = 1;
        }
      }
#pragma unroll
      for (int b = 0; b < B; ++b) {
        bool add = false;
        if (d > 0 && res[b] > 0 && w[b] != u && hw[b] < hu) {
          const int x = (int)min(d, (long long)res[b]);
          d -= x;
          add = push_to(g, a, w[b], key[b], hw[b], x, st);
        }
        queue_next(a, S, w[b], add);
      }
    }
  }
  if (lane64 == 0) st.slots += scanned;
  const unsigned long long best = block_min(keep.best);  // (its barriers: every wave is done)
  const long long pushed = (long long)s_pushed;
  if (lane == 0) {
    const unsigned long long prev = atomicMin(&gs[0], best);
    unsigned long long report = (1ull << 32) | (unsigned long long)pushed;
    order_after(report, prev);
    s_word = atomicAdd(&gs[1], report);
  }
  __syncthreads();
  const unsigned long long word = s_word;
  __syncthreads();
  if ((int)(word >> 32) != P - 1) return;  // uniform
  // the last part: every other part has reported
  const long long total = (long long)(word & 0xffffffffull) + pushed;
  if (lane == 0) {
    unsigned long long b0 = atomicExch(&gs[0], ~0ull);
    s_word = b0;
    atomicExch(&gs[1], 0ull);
    atomicExch(&gs[2], 0ull);
  }
  __syncthreads();
  Plan pl;
  plan_heights(s_word, hu, a.H, pl);
  long long extra = 0;
  if (!pl.dead && pl.hnew > hu) {
    // relabel (no part found a lower neighbour, so none pushed): this part's lowest arcs at the
    // new height are admissible now
    const int en = Group<kBlock>::first(lane == 0 ? ld_acq(&a.e[u]) : 0);
    const long long want = (keep.best != ~0ull && (int)(keep.best >> 32) < pl.hnew) ? keep.res : 0;
    long long sum;
    const long long incl = block_scan(want, sum);
    const int d = (int)max(0LL, min(want, (long long)en - (incl - want)));
    queue_next(a, S, keep.w, push_arc(g, a, keep, d, st));
    extra = min((long long)en, sum);
  }
  bool want_u = false;
  if (lane == 0) {
    st.active++;
    if (pl.dead || pl.hnew > hu) {
      a.h[u] = pl.hnew;
      st.relabels++;
    }
    if (!pl.dead) want_u = atomicSub(&a.e[u], (int)(total + extra)) - (total + extra) > 0;
  }
  queue_next(a, S, u, want_u);
  __syncthreads();  // the shared words are reused by the block's next part
}

// Giants in worklist class c: every giant has ceil(deg / kPartSlots) parts; part j of the list
// (counted over the giants in list order) goes to block j mod nblk.  Giants are few (the
// largest hubs), so each block prefix-sums the parts of up to kBlock * 4 of them in LDS.
template <class L>
__device__ __forceinline__ void cycle_giant(const L &g, const CycleArgs &a, Stage *S, int c,
                                            int cnt, int b, int nblk, Stats &st) {
  constexpr int kPer = 4, kChunk = kBlock * kPer;
  __shared__ int s_pre[kChunk + 1];
  const int base = a.in.cls[c];
  for (int k0 = 0; k0 < cnt; k0 += kChunk) {
    const int kc = min(kChunk, cnt - k0);
    int parts[kPer], mine = 0;
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
      const int x = (int)threadIdx.x * kPer + q;
      parts[q] = x < kc ? (slot_degree(g, a.in.items[base + k0 + x]) + kPartSlots - 1) / kPartSlots
                        : 0;
      mine += parts[q];
    }
    long long total;
    int pre = (int)(block_scan(mine, total) - mine);
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
      const int x = (int)threadIdx.x * kPer + q;
      if (x < kc) s_pre[x] = pre;
      pre += parts[q];
    }
    if (threadIdx.x == 0) s_pre[kc] = (int)total;
    __syncthreads();
    for (int j = b; j < (int)total; j += nblk) {
      int lo = 0, hi = kc;  // s_pre[lo] <= j < s_pre[hi]
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (s_pre[mid] <= j) lo = mid;
        else hi = mid;
      }
      discharge_part(g, a, S, a.in.items[base + k0 + lo], j - s_pre[lo], s_pre[lo + 1] - s_pre[lo],
                     k0 + lo, st);
    }
    __syncthreads();  // s_pre is rewritten for the next chunk
  }
}

template <int T, class L>
__device__ __forceinline__ void cycle_class(const L &g, const CycleArgs &a, Stage *S, int c,
                                            int cnt, int b, int nblk, Stats &st) {
  const int base = a.in.cls[c];
  if constexpr (T <= 64) {
    constexpr int tpb = kBlock / T;
    const int tile = b * tpb + (int)threadIdx.x / T;
    const int ntiles = nblk * tpb;
    const int lane = (int)threadIdx.x % T;
    for (int i = tile; i < cnt; i += ntiles) {
      int v = a.in.items[base + i];
      for (int hop = 0; v >= 0; ++hop)
        v = discharge<T>(g, a, S, v, lane, hop < a.fwd_hops ? c : -1, hop == 0 && a.fwd_hops > 0, st);
    }
  } else {
    for (int i = b; i < cnt; i += nblk) discharge_block(g, a, S, a.in.items[base + i], st);
  }
}

// The class, index within the class and class block count of this block.  Block rb serves
// degree class c when d.blk_begin[c] <= rb < d.blk_begin[c+1], rb counted from the end of the
// grid: blocks start in index order, so the classes of the longest blocks (hubs) start first and
// the short ones fill in behind them instead of leaving the hubs as the launch's tail.
struct BlockSlot {
  int c, b, nblk;
};
__device__ __forceinline__ BlockSlot block_slot(const Dispatch &d) {
  const int rb = (int)(gridDim.x - 1 - blockIdx.x);
  int c = 0;
#pragma unroll
  for (int q = 1; q < kNumClasses; ++q)
    if (q < d.nclass && rb >= d.blk_begin[q]) c = q;
  const int b0 = pick(d.blk_begin, c);
  return BlockSlot{c, rb - b0, pick(d.blk_begin, c + 1) - b0};
}

template <class L>
__global__ __launch_bounds__(kBlock) void k_cycle(L g, CycleArgs a, Dispatch d) {
  // counters of the list two cycles ahead: last read by the previous launch, next written by
  // the following one
  const long long t_start = a.work ? clock64() : 0;
  if (blockIdx.x == 0 && threadIdx.x < (unsigned)kNumClasses)
    a.clear_cnt[threadIdx.x * kCtlStride] = 0;
  const BlockSlot bs = block_slot(d);
  const int c = bs.c, b = bs.b, nblk = bs.nblk;
  const int cnt = list_count(a.in, c);
  // the grid is sized for the class populations; a block whose first tile has no item leaves
  // at once (small worklists: most of the grid)
  const int lanes = pick(d.lanes, c);
  if (lanes <= 64 ? b * (kBlock / lanes) >= cnt : (lanes == kBlock ? b >= cnt : cnt == 0)) return;
  __shared__ Stage S;
  stage_init(S);
  Stats st{0, 0, 0, 0};
  switch (lanes) {
    case 1: cycle_class<1>(g, a, &S, c, cnt, b, nblk, st); break;
    case 2: cycle_class<2>(g, a, &S, c, cnt, b, nblk, st); break;
    case 4: cycle_class<4>(g, a, &S, c, cnt, b, nblk, st); break;
    case 8: cycle_class<8>(g, a, &S, c, cnt, b, nblk, st); break;
    case 16: cycle_class<16>(g, a, &S, c, cnt, b, nblk, st); break;
    case 32: cycle_class<32>(g, a, &S, c, cnt, b, nblk, st); break;
    case 64: cycle_class<64>(g, a, &S, c, cnt, b, nblk, st); break;
    case kSplitLanes: cycle_giant(g, a, &S, c, cnt, b, nblk, st); break;
    default: cycle_class<kBlock>(g, a, &S, c, cnt, b, nblk, st); break;
  }
  if (a.work && __lane_id() == 0) {  // the wave's own work, before the block's flush
    const int wv = (int)blockIdx.x * (kBlock / 64) + (int)threadIdx.x / 64;
    a.work[wv] += (unsigned long long)(clock64() - t_start);  // one writer per wave slot
    a.work_cls[wv] = (unsigned char)c;
  }
  stage_flush(S, a.out, a.nclass);
  block_flush(st.active, st.slots, st.pushes, st.relabels, a.stats, ST_ACTIVE);
}

// ---- global relabel ------------------------------------------------------------------------

struct BfsArgs {
  int *h;
  const int *e;
  int H;
  const unsigned char *vcls;  // degree class of the frontier lists (always the VC classes)
  int nclass;
  const unsigned char *acls;  // degree class of the cycle worklists (null: one class)
  int nclass_act;
  ListSet fr[2];   // frontier items, double-buffered by level parity
  int *fcnt[3];    // frontier counters: level L reads fcnt[L%3], writes fcnt[(L+1)%3]
  ListSet active;  // reached vertices with excess (first worklist of the cycles)
  int *flags;      // [0] a source was reached, [1] a worklist overflowed
  int *state;      // [0] level, [1] frontier size (exit state of the one-workgroup kernel)
  unsigned long long *stats;
  int cls_mean[kNumClasses];  // mean residual degree of each frontier class
  int wg_slots;               // the one-workgroup kernel hands over above this many frontier slots
  // vertices a bottom-up level examined and did not reach, by class in the regions of the
  // vertex order (table rem_cls), double-buffered by level parity; the next bottom-up level walks
  // only these.  Counters by level % 3.  rem_items[0] null: every level walks all vertices.
  int *rem_items[2];
  const int *rem_cls;
  int *rcnt[3];
};

// LDS stages of a grid-wide BFS launch (frontier and worklist appends), or none.
struct BfsStages {
  Stage *f;
  Stage *a;
};

// cw, aw: the degree classes of w in the frontier lists and the cycle worklists (read by the
// caller, ahead of time where latency matters)
__device__ __forceinline__ void bfs_append(const BfsStages &S, const BfsArgs &a,
                                           const ListSet &out, int w, bool claim, bool act, int cw,
                                           int aw) {
  if (S.f) {
    staged_append(*S.f, out, w, claim, cw, a.nclass);
    staged_append(*S.a, a.active, w, act, aw, a.nclass_act);
  } else {
    append(out, w, claim, cw, a.nclass);
    append(a.active, w, act, aw, a.nclass_act);
  }
}

__device__ __forceinline__ void bfs_append(const BfsStages &S, const BfsArgs &a,
                                           const ListSet &out, int w, bool claim, bool act) {
  const int cw = (claim && a.vcls) ? a.vcls[w] : 0;
  const int aw = (act && a.acls) ? a.acls[w] : 0;
  bfs_append(S, a, out, w, claim, act, cw, aw);
}

// Slots one BFS lane keeps in flight.  Grid-wide levels (large frontiers, bandwidth matters)
// issue the dependent steps -- slot record, reverse residual, neighbour height, claim, class --
// one after the other for the whole batch; the one-workgroup kernel of a single graph (small
// frontiers, thousands of levels, latency matters) reads the neighbour's height, excess and
// classes together with the reverse residual, ahead of knowing whether the arc has residual
// capacity.  Excesses do not change during a global relabel.
constexpr int kBfsBatch = 4;

// Lanes lane, lane + step, ... of a group of `step` lanes (a tile, a block or a group of
// blocks: nothing here is shared inside the group) expand u's slots.
template <bool ahead, class L>
__device__ __forceinline__ void bfs_vertex(const L &g, const BfsArgs &a, const ListSet &out,
                                           const BfsStages &S, int u, int lane, int step,
                                           int level, long long &ns) {
  constexpr int B = kBfsBatch;
#pragma unroll
  for (int seg = 0; seg < L::kSeg; ++seg) {
    const int2 r = g.range(u, seg);
    if (lane == 0) ns += r.y - r.x;
    for (int base = r.x + lane; base < r.y; base += step * B) {
      int w[B], idx[B], rin[B], hw[B], ew[B], cw[B], aw[B];
      bool claim[B];
#pragma unroll
      for (int b = 0; b < B; ++b) {
        w[b] = -1;
        idx[b] = 0;
        if (base + b * step < r.y) g.bfs_slot(seg, base + b * step, w[b], idx[b]);
      }
      if constexpr (ahead) {
#pragma unroll
        for (int b = 0; b < B; ++b) {
          const bool ok = w[b] >= 0 && w[b] != u;
          rin[b] = ok ? g.bfs_rin(seg, idx[b]) : 0;
          hw[b] = ok ? a.h[w[b]] : -1;
          ew[b] = ok ? a.e[w[b]] : 0;
          cw[b] = (ok && a.vcls) ? a.vcls[w[b]] : 0;
          aw[b] = (ok && a.acls) ? a.acls[w[b]] : 0;
        }
      } else {
#pragma unroll
        for (int b = 0; b < B; ++b) rin[b] = (w[b] >= 0 && w[b] != u) ? g.bfs_rin(seg, idx[b]) : 0;
#pragma unroll
        for (int b = 0; b < B; ++b) hw[b] = rin[b] > 0 ? a.h[w[b]] : -1;
      }
#pragma unroll
      for (int b = 0; b < B; ++b) {
        claim[b] = false;
        if (rin[b] <= 0) continue;
        if (hw[b] == a.H + 1) a.flags[0] = 1;  // residual path from a source: invariant broken
        else if (hw[b] == a.H) claim[b] = atomicCAS(&a.h[w[b]], a.H, level + 1) == a.H;
      }
      if constexpr (ahead) {
#pragma unroll
        for (int b = 0; b < B; ++b)
          bfs_append(S, a, out, w[b], claim[b], claim[b] && ew[b] > 0, cw[b], aw[b]);
      } else {
#pragma unroll
        for (int b = 0; b < B; ++b) bfs_append(S, a, out, w[b], claim[b], claim[b] && a.e[w[b]] > 0);
      }
    }
  }
}

template <int T, bool ahead, class L>
__device__ __forceinline__ void bfs_class(const L &g, const BfsArgs &a, const ListSet &in,
                                          const ListSet &out, const BfsStages &S, int c, int cnt,
                                          int tile, int ntiles, int lane, int level, long long &nv,
                                          long long &ns) {
  const int base = in.cls[c];
  for (int i = tile; i < cnt; i += ntiles) {
    bfs_vertex<ahead>(g, a, out, S, in.items[base + i], lane, T, level, ns);
    if (lane == 0) ++nv;
  }
}

// ahead: read neighbour state ahead of the residual check (the latency-bound one-workgroup BFS
// of a single graph); grid-wide levels and the batched kernel, whose thousand workgroups share
// the memory system, read it only for arcs with residual capacity.
template <bool ahead, class L>
__device__ __forceinline__ void bfs_dispatch(const L &g, const BfsArgs &a, const ListSet &in,
                                             const ListSet &out, const BfsStages &S, int c,
                                             int lanes, int cnt, int tid, int nthreads, int level,
                                             long long &nv, long long &ns) {
  // A BFS lane covers kBfsBatch slots per step, so a class processed with `lanes` lanes by the
  // cycle kernels uses lanes / kBfsBatch lanes per vertex here (same slots per step, four times
  // as many vertices in flight).
  switch (lanes) {
#define MF_BFS_CASE(LANES, TT)                                                                \
  case LANES:                                                                                 \
    bfs_class<TT, ahead>(g, a, in, out, S, c, cnt, tid / TT, nthreads / TT, tid % TT, level, nv, ns); \
    break;
    MF_BFS_CASE(1, 1)
    MF_BFS_CASE(8, 2)
    MF_BFS_CASE(16, 4)
    MF_BFS_CASE(32, 8)
    MF_BFS_CASE(64, 16)
#undef MF_BFS_CASE
    default: {
      // hubs: a group of whole blocks per vertex, as many blocks as the class has per hub in
      // the frontier (one hub of 200,000 slots would keep a single tile busy for milliseconds)
      const int per = (int)blockDim.x;
      const BlockGroups bg = block_groups(tid / per, nthreads / per, cnt);
      const int base = in.cls[c];
      for (int i = bg.group; i < cnt; i += bg.ngroups) {
        bfs_vertex<ahead>(g, a, out, S, in.items[base + i], bg.lane, bg.width, level, ns);
        if (bg.lane == 0) ++nv;
      }
      break;
    }
  }
}

// Heights to "unreached", counters to zero.
__global__ void k_bfs_init(BfsArgs a, int *ccnt0, int *ccnt1) {
  const int stride = gridDim.x * blockDim.x;
  for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < a.H; v += stride) a.h[v] = a.H;
  if (blockIdx.x == 0 && threadIdx.x < (unsigned)kNumClasses) {
    const int c = threadIdx.x;
    a.fcnt[0][c * kCtlStride] = 0;
    a.fcnt[1][c * kCtlStride] = 0;
    a.fcnt[2][c * kCtlStride] = 0;
    ccnt0[c * kCtlStride] = 0;
    ccnt1[c * kCtlStride] = 0;
    if (c == 0) {
      a.flags[0] = 0;
      a.state[2] = 0;  // grid levels' marks (k_bfs_level)
      a.state[3] = 0;
      a.state[4] = 0;  // slots explored by grid-wide levels (a 64-bit counter in words 4 and 5)
      a.state[5] = 0;
    }
  }
}

// Terminal heights and the level-0 frontier (all sinks).
__global__ void k_bfs_seed(BfsArgs a, const int *src, const int *snk, int k) {
  ListSet f0 = a.fr[0];
  f0.cnt = a.fcnt[0];
  const int stride = gridDim.x * blockDim.x;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < k; i += stride) {
    const int t = snk[i];
    a.h[t] = 0;
    a.h[src[i]] = a.H + 1;
    append(f0, t, true, a.vcls ? a.vcls[t] : 0, a.nclass);
  }
}

// LDS frontier entries per group of the one-workgroup BFS (below), by degree class
constexpr int kWgCapA = 2048, kWgCapB = 256, kWgCapC = 32;
constexpr int kWgCap = kWgCapA + kWgCapB + kWgCapC;
__host__ __device__ constexpr int wg_group(int cls) { return cls <= 1 ? 0 : (cls <= 3 ? 1 : 2); }

// Where a batch of top-down levels stops: a level after the batch's first does nothing when its
// frontier would rather go bottom-up (estimated slots above stop_slots) or back to the
// one-workgroup kernel (at most small vertices and half its slot limit, fitting its LDS groups),
// or is empty.  Every block decides alike from the same complete counters.  A level that runs
// marks the next one (state[2 + (next & 1)] = next, a word it does not read itself); a level
// runs only if it was marked, so once one level of the batch does nothing, none after it runs
// (the counters are triple-buffered: three levels on, a skipped level's input counters would
// look valid again).
struct LevelStop {
  int first;             // the batch's first level (always runs)
  long long stop_slots;  // bottom-up switch point in estimated frontier slots
  int small;             // hand-back size to the one-workgroup kernel
};

template <class L>
__device__ __forceinline__ bool level_runs(const BfsArgs &a, const Dispatch &d, const ListSet &in,
                                           int level, const LevelStop &ls, long long &slots) {
  int size = 0, ga = 0, gb = 0, gc = 0;  // frontier size and its one-workgroup LDS groups
  slots = 0;
  for (int c = 0; c < d.nclass; ++c) {
    const int k = list_count(in, c);
    size += k;
    slots += (long long)k * pick(a.cls_mean, c);
    const int q = wg_group(c);
    ga += q == 0 ? k : 0;
    gb += q == 1 ? k : 0;
    gc += q == 2 ? k : 0;
  }
  const bool fits = ga <= kWgCapA && gb <= kWgCapB && gc <= kWgCapC;
  if (level == ls.first) return true;
  if (a.state[2 + (level & 1)] != level) return false;  // the level before did not run
  if (size == 0 || slots > ls.stop_slots) return false;
  if (size <= ls.small && slots <= a.wg_slots / 2 && fits) return false;
  return true;
}

// One BFS level over the whole grid.
template <class L>
__global__ __launch_bounds__(kBlock) void k_bfs_level(L g, BfsArgs a, Dispatch d, int level,
                                                      LevelStop ls) {
  {
    ListSet cur = a.fr[0];
    cur.cnt = pick(a.fcnt, level % 3);
    long long slots;
    if (!level_runs<L>(a, d, cur, level, ls, slots)) return;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      a.state[2 + ((level + 1) & 1)] = level + 1;  // the next level may run
      // estimated slots of the top-down frontiers expanded so far (the bottom-up rule's input)
      atomicAdd((unsigned long long *)(a.state + 4), (unsigned long long)slots);
    }
  }
  if (blockIdx.x == 0 && threadIdx.x < (unsigned)kNumClasses)
    pick(a.fcnt, (level + 2) % 3)[threadIdx.x * kCtlStride] = 0;
  const BlockSlot bs = block_slot(d);
  const int c = bs.c, b = bs.b, nblk = bs.nblk;
  // the two frontier lists share their class regions; only the item arrays alternate (selecting
  // whole descriptors would copy them to scratch memory)
  ListSet in = a.fr[0], out = a.fr[0];
  in.items = (level & 1) ? a.fr[1].items : a.fr[0].items;
  out.items = (level & 1) ? a.fr[0].items : a.fr[1].items;
  in.cnt = pick(a.fcnt, level % 3);
  out.cnt = pick(a.fcnt, (level + 1) % 3);
  const int cnt = list_count(in, c);
  // blocks whose first tile has no frontier vertex leave at once (tiles of lanes / 4 BFS lanes;
  // hub classes share their vertices among all their blocks)
  const int lanes = pick(d.lanes, c);
  const int tt = lanes == 1 ? 1 : ((lanes == 8 || lanes == 16 || lanes == 32 || lanes == 64) ? lanes / 4 : 0);
  if (cnt == 0 || (tt > 0 && b * (kBlock / tt) >= cnt)) return;
  __shared__ Stage sf, sa;
  stage_init(sf);
  stage_init(sa);
  long long nv = 0, ns = 0;
  bfs_dispatch<false>(g, a, in, out, BfsStages{&sf, &sa}, c, pick(d.lanes, c), cnt,
               b * kBlock + (int)threadIdx.x, nblk * kBlock, level, nv, ns);
  stage_flush(sf, out, a.nclass);
  stage_flush(sa, a.active, a.nclass_act);
  block_flush(nv, ns, 0, 0, a.stats, ST_BFS_VERTS);
}

// ---- cooperative top-down levels (mid-size frontiers) ----
//
// A frontier too large for the one-workgroup kernel and small next to the grid: kCoopBlocks
// co-resident workgroups (a cooperative launch) run level after level, each level as
// k_bfs_level runs it but over all classes with tiles numbered across the blocks, and a grid
// barrier between levels: no launch and no host round trip per level.  After each barrier every
// block reads the same complete counters and decides alike whether to go on (CoopStop).  The
// barrier's agent-scope release and acquire make a level's list items and counters, written by
// blocks on other XCDs, visible to the next level.  A barrier that does not complete within its
// spin limit sets *abort and every block leaves; the host then redoes the global relabel without
// cooperative levels.

struct CoopStop {
  int small;              // back to the one-workgroup kernel at or below this frontier size,
  long long small_slots;  //   this many estimated slots and fitting its LDS groups
  long long max_slots;    // back to the host (grid-wide or bottom-up levels) above this many
};

constexpr unsigned kCoopSpinLimit = 1u << 20;  // barrier polls before giving up (about a second)

// Grid barrier: blocks count arrivals on *bar (zero at launch); the barrier of generation gen
// waits for gen * gridDim.x arrivals.  False if it gave up.
__device__ __forceinline__ bool coop_barrier(unsigned *bar, unsigned target, int *abort) {
  __shared__ int s_ok;
  __syncthreads();
  if (threadIdx.x == 0) {
    int ok = 1;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");  // the block's level is visible first
    atomicAdd(bar, 1u);
    unsigned spins = 0;
    while (__hip_atomic_load(bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
      if (__hip_atomic_load(abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ||
          ++spins > kCoopSpinLimit) {
        atomicExch(abort, 1);
        ok = 0;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");  // every other block's level is visible
    s_ok = ok;
  }
  __syncthreads();
  return s_ok != 0;
}

template <class L>
__global__ __launch_bounds__(kBlock) void k_bfs_coop(L g, BfsArgs a, Dispatch d, int level,
                                                     CoopStop cs, unsigned *bar, int *abort) {
  __shared__ Stage sf, sa;
  __shared__ int s_go, s_size;
  __shared__ long long s_slots;
  const int gtid = (int)(blockIdx.x * kBlock + threadIdx.x);
  const int gthreads = (int)(gridDim.x * kBlock);
  long long nv = 0, ns = 0;
  for (unsigned gen = 1;; ++gen) {
    ListSet in = a.fr[0];
    in.items = (level & 1) ? a.fr[1].items : a.fr[0].items;
    in.cnt = pick(a.fcnt, level % 3);
    if (threadIdx.x == 0) {
      int size = 0, ga = 0, gb = 0, gc = 0;
      long long slots = 0;
      for (int c = 0; c < d.nclass; ++c) {
        const int k = list_count(in, c);
        size += k;
        slots += (long long)k * pick(a.cls_mean, c);
        const int q = wg_group(c);
        ga += q == 0 ? k : 0;
        gb += q == 1 ? k : 0;
        gc += q == 2 ? k : 0;
      }
      const bool small = size <= cs.small && slots <= cs.small_slots && ga <= kWgCapA &&
                         gb <= kWgCapB && gc <= kWgCapC;
      // the first level is the host's choice; later ones stop where the host would switch
      s_go = size > 0 && (gen == 1 || (slots <= cs.max_slots && !small));
      s_size = size;
      s_slots = slots;
    }
    __syncthreads();
    if (!s_go) {
      if (blockIdx.x == 0 && threadIdx.x == 0) {
        a.state[0] = level;
        a.state[1] = s_size;
      }
      break;
    }
    if (blockIdx.x == 0 && threadIdx.x < (unsigned)kNumClasses)
      pick(a.fcnt, (level + 2) % 3)[threadIdx.x * kCtlStride] = 0;
    if (blockIdx.x == 0 && threadIdx.x == 0)  // top-down slots explored (the bottom-up rule)
      atomicAdd((unsigned long long *)(a.state + 4), (unsigned long long)s_slots);
    ListSet out = a.fr[0];
    out.items = (level & 1) ? a.fr[0].items : a.fr[1].items;
    out.cnt = pick(a.fcnt, (level + 1) % 3);
    stage_init(sf);
    stage_init(sa);
    for (int c = 0; c < d.nclass; ++c) {
      const int cnt = list_count(in, c);
      if (cnt > 0)
        bfs_dispatch<true>(g, a, in, out, BfsStages{&sf, &sa}, c, pick(d.lanes, c), cnt, gtid,
                           gthreads, level, nv, ns);
    }
    stage_flush(sf, out, a.nclass);
    stage_flush(sa, a.active, a.nclass_act);
    if (!coop_barrier(bar, gen * gridDim.x, abort)) break;
    ++level;
  }
  block_flush(nv, ns, 0, 0, a.stats, ST_BFS_VERTS);
}

// ---- bottom-up levels (large frontiers) ----
//
// When the frontier holds a large part of the graph, a top-down level scans most slots of the
// graph with two random gathers each (reverse residual, neighbour height) and claims by CAS.
// Bottom-up, every still-unreached vertex w scans its own slots -- the residual of w -> u sits
// in w's slot record, so one coalesced load and one gather per slot -- and stops at the first
// residual arc into the current level.  The tile that owns w is its only writer: no CAS.

// Vertices of nonzero degree grouped by degree class (static for a graph).
struct VertexOrder {
  const int *items;
  int cls_off[kNumClasses];
  int cls_cnt[kNumClasses];
};

// Lanes of a tile that reported a hit (the tile is a contiguous group of T lanes of the wave).
template <int T>
__device__ __forceinline__ bool tile_any(bool hit) {
  const unsigned long long m = __ballot(hit);
  if constexpr (T == 64) {
    return m != 0;
  } else {
    const int shift = __lane_id() & ~(T - 1);
    return ((m >> shift) & ((1ull << T) - 1ull)) != 0;
  }
}

// Tile (T lanes) for the still unreached vertex w of class c.  Found: w joins the next frontier
// (and the cycle worklist if it has excess); not found: w goes to the next level's list of
// unreached vertices (rem, when kept).
template <int T, class L>
__device__ __forceinline__ void bu_vertex(const L &g, const BfsArgs &a, const ListSet &out,
                                          const ListSet &rem, const BfsStages &S, Stage *SR,
                                          int w, int c, int lane, int level, long long &nv,
                                          long long &ns) {
  constexpr int B = kBfsBatch;
  // reached already, or a source (tile-uniform: one address).  Sources are not scanned here:
  // the top-down levels check that no source reaches a sink
  if (a.h[w] != a.H) return;
  if (lane == 0) ++nv;
  bool found = false;
#pragma unroll
  for (int seg = 0; seg < L::kSeg; ++seg) {
    if (found) break;
    const int2 r = g.range(w, seg);
    for (int base = r.x + lane; base - lane < r.y; base += T * B) {
      int u[B], res[B], key[B], hu[B];
#pragma unroll
      for (int b = 0; b < B; ++b) {
        u[b] = w;
        res[b] = 0;
        if (base + b * T < r.y) g.scan(seg, base + b * T, u[b], res[b], key[b]);
      }
#pragma unroll
      for (int b = 0; b < B; ++b) hu[b] = (res[b] > 0 && u[b] != w) ? a.h[u[b]] : -1;
      bool hit = false;
#pragma unroll
      for (int b = 0; b < B; ++b) hit |= hu[b] == level;
      if (lane == 0) ns += min(r.y - (base - lane), T * B);
      if (tile_any<T>(hit)) {
        found = true;
        break;
      }
    }
  }
  const bool mine = found && lane == 0;
  if (mine) a.h[w] = level + 1;
  // excess and worklist class only for the reached: the unreached are most of the vertices a
  // level examines, and the level is bound by its memory requests
  const bool act = mine && a.e[w] > 0;
  bfs_append(S, a, out, w, mine, act, c, (act && a.acls) ? a.acls[w] : 0);
  if (rem.items) staged_append(*SR, rem, w, !found && lane == 0, c, a.nclass);
}

template <int T, class L>
__device__ __forceinline__ void bu_class(const L &g, const BfsArgs &a, const int *items, int base,
                                         int cnt, const ListSet &out, const ListSet &rem,
                                         const BfsStages &S, Stage *SR, int c, int tile,
                                         int ntiles, int lane, int level, long long &nv,
                                         long long &ns) {
  for (int i = tile; i < cnt; i += ntiles)
    bu_vertex<T>(g, a, out, rem, S, SR, items[base + i], c, lane, level, nv, ns);
}

// walk_all: the level walks every nonzero-degree vertex (the first bottom-up level after a
// top-down one, or no lists of unreached vertices are kept); otherwise those the level before
// left unreached.  Bottom-up levels are issued a few at a time between two read-backs: a level
// after the batch's first (first) runs only if the level before it ran (its mark, as in
// LevelStop) and the frontier still holds at least stay vertices, the size below which the
// host turns back to top-down levels; every block decides alike from the complete counters.
template <class L>
__global__ __launch_bounds__(kBlock) void k_bfs_bottom_up(L g, BfsArgs a, Dispatch d,
                                                          VertexOrder vo, int level,
                                                          int walk_all, int first, int stay) {
  {
    ListSet cur = a.fr[0];
    cur.cnt = pick(a.fcnt, level % 3);
    int size = 0;
    long long slots = 0;
    for (int c = 0; c < d.nclass; ++c) {
      const int k = list_count(cur, c);
      size += k;
      slots += (long long)k * pick(a.cls_mean, c);
    }
    if (level != first && (a.state[2 + (level & 1)] != level || size == 0 || size < stay)) return;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      a.state[2 + ((level + 1) & 1)] = level + 1;  // the next level may run
      // estimated slots of the frontiers expanded so far (the bottom-up rule's input)
      atomicAdd((unsigned long long *)(a.state + 4), (unsigned long long)slots);
    }
  }
  if (blockIdx.x == 0 && threadIdx.x < (unsigned)kNumClasses) {
    pick(a.fcnt, (level + 2) % 3)[threadIdx.x * kCtlStride] = 0;
    if (a.rem_items[0]) pick(a.rcnt, (level + 2) % 3)[threadIdx.x * kCtlStride] = 0;
  }
  __shared__ Stage sf, sa, sr;
  stage_init(sf);
  stage_init(sa);
  stage_init(sr);
  const BfsStages S{&sf, &sa};
  const BlockSlot bs = block_slot(d);
  const int c = bs.c;
  const int tid = bs.b * kBlock + (int)threadIdx.x;
  const int nthreads = bs.nblk * kBlock;
  ListSet out = a.fr[0];  // the frontier lists differ only in their item arrays
  out.items = (level & 1) ? a.fr[0].items : a.fr[1].items;
  out.cnt = pick(a.fcnt, (level + 1) % 3);
  const bool all = walk_all || !a.rem_items[0];
  const int base = pick(vo.cls_off, c);
  const int *items = all ? vo.items : ((level & 1) ? a.rem_items[1] : a.rem_items[0]);
  const int cnt = all ? pick(vo.cls_cnt, c) : pick(a.rcnt, level % 3)[c * kCtlStride];
  ListSet rem;
  rem.items = a.rem_items[0] ? ((level & 1) ? a.rem_items[0] : a.rem_items[1]) : nullptr;
  rem.cls = a.rem_cls;
  rem.cnt = pick(a.rcnt, (level + 1) % 3);
  rem.cstride = kCtlStride;
  rem.ovf = a.flags + 1;
  long long nv = 0, ns = 0;
  switch (pick(d.lanes, c)) {  // a quarter of the class's lanes (kBottomUpLanes) per vertex
#define MF_BU_CASE(LANES, TT)                                                                  \
  case LANES:                                                                                  \
    bu_class<TT>(g, a, items, base, cnt, out, rem, S, &sr, c, tid / TT, nthreads / TT,         \
                 tid % TT, level, nv, ns);                                                     \
    break;
    MF_BU_CASE(1, 1)
    MF_BU_CASE(4, 1)
    MF_BU_CASE(8, 2)
    MF_BU_CASE(16, 4)
    MF_BU_CASE(32, 8)
    MF_BU_CASE(64, 16)
    default:
      bu_class<64>(g, a, items, base, cnt, out, rem, S, &sr, c, tid / 64, nthreads / 64,
                   tid % 64, level, nv, ns);
      break;
#undef MF_BU_CASE
  }
  stage_flush(sf, out, a.nclass);
  stage_flush(sa, a.active, a.nclass_act);
  if (rem.items) stage_flush(sr, rem, a.nclass);
  block_flush(nv, ns, 0, 0, a.stats, ST_BU_VERTS);
}

// ---- one-workgroup BFS ----
//
// All BFS levels inside one workgroup while the frontier is small: workgroup barriers replace
// kernel launches, which is what long-diameter graphs (thousands of levels) need, and a level
// costs only its dependent memory round trips.  The frontier lives in LDS in three groups by
// degree class -- up to 16 slots (tiles of 4 lanes), up to 1,024 (one wave per vertex), more
// (the whole workgroup per vertex) -- and, for one-segment layouts, every entry carries its slot
// range, loaded by the lane that claims the vertex together with its claim.  A level is then
// three round trips (slot records; reverse residuals with the neighbours' heights, excesses and
// classes; claims with the claimed vertices' ranges) instead of five (frontier item and offsets
// first).  When the next frontier outgrows a group or the hand-over limits, its vertices go to
// the device frontier lists by class, as the grid-wide kernels read them, and the kernel
// returns.  No device-scope fence is needed: every value this kernel reads that another kernel
// wrote was written before the launch.

__device__ __forceinline__ int wg_cap(int q) { return q == 0 ? kWgCapA : (q == 1 ? kWgCapB : kWgCapC); }
__device__ __forceinline__ int wg_off(int q) { return q == 0 ? 0 : (q == 1 ? kWgCapA : kWgCapA + kWgCapB); }

// Where a level's claimed vertices go: the LDS groups of the next buffer, or (when a group is
// full) the device list of the next level, counted in LDS.
struct WgOut {
  int *u, *b, *e;  // next buffer's entries (b, e: slot ranges, one-segment layouts only)
  int *n;          // entries per group (may count past the capacity: readers clamp)
  ListSet dev;     // device list of the next level (items in global memory, counters in LDS)
  ListSet act;     // the cycle worklist (items in global memory, counters in LDS for the launch)
  int *ovf;        // set when any vertex went to the device list
  int nclass;
};

// Wave-aggregated append of the claimed w (class cw, slot range [rb, re)) to its LDS group;
// past the group's capacity it goes to the device list.  Every lane of the wave calls this.
template <bool kPre>
__device__ __forceinline__ void wg_append(const WgOut &o, int w, bool want, int cw, int rb,
                                          int re) {
  if (__ballot(want) == 0) return;
  const int lane = __lane_id();
  const int grp = wg_group(cw);
  bool spill = false;
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const bool mine = want && grp == q;
    const unsigned long long mask = __ballot(mine);
    if (mask == 0) continue;
    const int leader = __ffsll((long long)mask) - 1;
    const int rank = __popcll(mask & ((1ull << lane) - 1ull));
    int base = 0;
    if (lane == leader) base = atomicAdd(&o.n[q], __popcll(mask));
    base = wave_lane(base, leader);
    if (mine) {
      const int pos = base + rank;
      if (pos < wg_cap(q)) {
        o.u[wg_off(q) + pos] = w;
        if constexpr (kPre) {
          o.b[wg_off(q) + pos] = rb;
          o.e[wg_off(q) + pos] = re;
        }
      } else {
        spill = true;
      }
    }
  }
  append(o.dev, w, spill, cw, o.nclass);
  if (spill) *o.ovf = 1;
}

// T lanes (lane = 0..T-1) expand frontier vertex u, whose slot range is [rb, re) for one-segment
// layouts; est collects the claimed vertices' estimated slots.
template <int T, class L>
__device__ __forceinline__ void wg_expand(const L &g, const BfsArgs &a, const WgOut &o, int u,
                                          int rb, int re, int lane, int level, long long &ns,
                                          long long &est) {
  constexpr bool kPre = L::kSeg == 1;
  constexpr int B = kBfsBatch;
#pragma unroll 1
  for (int seg = 0; seg < L::kSeg; ++seg) {
    const int2 r = kPre ? make_int2(rb, re) : g.range(u, seg);
    if (lane == 0) ns += r.y - r.x;
    for (int base = r.x + lane; base < r.y; base += T * B) {
      int w[B], idx[B], rin[B], hw[B], ew[B], cw[B], aw[B], wb[B], we[B];
      bool claim[B];
#pragma unroll
      for (int k = 0; k < B; ++k) {
        w[k] = -1;
        idx[k] = 0;
        if (base + k * T < r.y) g.bfs_slot(seg, base + k * T, w[k], idx[k]);
      }
#pragma unroll
      for (int k = 0; k < B; ++k) {
        const bool ok = w[k] >= 0 && w[k] != u;
        rin[k] = ok ? g.bfs_rin(seg, idx[k]) : 0;
        hw[k] = ok ? a.h[w[k]] : -1;
        ew[k] = ok ? a.e[w[k]] : 0;
        cw[k] = (ok && a.vcls) ? a.vcls[w[k]] : 0;
        aw[k] = (ok && a.acls) ? a.acls[w[k]] : 0;
      }
#pragma unroll
      for (int k = 0; k < B; ++k) {
        const bool cand = rin[k] > 0 && hw[k] == a.H;
        wb[k] = 0;
        we[k] = 0;
        if constexpr (kPre) {
          if (cand) {  // the range travels with the entry: loaded beside the claim
            const int2 rr = g.range(w[k], 0);
            wb[k] = rr.x;
            we[k] = rr.y;
          }
        }
        if (rin[k] > 0 && hw[k] == a.H + 1) a.flags[0] = 1;  // residual path from a source
        claim[k] = cand && atomicCAS(&a.h[w[k]], a.H, level + 1) == a.H;
        if (claim[k]) est += pick(a.cls_mean, cw[k]);
      }
#pragma unroll
      for (int k = 0; k < B; ++k) {
        wg_append<kPre>(o, w[k], claim[k], cw[k], wb[k], we[k]);
        append(o.act, w[k], claim[k] && ew[k] > 0, aw[k], a.nclass_act);
      }
    }
  }
}

template <class L>
__global__ __launch_bounds__(kWgBlock) void k_bfs_wg(L g, BfsArgs a, Dispatch d, int level,
                                                     int fmax) {
  constexpr bool kPre = L::kSeg == 1;
  __shared__ int s_u[2][kWgCap];
  __shared__ int s_b[kPre ? 2 : 1][kPre ? kWgCap : 1];
  __shared__ int s_e[kPre ? 2 : 1][kPre ? kWgCap : 1];
  __shared__ int s_n[2][3];
  __shared__ int s_ocnt[kNumClasses];
  __shared__ int s_acnt[kNumClasses];  // the cycle worklist's counters while this kernel runs
  __shared__ int s_ovf, s_go, s_size;
  __shared__ long long s_est;
  const int tid = (int)threadIdx.x;
  long long nv = 0, ns = 0;
  fmax = min(fmax, kWgCap);
  // take over the current frontier from the device lists if it fits
  ListSet in = a.fr[0];
  in.items = (level & 1) ? a.fr[1].items : a.fr[0].items;
  in.cnt = pick(a.fcnt, level % 3);
  if (tid < 3) {
    s_n[0][tid] = 0;
    s_n[1][tid] = 0;
  }
  // nothing else appends to the cycle worklist while this kernel runs: its counters move to LDS
  // (no returning global atomic per appending wave) and back at the end
  if (tid < kNumClasses) s_acnt[tid] = a.active.cnt[tid * a.active.cstride];
  if (tid == 0) {
    int n[3] = {0, 0, 0}, sz = 0;
    long long slots = 0;  // estimated from the class means: hubs make a few vertices heavy
    for (int c = 0; c < d.nclass; ++c) {
      const int k = list_count(in, c);
      sz += k;
      slots += (long long)k * pick(a.cls_mean, c);
      const int q = wg_group(c);
      n[q] += k;
    }
    s_size = sz;
    s_go = sz > 0 && sz <= fmax && slots <= a.wg_slots && n[0] <= kWgCapA && n[1] <= kWgCapB &&
           n[2] <= kWgCapC;
  }
  __syncthreads();
  if (!s_go) {  // leave the device lists as they are
    if (tid == 0) {
      a.state[0] = level;
      a.state[1] = s_size;
    }
    return;
  }
  for (int c = 0; c < d.nclass; ++c) {
    const int k = list_count(in, c), q = wg_group(c), off = in.cls[c];
    for (int i = tid; i < k; i += kWgBlock) {
      const int u = in.items[off + i];
      const int pos = wg_off(q) + atomicAdd(&s_n[0][q], 1);  // fits: checked above
      s_u[0][pos] = u;
      if constexpr (kPre) {
        const int2 r = g.range(u, 0);
        s_b[0][pos] = r.x;
        s_e[0][pos] = r.y;
      }
    }
  }
  int p = 0;
  for (;;) {
    // level `level` sits in buffer p; its claims go to buffer p ^ 1 (or the device list)
    if (tid < 3) s_n[p ^ 1][tid] = 0;
    if (tid < kNumClasses) s_ocnt[tid] = 0;
    if (tid == 0) {
      s_ovf = 0;
      s_est = 0;
    }
    __syncthreads();
    const int nA = min(s_n[p][0], kWgCapA), nB = min(s_n[p][1], kWgCapB),
              nC = min(s_n[p][2], kWgCapC);
    WgOut o;
    o.u = s_u[p ^ 1];
    o.b = kPre ? s_b[p ^ 1] : nullptr;
    o.e = kPre ? s_e[p ^ 1] : nullptr;
    o.n = s_n[p ^ 1];
    o.dev = a.fr[0];
    o.dev.items = (level & 1) ? a.fr[0].items : a.fr[1].items;  // the next level's item array
    o.dev.cnt = s_ocnt;
    o.dev.cstride = 1;
    o.act = a.active;
    o.act.cnt = s_acnt;
    o.act.cstride = 1;
    o.ovf = &s_ovf;
    o.nclass = a.nclass;
    long long est = 0;
    const int* const cu = s_u[p];
    const int* const cb = kPre ? s_b[p] : nullptr;
    const int* const ce = kPre ? s_e[p] : nullptr;
    for (int i = tid / 4; i < nA; i += kWgBlock / 4) {
      const int x = wg_off(0) + i;
      wg_expand<4>(g, a, o, cu[x], kPre ? cb[x] : 0, kPre ? ce[x] : 0, tid % 4, level, ns, est);
      if (tid % 4 == 0) ++nv;
    }
    for (int i = tid / 64; i < nB; i += kWgBlock / 64) {
      const int x = wg_off(1) + i;
      wg_expand<64>(g, a, o, cu[x], kPre ? cb[x] : 0, kPre ? ce[x] : 0, tid % 64, level, ns, est);
      if (tid % 64 == 0) ++nv;
    }
    for (int i = 0; i < nC; ++i) {
      const int x = wg_off(2) + i;
      wg_expand<kWgBlock>(g, a, o, cu[x], kPre ? cb[x] : 0, kPre ? ce[x] : 0, tid, level, ns, est);
      if (tid == 0) ++nv;
    }
    est = tile_sum<64>(est);
    if (__lane_id() == 0 && est) atomicAdd((unsigned long long *)&s_est, (unsigned long long)est);
    __syncthreads();
    ++level;
    p ^= 1;
    // the next frontier (now `level`) stays in LDS only if nothing spilled and it is small
    if (tid == 0) {
      const int sz = s_n[p][0] + s_n[p][1] + s_n[p][2];
      s_size = sz;
      s_go = !s_ovf && sz > 0 && sz <= fmax && s_est <= a.wg_slots;
    }
    __syncthreads();
    if (s_go) continue;
    break;
  }
  // hand the frontier of `level` to the device lists: what spilled is there already, the LDS
  // entries follow (classes read back from the class array)
  {
    ListSet dev = a.fr[0];
    dev.items = (level & 1) ? a.fr[1].items : a.fr[0].items;
    dev.cnt = s_ocnt;
    dev.cstride = 1;
    for (int q = 0; q < 3; ++q) {
      const int k = min(s_n[p][q], wg_cap(q));
      for (int i0 = 0; i0 < k; i0 += kWgBlock) {
        const int i = i0 + tid;
        const bool want = i < k;
        const int u = want ? s_u[p][wg_off(q) + i] : 0;
        append(dev, u, want, (want && a.vcls) ? a.vcls[u] : 0, a.nclass);
      }
    }
  }
  __syncthreads();
  if (tid < kNumClasses) {
    const int c = tid;
    a.active.cnt[c * a.active.cstride] = s_acnt[c];
    pick(a.fcnt, level % 3)[c * kCtlStride] = s_ocnt[c];
    pick(a.fcnt, (level + 1) % 3)[c * kCtlStride] = 0;
    pick(a.fcnt, (level + 2) % 3)[c * kCtlStride] = 0;
  }
  if (tid == 0) {
    int sz = 0;
    for (int c = 0; c < kNumClasses; ++c) sz += s_ocnt[c];
    a.state[0] = level;
    a.state[1] = sz;
  }
  block_flush(nv, ns, 0, 0, a.stats, ST_BFS_VERTS);
}

// ---- reset, preflow, results ---------------------------------------------------------------

__global__ void k_reset_pairs(int2 *pairs, const int *cap0, long long count, int field) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long j = blockIdx.x * (long long)blockDim.x + threadIdx.x; j < count; j += stride) {
    if (field == 0) pairs[j].x = cap0[j];
    else pairs[j].y = cap0[j];
  }
}

// The same restore over the lines of slot pairs marked dirty by the last solve's pushes
// (dirty[i] covers slots 8i..8i+7); clears the marks.  A full line is rewritten whole.
__global__ void k_reset_dirty(int2 *pairs, const int *cap0, unsigned char *dirty, long long lines,
                              long long count, int field) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < lines; i += stride) {
    if (!dirty[i]) continue;
    dirty[i] = 0;
    const long long j0 = i * 8;
    if (j0 + 8 <= count) {
      int4 *line = reinterpret_cast<int4 *>(pairs + j0);  // 64-byte aligned: j0 % 8 == 0
      const int4 *cap = reinterpret_cast<const int4 *>(cap0 + j0);
      int4 q[4];
      const int4 c0 = cap[0], c1 = cap[1];
#pragma unroll
      for (int k = 0; k < 4; ++k) q[k] = line[k];
      const int cv[8] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w};
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        if (field == 0) {
          q[k].x = cv[2 * k];
          q[k].z = cv[2 * k + 1];
        } else {
          q[k].y = cv[2 * k];
          q[k].w = cv[2 * k + 1];
        }
        line[k] = q[k];
      }
    } else {
      for (long long j = j0; j < count; ++j) {
        if (field == 0) pairs[j].x = cap0[j];
        else pairs[j].y = cap0[j];
      }
    }
  }
}

__global__ void k_reset_vertices(int *h, int *e, int *stamp, int n, int *bcf, long long m) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long v = blockIdx.x * (long long)blockDim.x + threadIdx.x; v < n; v += stride) {
    h[v] = 0;
    e[v] = 0;
    stamp[v] = -1;
  }
  if (bcf)
    for (long long j = blockIdx.x * (long long)blockDim.x + threadIdx.x; j < m; j += stride)
      bcf[j] = 0;
}

// preflow (maxflow-cuda/src/preflow.cu:3-31), one block per source: saturate every residual arc
// out of the source.  Parallel s -> v arcs add to e(v) (the reference overwrites it).
template <class L>
__global__ __launch_bounds__(kBlock) void k_preflow(L g, int *e, const int *src, int k,
                                                    unsigned long long *stats) {
  long long sum = 0;
  // a group of blocks per source: a single source of high degree uses the whole grid
  const BlockGroups bg = block_groups((int)blockIdx.x, (int)gridDim.x, k);
  for (int i = bg.group; i < k; i += bg.ngroups) {
    const int s = src[i];
#pragma unroll
    for (int seg = 0; seg < L::kSeg; ++seg) {
      const int2 r = g.range(s, seg);
      for (int j = r.x + bg.lane; j < r.y; j += bg.width) {
        int w, res, key;
        g.scan(seg, j, w, res, key);
        if (res > 0 && w != s) {
          g.push(key, res);
          atomicAdd(&e[w], res);
          sum += res;
        }
      }
    }
  }
  block_flush(sum, 0, 0, 0, stats, ST_SOURCE_OUT);
}

__global__ void k_set_int(int *p, int v) { *p = v; }

__global__ void k_fill_int(int *p, int v, int count) {
  for (int i = (int)threadIdx.x; i < count; i += (int)blockDim.x) p[i] = v;
}

__global__ void k_gather_values(const int *e, const int *snk, int k, long long *out) {
  const int stride = gridDim.x * blockDim.x;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < k; i += stride) out[i] = e[snk[i]];
}

// ---- batched solve: whole instances per workgroup --------------------------------------------
//
// A persistent grid of workgroups pulls instances from a device queue (largest first) and solves
// each one entirely inside the workgroup: preflow, global relabels (BFS level by level) and
// push-relabel cycles, with workgroup barriers in place of kernel launches and all worklist
// counters in LDS.  Only this workgroup touches the instance's vertices and slots, so the
// barriers give the same ordering the single-instance path gets from kernel boundaries.  The
// relabel cadence is the adaptive one of the single-instance path, timed with wall_clock64.

struct BatchArgs {
  int *h, *e;
  int *lists;                      // 4 sets x N items (instance regions at their vertex ranges)
  int N;
  const int *vbase, *nver, *src, *snk;
  const int *pop;                  // k x kNumClasses, cycle classes
  const int *bpop;                 // k x kNumClasses, BFS (VC) classes
  const unsigned char *vcls;       // VC class of every vertex
  const int *order;                // instances, largest first
  int k;
  int *next;                       // queue head
  long long *values;
  int *status;                     // per instance: 0 ok, 1 round limit, 2 invariant broken
  int *ovf;
  unsigned long long *stats;
  int nclass;                      // cycle classes (1 for TC)
  int lanes[kNumClasses];          // cycle lanes per class
  int blanes[kNumClasses];         // BFS (VC) lanes per class
  int use_vcls;                    // cycle classes follow vcls
  float ratio;                     // cycle time / relabel time per round
  int max_rounds;
  int max_cycles;                  // per round
  int fwd_hops;                    // push forwarding hops per cycle
  int *stamp;                      // per vertex: last cycle that discharged it
};

template <class L>
__device__ __forceinline__ void batch_cycle_class(const L &g, const CycleArgs &ca, int c, int lanes,
                                                  int cnt, Stats &st) {
  const int base = ca.in.cls[c];
  const int tid = (int)threadIdx.x;
  switch (lanes) {
#define MF_BATCH_CASE(TT)                                                          \
  case TT:                                                                         \
    for (int i = tid / TT; i < cnt; i += kBlock / TT) {                             \
      int v = ca.in.items[base + i];                                               \
      for (int hop = 0; v >= 0; ++hop)                                             \
        v = discharge<TT>(g, ca, nullptr, v, tid % TT, hop < ca.fwd_hops ? c : -1,  \
                          hop == 0 && ca.fwd_hops > 0, st);                        \
    }                                                                              \
    break;
    MF_BATCH_CASE(1)
    MF_BATCH_CASE(2)
    MF_BATCH_CASE(4)
    MF_BATCH_CASE(8)
    MF_BATCH_CASE(16)
    MF_BATCH_CASE(32)
    MF_BATCH_CASE(64)
#undef MF_BATCH_CASE
    default:
      for (int i = 0; i < cnt; ++i) discharge_block(g, ca, nullptr, ca.in.items[base + i], st);
      break;
  }
}

// One instance of a batch: vertices [vb, vb + n), source s and sink t (ids in the packed
// space), and its class populations (cycle classes, BFS classes; kNumClasses each).
struct BatchInst {
  int vb, n, s, t;
  const int *pop, *bpop;
};

// Solves one instance inside the calling workgroup; returns its value (the sink's excess) and
// *status_out (0 ok, 1 round limit, 2 a source was reached: invariant broken).
template <class L>
__device__ __forceinline__ long long batch_instance(const L &g, const BatchArgs &b,
                               const BatchInst &I, int *status_out,
                               int (*s_f)[kNumClasses], int (*s_c)[kNumClasses], int *s_flag,
                               int *s_dec, int (*s_tab)[2 * kNumClasses], Stats &st,
                               long long &nv, long long &ns) {
  const int tid = (int)threadIdx.x;
  const int vb = I.vb, n = I.n, s = I.s, t = I.t;
  const int H = n;  // heights are distances inside the instance
  // class regions of the instance's frontier lists (table 0) and cycle worklists (table 1) in
  // the packed list arrays, in LDS
  if (tid < 2) {
    const int *pp = tid == 0 ? I.bpop : I.pop;
    const int nc = tid == 0 ? kNumClasses : b.nclass;
    int acc = vb;
    for (int c = 0; c < kNumClasses; ++c) {
      s_tab[tid][c] = acc;
      s_tab[tid][kNumClasses + c] = c < nc ? pp[c] : 0;
      if (c < nc) acc += pp[c];
    }
  }
  __syncthreads();
  // frontier lists F0, F1 and cycle worklists S0, S1
  auto make_set = [&](int q) {
    ListSet ls;
    ls.items = b.lists + (size_t)q * b.N;
    ls.cls = s_tab[q < 2 ? 0 : 1];
    ls.cnt = nullptr;
    ls.cstride = 1;  // the workgroup's counters live in LDS
    ls.ovf = b.ovf;
    return ls;
  };
  const ListSet F0 = make_set(0), F1 = make_set(1), S0 = make_set(2), S1 = make_set(3);
  BfsArgs ba;
  ba.h = b.h;
  ba.e = b.e;
  ba.H = H;
  ba.vcls = b.vcls;
  ba.nclass = kNumClasses;
  ba.acls = b.use_vcls ? b.vcls : nullptr;
  ba.nclass_act = b.nclass;
  ba.fr[0] = F0;
  ba.fr[1] = F1;
  for (int q = 0; q < 3; ++q) ba.fcnt[q] = s_f[q];
  ba.active = S0;
  ba.active.cnt = s_c[0];
  ba.flags = s_flag;
  ba.state = nullptr;
  ba.stats = nullptr;
  for (int c = 0; c < kNumClasses; ++c) ba.cls_mean[c] = 0;  // (used by the grid-wide BFS only)
  ba.wg_slots = INT_MAX;
  ba.rem_items[0] = ba.rem_items[1] = nullptr;
  ba.rem_cls = nullptr;
  for (int q = 0; q < 3; ++q) ba.rcnt[q] = nullptr;
  CycleArgs ca;
  ca.h = b.h;
  ca.e = b.e;
  ca.H = H;
  ca.vcls = b.use_vcls ? b.vcls : nullptr;
  ca.nclass = b.nclass;
  ca.clear_cnt = nullptr;
  ca.stats = nullptr;
  ca.fwd_hops = b.fwd_hops;
  ca.stamp = b.stamp;
  ca.cycle_base = nullptr;
  ca.work = nullptr;
  ca.work_cls = nullptr;
  ca.giant = nullptr;  // giants are discharged by the whole workgroup like other hubs
  ca.cycle = 0;

  // preflow: saturate the source's residual arcs
#pragma unroll
  for (int seg = 0; seg < L::kSeg; ++seg) {
    const int2 r = g.range(s, seg);
    for (int j = r.x + tid; j < r.y; j += kBlock) {
      int w, res, key;
      g.scan(seg, j, w, res, key);
      if (res > 0 && w != s) {
        g.push(key, res);
        atomicAdd(&b.e[w], res);
      }
    }
  }
  __syncthreads();
  int status = 0;
  for (int round = 0;; ++round) {
    if (round >= b.max_rounds) {
      status = 1;
      break;
    }
    // ---- global relabel inside the workgroup ----
    const long long t_bfs = wall_clock64();
    for (int v = vb + tid; v < vb + n; v += kBlock) b.h[v] = H;
    if (tid < 3 * kNumClasses) {
      s_f[tid / kNumClasses][tid % kNumClasses] = 0;
      s_c[tid / kNumClasses][tid % kNumClasses] = 0;
    }
    if (tid == 0) s_flag[0] = 0;
    __syncthreads();
    if (tid == 0) {
      b.h[t] = 0;
      b.h[s] = H + 1;
      ListSet f0 = F0;
      f0.cnt = s_f[0];
      append(f0, t, true, b.vcls[t], kNumClasses);
    }
    __syncthreads();
    for (int level = 0;; ++level) {
      if (tid == 0) {
        int sz = 0;
        for (int c = 0; c < kNumClasses; ++c) sz += s_f[level % 3][c];
        *s_dec = sz;
      }
      if (tid < kNumClasses) s_f[(level + 2) % 3][tid] = 0;
      __syncthreads();
      const int size = *s_dec;
      __syncthreads();
      if (size == 0) break;
      ListSet in = (level & 1) ? F1 : F0;
      in.cnt = s_f[level % 3];
      ListSet out = (level & 1) ? F0 : F1;
      out.cnt = s_f[(level + 1) % 3];
      for (int c = 0; c < kNumClasses; ++c) {
        const int cnt = list_count(in, c);
        if (cnt)
          bfs_dispatch<false>(g, ba, in, out, BfsStages{nullptr, nullptr}, c, pick(b.blanes, c), cnt,
                       tid, kBlock, level, nv, ns);
      }
      __syncthreads();
    }
    if (s_flag[0]) {
      status = 2;
      break;
    }
    int active = 0;
    for (int c = 0; c < b.nclass; ++c) active += s_c[0][c];
    if (active == 0) break;
    const long long bfs_ticks = wall_clock64() - t_bfs;
    // ---- push-relabel cycles ----
    const long long t_cyc = wall_clock64();
    for (int q = 0; q < b.max_cycles; ++q) {
      ++ca.cycle;  // increasing over the whole solve (stamps of earlier rounds never match)
      ca.in = (q & 1) ? S1 : S0;
      ca.in.cnt = s_c[q % 3];
      ca.out = (q & 1) ? S0 : S1;
      ca.out.cnt = s_c[(q + 1) % 3];
      if (tid < kNumClasses) s_c[(q + 2) % 3][tid] = 0;  // last read in cycle q - 1
      for (int c = 0; c < b.nclass; ++c) {
        const int cnt = list_count(ca.in, c);
        if (cnt) batch_cycle_class(g, ca, c, pick(b.lanes, c), cnt, st);
      }
      __syncthreads();
      if (tid == 0) {
        int left = 0;
        for (int c = 0; c < b.nclass; ++c) left += s_c[(q + 1) % 3][c];
        const long long spent = wall_clock64() - t_cyc;
        *s_dec = (left == 0 || (double)spent >= b.ratio * (double)bfs_ticks) ? 1 : 0;
      }
      __syncthreads();
      const int stop = *s_dec;
      __syncthreads();
      if (stop) break;
    }
  }
  const long long value = b.e[t];  // every thread: the last cycle's atomics are behind a barrier
  *status_out = status;
  __syncthreads();
  return value;
}

template <class L>
__global__ __launch_bounds__(kBlock) void k_batch_wg(L g, BatchArgs b) {
  __shared__ int s_f[3][kNumClasses];
  __shared__ int s_c[3][kNumClasses];
  __shared__ int s_flag[2];
  __shared__ int s_dec;
  __shared__ int s_inst;
  __shared__ int s_tab[2][2 * kNumClasses];
  Stats st{0, 0, 0, 0};
  long long nv = 0, ns = 0;
  for (;;) {
    if (threadIdx.x == 0) s_inst = atomicAdd(b.next, 1);
    __syncthreads();
    const int i = s_inst;
    __syncthreads();
    if (i >= b.k) break;
    const int inst = b.order[i];
    const BatchInst I{b.vbase[inst], b.nver[inst], b.src[inst], b.snk[inst],
                      b.pop + (size_t)inst * kNumClasses, b.bpop + (size_t)inst * kNumClasses};
    int status = 0;
    const long long value = batch_instance(g, b, I, &status, s_f, s_c, s_flag, &s_dec, s_tab, st, nv, ns);
    if (threadIdx.x == 0) {
      b.values[inst] = value;
      b.status[inst] = status;
    }
    block_flush(st.active, st.slots, st.pushes, st.relabels, b.stats, ST_ACTIVE);
    st = Stats{0, 0, 0, 0};
  }
  block_flush(st.active, st.slots, st.pushes, st.relabels, b.stats, ST_ACTIVE);
  block_flush(nv, ns, 0, 0, b.stats, ST_BFS_VERTS);
}

// ---- serving: instances admitted while the kernel runs ---------------------------------------
//
// A persistent grid that serves a queue the host keeps filling.  The host copies an admitted
// request's residual slots, offsets and vertex classes into its ranges of the server's device
// arena, then publishes queue entry q: it writes the entry (vertex range, terminals, class
// populations) and then the published count, both in host memory that the GPU reads uncached.
// A workgroup claims the next published entry on a device counter, solves the instance exactly
// like a batched one (batch_instance), and writes the value, status and device times into the
// entry's report in host memory, the entry number last.  An idle workgroup sleeps between looks
// at the published count and leaves once it has been idle for idle_ticks, or when the host closes
// the queue, so the grid drains by itself; the host launches it again when entries are published
// that no running workgroup will take.
//
// Arena ranges are reused by later requests, and those copies happen while a workgroup may hold
// lines of the range in its caches: a workgroup invalidates its caches (agent-scope acquire)
// before it reads an admitted instance, and writes its dirty lines back (agent-scope release)
// before it reports the instance, after which the host may copy a new request over the range.
// Idle workgroups read the host's published count at most once per kHostPollTicks between them
// (the last value read is kept in device memory), so the grid does not flood the host link.

struct ServeEntry {
  int vb, n, s, t;        // vertex range [vb, vb + n) in the arena, source, sink
  int pop[kNumClasses];   // cycle-class populations of the instance
  int bpop[kNumClasses];  // BFS-class populations
};
constexpr int kServeEntryInts = 4 + 2 * kNumClasses;
static_assert(sizeof(ServeEntry) == sizeof(int) * kServeEntryInts, "ServeEntry is plain ints");

struct ServeReport {
  long long value;
  long long t_start, t_end;  // wall_clock64 ticks (100 MHz) around the instance's solve
  int status;                // 0 ok, 1 round limit, 2 invariant broken
  int seq;                   // entry number + 1, written last
};

struct ServeArgs {
  BatchArgs b;                 // arena arrays and solver options (per-instance fields unused)
  const int *entries;          // host memory: qcap entries of kServeEntryInts ints
  ServeReport *reports;        // host memory: qcap reports
  const int *published;        // host memory: entries published so far
  const int *closed;           // host memory: nonzero asks idle workgroups to leave
  int *claimed;                // device memory: entries claimed so far
  int *pub_seen;               // device memory: the largest published count read so far
  long long *poll_at;          // device memory: wall_clock64 of the last host read
  long long *last_claim;       // device memory: wall_clock64 of the last claim or finish
  int *busy;                   // device memory: workgroups solving an instance now
  int qcap;
  long long idle_ticks;        // ticks without any claim or finish after which idle workgroups
                               // leave (none may be busy)
};
constexpr long long kHostPollTicks = 500;  // 5 us at the 100 MHz wall clock
constexpr int kServeCoreStride = 4;        // one serving workgroup in four stays while any works

// Published entries, as one idle workgroup's thread sees them: the host's count when no
// workgroup has read it for kHostPollTicks, else the last count read.
__device__ __forceinline__ int serve_published(const ServeArgs &a) {
  const long long now = (long long)wall_clock64();
  unsigned long long last =
      __hip_atomic_load((unsigned long long *)a.poll_at, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (now - (long long)last >= kHostPollTicks &&
      atomicCAS((unsigned long long *)a.poll_at, last, (unsigned long long)now) == last) {
    const int pub = __hip_atomic_load(a.published, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_fetch_max(a.pub_seen, pub, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    return pub;
  }
  return __hip_atomic_load(a.pub_seen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}

template <class L>
__global__ __launch_bounds__(kBlock) void k_serve_wg(L g, ServeArgs a) {
  __shared__ int s_f[3][kNumClasses];
  __shared__ int s_c[3][kNumClasses];
  __shared__ int s_flag[2];
  __shared__ int s_dec;
  __shared__ int s_q;
  __shared__ int s_tab[2][2 * kNumClasses];
  __shared__ int s_entry[kServeEntryInts];
  const int tid = (int)threadIdx.x;
  Stats st{0, 0, 0, 0};
  long long nv = 0, ns = 0;
  long long idle_from = (long long)wall_clock64();
  for (;;) {
    if (tid == 0) {
      int q = -1;
      const int pub = serve_published(a);
      int c = __hip_atomic_load(a.claimed, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      while (c < pub) {  // every failed exchange means another workgroup claimed entry c
        const int old = atomicCAS(a.claimed, c, c + 1);
        if (old == c) {
          q = c;
          break;
        }
        c = old;
      }
      const long long now = (long long)wall_clock64();
      if (q >= 0) {
        atomicAdd(a.busy, 1);
        __hip_atomic_fetch_max(a.last_claim, now, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      // Leaving.  A workgroup leaves after idle_ticks without work of its own, except the core
      // (every kServeCoreStride-th workgroup), which leaves only when no workgroup is solving an
      // instance and none has claimed or finished one for idle_ticks.  Requests arriving one by
      // one, or beside a long instance, then always find the core, not only the workgroups that
      // happened to be busy; and the idle rest does not hold registers that solves running
      // alone beside the serving kernel can use.
      if (q < 0) {
        bool quiet = now - idle_from > a.idle_ticks;
        if (quiet && blockIdx.x % kServeCoreStride == 0) {
          const long long last = __hip_atomic_load(a.last_claim, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT);
          quiet = __hip_atomic_load(a.busy, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0 &&
                  now - last > a.idle_ticks;
        }
        if (quiet || __hip_atomic_load(a.closed, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM))
          q = -2;
      }
      s_q = q;
    }
    __syncthreads();
    const int q = s_q;
    __syncthreads();
    if (q == -2) break;
    if (q == -1) {
      __builtin_amdgcn_s_sleep(127);
      __builtin_amdgcn_s_sleep(127);
      continue;
    }
    const long long t0 = (long long)wall_clock64();
    const int slot = q % a.qcap;
    if (tid < kServeEntryInts)
      s_entry[tid] = __hip_atomic_load(a.entries + (size_t)slot * kServeEntryInts + tid,
                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    // the request's arena ranges were written by a copy after this kernel may have cached them
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    __syncthreads();
    const BatchInst I{s_entry[0], s_entry[1], s_entry[2], s_entry[3], s_entry + 4,
                      s_entry + 4 + kNumClasses};
    // host state of the instance's vertices: no excess, never discharged
    for (int v = I.vb + tid; v < I.vb + I.n; v += kBlock) {
      a.b.e[v] = 0;
      a.b.stamp[v] = -1;
    }
    __syncthreads();
    int status = 0;
    const long long value =
        batch_instance(g, a.b, I, &status, s_f, s_c, s_flag, &s_dec, s_tab, st, nv, ns);
    // every write of the instance reaches memory before the host may reuse its ranges
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    if (tid == 0) {
      ServeReport *r = a.reports + slot;
      __hip_atomic_store(&r->value, value, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      __hip_atomic_store(&r->t_start, t0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      __hip_atomic_store(&r->t_end, (long long)wall_clock64(), __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_SYSTEM);
      __hip_atomic_store(&r->status, status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      __hip_atomic_store(&r->seq, q + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
      __hip_atomic_fetch_max(a.last_claim, (long long)wall_clock64(), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
      atomicSub(a.busy, 1);
    }
    block_flush(st.active, st.slots, st.pushes, st.relabels, a.b.stats, ST_ACTIVE);
    st = Stats{0, 0, 0, 0};
    idle_from = (long long)wall_clock64();
  }
  block_flush(st.active, st.slots, st.pushes, st.relabels, a.b.stats, ST_ACTIVE);
  block_flush(nv, ns, 0, 0, a.b.stats, ST_BFS_VERTS);
}

}  // namespace mf

// ============================================================================================
// host side
// ============================================================================================
namespace mf {

static thread_local char g_last_error[256];

const char *last_device_error() { return g_last_error; }

#define HIP_TRY(expr)                                                                         \
  do {                                                                                        \
    hipError_t _e = (expr);                                                                   \
    if (_e != hipSuccess) {                                                                   \
      snprintf(g_last_error, sizeof(g_last_error), "%s:%d %s", __FILE__, __LINE__,            \
               hipGetErrorString(_e));                                                        \
      return _e == hipErrorOutOfMemory ? MF_ERR_OUT_OF_MEMORY : MF_ERR_DEVICE;                \
    }                                                                                         \
  } while (0)

static double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

// Global-relabel cadence.  Adaptive by default: after a global relabel that took T ms, cycles
// run until about ratio * T ms have passed (checked between chunks of cycles), so the two phases
// stay balanced whether relabels are cheap (small graphs) or dominated by long BFS chains.
// Excess that can no longer reach a sink only stops moving at the next global relabel, which is
// why a fixed large cycle count is wasteful on graphs that strand excess (bipartite, RLG).
constexpr float kDefaultRelabelRatio = 2.0f;
constexpr int kMaxAdaptiveCycles = 1 << 14;  // cap of cycles per round in adaptive mode
constexpr int kDefaultBfsFrontier = 4096;    // one-workgroup BFS up to this frontier size
constexpr int kWgBfsSlots = 32768;           // ... and up to about this many frontier slots
constexpr int kGridLevelBatch = 4;           // top-down grid levels issued per read-back
constexpr int kBottomUpBatch = 3;            // bottom-up levels issued per read-back
constexpr int kDefaultForwardHops = 2;       // push forwarding: vertices a tile chains per cycle
constexpr int kCycleChunk = 64;              // most cycles launched between two worklist checks
// graphs of at least this many residual slots track the lines pushes change, so a re-solve
// restores only those (a full restore of R-MAT 22's 130M slots takes 0.55 ms)
constexpr int64_t kDirtyMinSlots = int64_t(1) << 24;
constexpr int kFirstChunk = 8;               // first chunk, before a cycle has been timed
constexpr int kGraphPeriod = 6;              // cycles after which lists and counters repeat
constexpr int kGraphChunkMax = 60;           // longest captured chunk
// Direction of the grid-wide BFS levels (Beamer's rule): bottom-up once the frontier's slots
// exceed 1/kBottomUpAlpha of the slots not yet explored, top-down again once the frontier holds
// fewer than 1/kBottomUpBeta of the nonzero-degree vertices; never below kBottomUpMinVertices.
constexpr int64_t kBottomUpAlpha = 14;
constexpr int64_t kBottomUpBeta = 24;
constexpr int64_t kBottomUpMinVertices = 1 << 16;
// Lanes per class of the bottom-up levels (a tile of lanes / 4 BFS lanes, 1 for 4), for the
// classes <= 8, <= 16, <= 32 and <= 1,024 slots.  R-MAT 22 relabels: 2.03 ms with the top-down
// levels' 8 / 16 / 32 / 64, 1.81 ms with these; other splits within 1.79-1.81 ms.
constexpr int kBottomUpLanes[4] = {4, 8, 16, 32};
constexpr int kBatchGroupsPerCU = 4;         // at most this many batched workgroups per CU
// Top-down grid levels: blocks per CU per class.  Most of a batch's levels after the first find
// nothing to do (LevelStop), and every block of such a launch still starts and reads the
// counters: R-MAT 22 solves take 7.63 / 7.48 / 7.45 / 7.40 ms at 8 / 4 / 2 / 1 blocks, Genrmf 64
// is level within its spread.
constexpr int kBfsBlocksPerCU = 2;
// Cooperative top-down levels (k_bfs_coop): workgroups of the launch, and the largest frontier
// (estimated slots) they take; larger frontiers go to one grid-wide launch per level.
constexpr int kCoopBlocks = 64;
constexpr long long kCoopSlots = 1 << 17;

template <class T>
static mf_status dmalloc(T **p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  HIP_TRY(hipMalloc((void **)p, sizeof(T) * count));
  return MF_OK;
}

static_assert(kClasses == kNumClasses, "host and device class counts differ");
static_assert(kTunedClasses == sizeof(((mf_options *)nullptr)->class_lanes) / sizeof(uint16_t),
              "mf_options.class_lanes covers the tunable classes");

ClassConfig class_config(mf_algo algo, int tile_lanes, const uint16_t *class_lanes) {
  ClassConfig c;
  if (algo == MF_ALGO_TC) {
    c.nclass = 1;
    c.lanes[0] = 1;
    c.maxdeg[0] = INT32_MAX;
  } else if (tile_lanes > 0) {
    c.nclass = 1;
    c.lanes[0] = tile_lanes;
    c.maxdeg[0] = INT32_MAX;
  } else {
    c.nclass = kClasses;
    const int lanes[kClasses] = {8, 16, 32, 64, kBlock, kSplitLanes};
    const int64_t maxdeg[kClasses] = {8, 16, 32, 1024, kGiantSlots, INT64_MAX};
    for (int k = 0; k < kClasses; ++k) {
      c.lanes[k] = (k < kTunedClasses && class_lanes && class_lanes[k]) ? class_lanes[k] : lanes[k];
      c.maxdeg[k] = maxdeg[k];
    }
  }
  return c;
}

int class_of_degree(const ClassConfig &c, int64_t deg) {
  for (int k = 0; k < c.nclass; ++k)
    if (deg <= c.maxdeg[k]) return k;
  return c.nclass - 1;
}

// Blocks per class: enough tiles for the class population, at most per_cu per CU; the split
// class (giants) gets one block per part (split_units parts in all).
static Dispatch make_dispatch(const ClassConfig &c, const int64_t *pop, int num_cu,
                              int64_t split_units, int per_cu = 8) {
  Dispatch d;
  memset(&d, 0, sizeof(d));
  d.nclass = c.nclass;
  int b = 0;
  for (int k = 0; k < c.nclass; ++k) {
    d.lanes[k] = c.lanes[k];
    d.blk_begin[k] = b;
    if (pop[k] == 0) continue;
    int64_t nb;
    if (c.lanes[k] <= 64) {
      const int64_t tpb = kBlock / c.lanes[k];
      nb = std::min<int64_t>((pop[k] + tpb - 1) / tpb, (int64_t)per_cu * num_cu);
    } else if (c.lanes[k] == kSplitLanes) {
      nb = std::min<int64_t>(std::max<int64_t>(split_units, pop[k]), (int64_t)per_cu * num_cu);
    } else {
      nb = std::min<int64_t>(pop[k], (int64_t)per_cu * num_cu);
    }
    b += (int)std::max<int64_t>(nb, 1);
  }
  for (int k = c.nclass; k <= kNumClasses; ++k) d.blk_begin[k] = b;
  return d;
}

int64_t algorithmic_bytes_cycle(mf_layout layout, int64_t active, int64_t slots, int64_t pushes,
                                int64_t relabels) {
  // per active vertex: worklist item 4 + offsets 8 (RCSR 16) + own height 4 + append 4
  // per slot: slot record (BCSR 8, BCSR2 12, RCSR 8 forward / 12 reverse -> 10) + height 4
  // per push: excess 4 + residual re-read 4 + four int32 read-modify-writes 32
  // per relabel: height store 4
  const int64_t per_v = layout == MF_LAYOUT_RCSR ? 28 : 20;
  const int64_t per_slot = layout == MF_LAYOUT_BCSR ? 12 : (layout == MF_LAYOUT_BCSR2 ? 16 : 14);
  return active * per_v + slots * per_slot + pushes * 40 + relabels * 4;
}

int64_t algorithmic_bytes_bfs(mf_layout layout, int32_t n, int64_t rounds, int64_t verts,
                              int64_t slots, int64_t bu_verts, int64_t bu_slots) {
  // per round: height initialisation 4n; per frontier vertex: item 4 + offsets 8;
  // per slot: slot record + reverse residual 4 + neighbour height 4.
  // bottom-up: per unreached vertex examined: height 4 + offsets 8; per slot: slot record +
  // neighbour height 4
  const int64_t per_slot = layout == MF_LAYOUT_BCSR ? 16 : (layout == MF_LAYOUT_BCSR2 ? 20 : 14);
  const int64_t bu_slot = layout == MF_LAYOUT_BCSR ? 12 : (layout == MF_LAYOUT_BCSR2 ? 16 : 14);
  return rounds * 4 * (int64_t)n + verts * 12 + slots * per_slot + bu_verts * 12 +
         bu_slots * bu_slot;
}

}  // namespace mf

struct mf_context {
  int device = 0;
  hipStream_t stream = nullptr;
  int num_cu = 0;
  char name[256] = {0};
  char arch[64] = {0};
  std::vector<hipEvent_t> prof_ev;
  hipEvent_t ev_solve[4] = {nullptr, nullptr, nullptr, nullptr};  // brackets of the solve's phases
  void *staging = nullptr;  // pinned host buffer for packed batches (grown on demand)
  size_t staging_bytes = 0;
  bool cooperative = false;  // the device runs cooperative launches
};

namespace mf {

// Device-resident residual graph of k packed instances.
struct DevGraph {
  mf_layout layout = MF_LAYOUT_BCSR;
  int32_t n = 0;  // vertices of all instances
  int64_t m = 0, S = 0;
  int32_t k = 0;
  // layout arrays
  int *off = nullptr;
  int2 *dc = nullptr;   // BCSR {dst, residual}
  int *rev = nullptr;   // BCSR reverse slot
  int *dst = nullptr;   // BCSR2 dst
  int2 *cr = nullptr;   // BCSR2 {residual, reverse slot}
  int2 *fdc = nullptr;  // RCSR {dst, forward residual}
  int *bcf = nullptr;   // RCSR backward residual
  int *roff = nullptr;  // RCSR reverse offsets
  int2 *rdi = nullptr;  // RCSR {tail, arc}
  int *cap0 = nullptr;  // initial residual per slot (per arc for RCSR)
  int *fslot = nullptr; // BCSR, BCSR2: forward slot of every arc
  unsigned long long *work = nullptr;  // record_workload: busy cycles per wave of k_cycle
  unsigned char *work_cls = nullptr;   //   and the degree class of each wave's block
  int64_t work_waves = 0;
  int *cycle_base = nullptr;  // graph replays: first cycle number of the replayed chunk
  int *cls_tab = nullptr;     // class tables of the frontier lists and the cycle worklists
  // captured cycle chunks, kept across solves: one set per configuration of the launches
  struct CycleGraphs {
    std::array<int, 9> key;
    std::vector<hipGraphExec_t> exec;  // index: chunk length / kGraphPeriod
  };
  std::vector<CycleGraphs> cycle_graphs;
  // terminals
  int *src = nullptr, *snk = nullptr;
  long long *values = nullptr;
  // state
  int *h = nullptr, *e = nullptr;
  int *stamp = nullptr;  // last cycle that discharged each vertex (push forwarding)
  int *lists = nullptr;  // 4 sets x n items
  int *ctl = nullptr;    // counters and flags
  unsigned long long *stats = nullptr;
  unsigned char *vcls = nullptr;  // VC degree class per vertex
  int64_t vc_pop[kNumClasses] = {0};
  int *vorder = nullptr;           // nonzero-degree vertices grouped by VC class
  VertexOrder vo{};                // vorder's class regions
  int64_t bu_pop[kNumClasses] = {0};
  int64_t nonzero = 0;             // vertices of nonzero degree
  int64_t class_slots[kNumClasses] = {0};  // residual slots of the vertices of each VC class
  int64_t max_src_deg = 0;         // largest residual degree of a source
  int64_t giant_parts = 0;         // parts of all giants (kPartSlots slots each)
  unsigned long long *giant = nullptr;  // per giant-list position: state of the parts
  // host copies for result read-back (single solves; batched uploads keep none)
  std::vector<int32_t> h_off, h_dst, h_cap, h_fslot, h_src, h_snk;
  std::vector<int32_t> degree;  // residual slots of every vertex (out + in arcs)
  bool solved = false;
  bool fresh = false;  // residuals still equal the capacities (built, not solved since)
  unsigned char *dirty = nullptr;  // BCSR/BCSR2 lines of slot pairs pushes changed (large graphs)
  int *bu_rem = nullptr;  // two lists of unreached vertices (2 x nonzero) + their class table

  LayoutBCSR bcsr() const { return LayoutBCSR{off, dc, rev, dirty}; }
  LayoutBCSR2 bcsr2() const { return LayoutBCSR2{off, dst, cr, dirty}; }
  LayoutRCSR rcsr() const { return LayoutRCSR{off, fdc, bcf, roff, rdi, (int)m}; }
};

// control block (int offsets); every counter on a 128-byte line of its own (kCtlStride)
enum {
  CTL_FCNT = 0,                                        // 3 x kNumClasses frontier counters
  CTL_CCNT = CTL_FCNT + 3 * kNumClasses * kCtlStride,  // 3 x kNumClasses worklist counters
  CTL_FLAGS = CTL_CCNT + 3 * kNumClasses * kCtlStride, // [0] source reached, [1] list overflow
  CTL_STATE = CTL_FLAGS + kCtlStride,                  // [0] level, [1] frontier size,
                                                       // [2..3] grid levels' marks,
                                                       // [4..5] slots explored
  CTL_RCNT = CTL_STATE + kCtlStride,                   // 3 x kNumClasses unreached-list counters
  CTL_READ = CTL_RCNT,                                 // the part the host reads back
  CTL_SIZE = CTL_RCNT + 3 * kNumClasses * kCtlStride
};

// counter c of triple-buffer slot q of the frontier (CTL_FCNT) or worklist (CTL_CCNT) set
constexpr int ctl_cnt(int set, int q, int c) { return set + (q * kNumClasses + c) * kCtlStride; }

static void free_dev(DevGraph *g) {
  for (auto &cg : g->cycle_graphs)
    for (hipGraphExec_t x : cg.exec)
      if (x) (void)hipGraphExecDestroy(x);
  g->cycle_graphs.clear();
  void *ptrs[] = {g->off, g->dc,  g->rev, g->dst,    g->cr,    g->fdc,   g->bcf,
                  g->roff, g->rdi, g->cap0, g->src,   g->snk,   g->values, g->h,
                  g->e,   g->lists, g->ctl, g->stats, g->vcls, g->vorder, g->stamp, g->fslot,
                  g->work, g->work_cls, g->cycle_base, g->cls_tab, g->giant, g->dirty,
                  g->bu_rem};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
}

// Sums the spread statistics copies (see block_flush).
static hipError_t read_stats(const unsigned long long *dev, unsigned long long out[kStatSlot]) {
  std::vector<unsigned long long> all((size_t)kStatSlot * kStatSpread);
  hipError_t err = hipMemcpy(all.data(), dev, sizeof(unsigned long long) * all.size(),
                             hipMemcpyDeviceToHost);
  for (int i = 0; i < kStatSlot; ++i) {
    out[i] = 0;
    for (int q = 0; q < kStatSpread; ++q) out[i] += all[(size_t)q * kStatSlot + i];
  }
  return err;
}

// Degree classes of every vertex (VC buckets) and the class-grouped order of the vertices of
// nonzero degree that bottom-up BFS levels walk.  degree counts residual slots.
static mf_status set_vertex_classes(DevGraph &G, const std::vector<int64_t> &degree,
                                    hipStream_t st) {
  const int n = G.n;
  mf_status rc;
  // (arrays already placed by the caller, as in the static-scratch upload, are used as they are)
  if ((!G.vcls && (rc = dmalloc(&G.vcls, (size_t)n))) ||
      (!G.vorder && (rc = dmalloc(&G.vorder, (size_t)std::max(n, 1)))))
    return rc;
  const ClassConfig vc = class_config(MF_ALGO_VC, 0);
  std::vector<unsigned char> cls((size_t)n);
  for (int q = 0; q < kNumClasses; ++q) G.vc_pop[q] = G.bu_pop[q] = 0;
  for (int q = 0; q < kNumClasses; ++q) G.class_slots[q] = 0;
  G.giant_parts = 0;
  // classes in parallel over contiguous vertex chunks; per chunk the counts of nonzero-degree
  // vertices by class, so the chunks fill their parts of the class-ordered list in vertex order
  constexpr int64_t kChunk = 1 << 20;
  const int64_t nchunk = std::max<int64_t>(1, (n + kChunk - 1) / kChunk);
  std::vector<std::array<int64_t, 4 * kNumClasses + 1>> part((size_t)nchunk);
  parallel_for(
      nchunk,
      [&](int64_t b, int64_t e) {
        for (int64_t ch = b; ch < e; ++ch) {
          auto &pc = part[(size_t)ch];
          pc.fill(0);
          for (int64_t v = ch * kChunk; v < std::min<int64_t>(n, (ch + 1) * kChunk); ++v) {
            const int64_t d = degree[(size_t)v];
            const int q = class_of_degree(vc, d);
            cls[(size_t)v] = (unsigned char)q;
            pc[q]++;                                      // vertices
            pc[kNumClasses + q] += d;                     // slots
            if (d > 0) pc[2 * kNumClasses + q]++;         // nonzero degree
            if (vc.lanes[q] == kSplitLanes) pc[4 * kNumClasses] += (d + kPartSlots - 1) / kPartSlots;
          }
        }
      },
      1);
  for (const auto &pc : part) {
    for (int q = 0; q < kNumClasses; ++q) {
      G.vc_pop[q] += pc[q];
      G.class_slots[q] += pc[kNumClasses + q];
      G.bu_pop[q] += pc[2 * kNumClasses + q];
    }
    G.giant_parts += pc[4 * kNumClasses];
  }
  // giant-list state, one entry per possible list position: no minimum, no parts, nothing
  // reserved (each giant's last part restores this)
  const size_t giants = (size_t)std::max<int64_t>(G.vc_pop[kNumClasses - 1], 1);
  if (!G.giant && (rc = dmalloc(&G.giant, giants * kGiantWords))) return rc;
  std::vector<unsigned long long> gs0(giants * kGiantWords, 0ull);
  for (size_t q = 0; q < giants; ++q) gs0[q * kGiantWords] = ~0ull;
  HIP_TRY(hipMemcpy(G.giant, gs0.data(), sizeof(unsigned long long) * gs0.size(),
                    hipMemcpyHostToDevice));
  G.nonzero = 0;
  int fill[kNumClasses];
  for (int q = 0; q < kNumClasses; ++q) {
    G.vo.cls_off[q] = fill[q] = (int)G.nonzero;
    G.vo.cls_cnt[q] = (int)G.bu_pop[q];
    G.nonzero += G.bu_pop[q];
  }
  std::vector<int> order((size_t)G.nonzero);
  for (auto &pc : part)  // each chunk's first position per class (exclusive prefix over chunks)
    for (int q = 0; q < kNumClasses; ++q) {
      pc[3 * kNumClasses + q] = fill[q];
      fill[q] += (int)pc[2 * kNumClasses + q];
    }
  parallel_for(
      nchunk,
      [&](int64_t b, int64_t e) {
        for (int64_t ch = b; ch < e; ++ch) {
          int64_t at[kNumClasses];
          for (int q = 0; q < kNumClasses; ++q) at[q] = part[(size_t)ch][3 * kNumClasses + q];
          for (int64_t v = ch * kChunk; v < std::min<int64_t>(n, (ch + 1) * kChunk); ++v)
            if (degree[(size_t)v] > 0) order[(size_t)at[cls[(size_t)v]]++] = (int)v;
        }
      },
      1);
  G.vo.items = G.vorder;
  HIP_TRY(hipMemcpyAsync(G.vcls, cls.data(), (size_t)n, hipMemcpyHostToDevice, st));
  if (!order.empty())
    HIP_TRY(hipMemcpyAsync(G.vorder, order.data(), sizeof(int) * order.size(),
                           hipMemcpyHostToDevice, st));
  HIP_TRY(hipStreamSynchronize(st));
  return MF_OK;
}

// Uploads a (packed) graph: c is the union CSR, src/snk the k terminal pairs.
// keep_host: keep host copies of the CSR for flow read-back (single solves); batched uploads,
// whose flows are never read, skip them.
static mf_status upload(mf_context *ctx, const mf_csr *c, const int32_t *src, const int32_t *snk,
                        int32_t k, mf_layout layout, DevGraph &G, bool keep_host = true) {
  G.layout = layout;
  G.n = c->n;
  G.m = c->m;
  G.k = k;
  std::vector<int32_t> unit_cap;
  const int32_t *cap = c->cap;
  if (!cap) {  // unit capacities (SNAP inputs)
    unit_cap.assign((size_t)c->m, 1);
    cap = unit_cap.data();
  }
  if (keep_host) {
    G.h_off.assign(c->offsets, c->offsets + c->n + 1);
    G.h_dst.assign(c->dst, c->dst + c->m);
    G.h_cap.assign(cap, cap + c->m);
  }
  G.h_src.assign(src, src + k);
  G.h_snk.assign(snk, snk + k);
  if ((int64_t)c->n + 1 >= INT32_MAX) return MF_ERR_OVERFLOW;
  for (int32_t i = 0; i < k; ++i) {
    int64_t sc = 0;
    for (int32_t a = c->offsets[src[i]]; a < c->offsets[src[i] + 1]; ++a)
      if (c->dst[a] != src[i]) sc += cap[a];
    if (sc > INT32_MAX) return MF_ERR_OVERFLOW;  // every excess must fit int32
  }
  hipStream_t st = ctx->stream;
  if (2 * c->m > INT32_MAX) return MF_ERR_OVERFLOW;
  std::vector<int64_t> degree((size_t)c->n, 0);
  mf_status rc;
  // the CSR goes to the device as it is; the residual layout is built there (build.hip)
  int *d_off = nullptr, *d_dst = nullptr, *d_cap = nullptr;
  auto free_raw = [&]() {
    if (d_off && d_off != G.off) (void)hipFree(d_off);
    if (d_dst) (void)hipFree(d_dst);
    if (d_cap) (void)hipFree(d_cap);
  };
  if ((rc = dmalloc(&d_off, (size_t)c->n + 1)) || (rc = dmalloc(&d_dst, (size_t)c->m)) ||
      (rc = dmalloc(&d_cap, (size_t)c->m))) {
    free_raw();
    return rc;
  }
  // (a graph without arcs may pass NULL arc arrays: nothing to copy then)
  if (hipMemcpyAsync(d_off, c->offsets, sizeof(int) * (c->n + 1), hipMemcpyHostToDevice, st) ||
      (c->m > 0 &&
       (hipMemcpyAsync(d_dst, c->dst, sizeof(int) * c->m, hipMemcpyHostToDevice, st) ||
        hipMemcpyAsync(d_cap, cap, sizeof(int) * c->m, hipMemcpyHostToDevice, st)))) {
    free_raw();
    return MF_ERR_DEVICE;
  }
  DeviceLayout out;
  if (layout == MF_LAYOUT_RCSR) {
    G.S = c->m;
    G.off = d_off;  // RCSR keeps the arc offsets
    if ((rc = dmalloc(&G.fdc, (size_t)c->m)) || (rc = dmalloc(&G.bcf, (size_t)c->m)) ||
        (rc = dmalloc(&G.roff, (size_t)c->n + 1)) || (rc = dmalloc(&G.rdi, (size_t)c->m)) ||
        (rc = dmalloc(&G.cap0, (size_t)c->m))) {
      free_raw();
      return rc;
    }
    out.fdc = G.fdc;
    out.roff = G.roff;
    out.rdi = G.rdi;
    out.cap0 = G.cap0;
  } else {
    G.S = 2 * c->m;
    if ((rc = dmalloc(&G.off, (size_t)c->n + 1)) || (rc = dmalloc(&G.cap0, (size_t)G.S)) ||
        (rc = dmalloc(&G.fslot, (size_t)c->m))) {
      free_raw();
      return rc;
    }
    out.soff = G.off;
    out.cap0 = G.cap0;
    out.fslot = G.fslot;
    if (layout == MF_LAYOUT_BCSR) {
      rc = dmalloc(&G.dc, (size_t)G.S);
      if (rc == MF_OK) rc = dmalloc(&G.rev, (size_t)G.S);
      out.dc = G.dc;
      out.rev = G.rev;
    } else {
      rc = dmalloc(&G.dst, (size_t)G.S);
      if (rc == MF_OK) rc = dmalloc(&G.cr, (size_t)G.S);
      out.dst2 = G.dst;
      out.cr = G.cr;
    }
    if (rc != MF_OK) {
      free_raw();
      return rc;
    }
  }
  std::vector<int> in_off((size_t)c->n + 1);
  rc = build_layout_device(st, c->n, c->m, d_off, d_dst, d_cap, layout, out, in_off.data());
  free_raw();
  if (rc != MF_OK) return rc;
  parallel_for(c->n, [&](int64_t b, int64_t e) {
    for (int64_t v = b; v < e; ++v)
      degree[(size_t)v] = (int64_t)(c->offsets[v + 1] - c->offsets[v]) +
                          (in_off[(size_t)v + 1] - in_off[(size_t)v]);
  });
  G.degree.resize((size_t)c->n);
  parallel_for(c->n, [&](int64_t b, int64_t e) {
    for (int64_t v = b; v < e; ++v) G.degree[(size_t)v] = (int32_t)degree[(size_t)v];
  });
  if ((rc = dmalloc(&G.h, (size_t)c->n)) || (rc = dmalloc(&G.e, (size_t)c->n)) ||
      (rc = dmalloc(&G.stamp, (size_t)c->n)) || (rc = dmalloc(&G.cycle_base, 1)) ||
      (rc = dmalloc(&G.cls_tab, (size_t)4 * kNumClasses)) ||
      (rc = dmalloc(&G.lists, (size_t)4 * c->n)) || (rc = dmalloc(&G.ctl, (size_t)CTL_SIZE)) ||
      (rc = dmalloc(&G.stats, (size_t)kStatSlot * kStatSpread)) || (rc = dmalloc(&G.src, (size_t)k)) ||
      (rc = dmalloc(&G.snk, (size_t)k)) || (rc = dmalloc(&G.values, (size_t)k)))
    return rc;
  if ((rc = set_vertex_classes(G, degree, st)) != MF_OK) return rc;
  G.max_src_deg = 0;
  for (int32_t i = 0; i < k; ++i) G.max_src_deg = std::max(G.max_src_deg, degree[(size_t)src[i]]);
  HIP_TRY(hipMemcpyAsync(G.src, src, sizeof(int) * k, hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(G.snk, snk, sizeof(int) * k, hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemsetAsync(G.ctl, 0, sizeof(int) * CTL_SIZE, st));
  if (keep_host && G.nonzero >= kBottomUpMinVertices) {  // bottom-up levels' unreached lists
    if ((rc = dmalloc(&G.bu_rem, (size_t)2 * G.nonzero + 2 * kNumClasses))) return rc;
    int tab[2 * kNumClasses];
    for (int c = 0; c < kNumClasses; ++c) {
      tab[c] = G.vo.cls_off[c];
      tab[kNumClasses + c] = G.vo.cls_cnt[c];
    }
    HIP_TRY(hipMemcpyAsync(G.bu_rem + 2 * G.nonzero, tab, sizeof(tab), hipMemcpyHostToDevice, st));
  }
  if (keep_host && layout != MF_LAYOUT_RCSR && G.S >= kDirtyMinSlots) {  // single graphs only
    if ((rc = dmalloc(&G.dirty, (size_t)(G.S + 7) / 8))) return rc;
    HIP_TRY(hipMemsetAsync(G.dirty, 0, (size_t)(G.S + 7) / 8, st));
  }
  HIP_TRY(hipStreamSynchronize(st));
  G.fresh = true;  // the layout build wrote every residual as its capacity
  return MF_OK;
}

template <class L>
static mf_status run_solve(mf_context *ctx, DevGraph &G, const L &lay, mf_algo algo,
                           const mf_options &o, mf_result *res) {
  hipStream_t st = ctx->stream;
  const int n = G.n;
  const ClassConfig cfg = class_config(algo, o.tile_lanes, o.class_lanes);
  int64_t pop[kNumClasses] = {0};
  if (cfg.nclass == 1) pop[0] = n;
  else
    for (int q = 0; q < kNumClasses; ++q) pop[q] = G.vc_pop[q];
  const unsigned char *vcls = cfg.nclass == 1 ? nullptr : G.vcls;
  const Dispatch disp = make_dispatch(cfg, pop, ctx->num_cu, G.giant_parts);
  const int grid = disp.blk_begin[cfg.nclass];
  // the global relabel always splits its frontier by degree class, whatever the cycle kernels do
  const ClassConfig bcfg = class_config(MF_ALGO_VC, 0);
  const Dispatch bdisp = make_dispatch(bcfg, G.vc_pop, ctx->num_cu, G.giant_parts, kBfsBlocksPerCU);
  const int bgrid = bdisp.blk_begin[bcfg.nclass];
  // bottom-up levels walk the nonzero-degree vertices in class order, with tiles half as wide as
  // the top-down levels': a vertex stops at its first residual arc into the frontier, so most of
  // a level's memory requests are per vertex, not per slot
  ClassConfig bucfg = bcfg;
  for (int c = 0; c < 4 && c < bucfg.nclass; ++c) bucfg.lanes[c] = kBottomUpLanes[c];
  const Dispatch budisp = make_dispatch(bucfg, G.bu_pop, ctx->num_cu, 0);
  const int bugrid = budisp.blk_begin[bcfg.nclass];
  const bool bu_allowed = o.bfs_direction == 0 && G.nonzero >= kBottomUpMinVertices;
  const bool bu_always = o.bfs_direction == 2;

  // list sets: F0, F1 (BFS frontier), S0, S1 (cycle worklists); class regions by population
  // class tables: 0 for the frontier lists, 1 for the cycle worklists (the lists of a pair
  // share their class regions)
  int tab[2][2 * kNumClasses];
  for (int q = 0; q < 2; ++q) {
    const int nc = q == 0 ? bcfg.nclass : cfg.nclass;
    const int64_t *pp = q == 0 ? G.vc_pop : pop;
    int64_t acc = 0;
    for (int c = 0; c < kNumClasses; ++c) {
      tab[q][c] = (int)acc;
      tab[q][kNumClasses + c] = c < nc ? (int)pp[c] : 0;
      if (c < nc) acc += pp[c];
    }
  }
  HIP_TRY(hipMemcpy(G.cls_tab, tab, sizeof(tab), hipMemcpyHostToDevice));
  ListSet sets[4];
  for (int q = 0; q < 4; ++q) {
    sets[q].items = G.lists + (size_t)q * n;
    sets[q].cls = G.cls_tab + (q < 2 ? 0 : 2 * kNumClasses);
    sets[q].cnt = nullptr;
    sets[q].cstride = kCtlStride;
    sets[q].ovf = G.ctl + CTL_FLAGS + 1;
  }
  int *fcnt[3], *ccnt[3];
  for (int q = 0; q < 3; ++q) {
    fcnt[q] = G.ctl + ctl_cnt(CTL_FCNT, q, 0);
    ccnt[q] = G.ctl + ctl_cnt(CTL_CCNT, q, 0);
  }

  BfsArgs ba;
  ba.h = G.h;
  ba.e = G.e;
  ba.H = n;
  ba.vcls = G.vcls;
  ba.nclass = bcfg.nclass;
  ba.acls = vcls;
  ba.nclass_act = cfg.nclass;
  ba.fr[0] = sets[0];
  ba.fr[1] = sets[1];
  for (int q = 0; q < 3; ++q) ba.fcnt[q] = fcnt[q];
  ba.active = sets[2];
  ba.active.cnt = ccnt[0];
  ba.flags = G.ctl + CTL_FLAGS;
  ba.state = G.ctl + CTL_STATE;
  for (int c = 0; c < kNumClasses; ++c)
    ba.cls_mean[c] = G.vc_pop[c] ? (int)((G.class_slots[c] + G.vc_pop[c] - 1) / G.vc_pop[c]) : 0;
  // a giant in the frontier always goes to the grid-wide levels: one workgroup would scan its
  // (up to millions of) slots alone
  for (int c = 0; c < bcfg.nclass; ++c)
    if (bcfg.lanes[c] == kSplitLanes) ba.cls_mean[c] = std::max(ba.cls_mean[c], kWgBfsSlots + 1);
  ba.wg_slots = kWgBfsSlots;
  ba.stats = G.stats;
  ba.rem_items[0] = G.bu_rem;
  ba.rem_items[1] = G.bu_rem ? G.bu_rem + G.nonzero : nullptr;
  ba.rem_cls = G.bu_rem ? G.bu_rem + 2 * G.nonzero : nullptr;
  for (int q = 0; q < 3; ++q) ba.rcnt[q] = G.ctl + CTL_RCNT + q * kNumClasses * kCtlStride;

  CycleArgs ca;
  ca.h = G.h;
  ca.e = G.e;
  ca.H = n;
  ca.vcls = vcls;
  ca.nclass = cfg.nclass;
  ca.stats = G.stats;
  ca.fwd_hops = o.forward_hops < 0 ? 0 : (o.forward_hops > 0 ? o.forward_hops : kDefaultForwardHops);
  ca.stamp = G.stamp;
  ca.cycle_base = nullptr;
  ca.work = nullptr;
  ca.work_cls = nullptr;
  ca.giant = G.giant;
  if (o.record_workload) {  // per-wave busy cycles, summed over this solve's cycle launches
    const int64_t waves = (int64_t)grid * (kBlock / 64);
    if (G.work_waves != waves) {
      if (G.work) (void)hipFree(G.work);
      if (G.work_cls) (void)hipFree(G.work_cls);
      G.work = nullptr;
      G.work_cls = nullptr;
      G.work_waves = 0;
      mf_status rc;
      if ((rc = dmalloc(&G.work, (size_t)waves)) || (rc = dmalloc(&G.work_cls, (size_t)waves)))
        return rc;
      G.work_waves = waves;
    }
    HIP_TRY(hipMemsetAsync(G.work, 0, sizeof(unsigned long long) * waves, st));
    HIP_TRY(hipMemsetAsync(G.work_cls, 0, (size_t)waves, st));
    ca.work = G.work;
    ca.work_cls = G.work_cls;
  }

  const int fmax = o.bfs_block_frontier > 0 ? o.bfs_block_frontier : kDefaultBfsFrontier;
  const int fmax_low = std::max(1, fmax / 2);
  bool use_coop = ctx->cooperative && o.bfs_cooperative == 1;
  const bool adaptive = o.cycles_per_relabel == 0;
  const int K = adaptive ? kMaxAdaptiveCycles : o.cycles_per_relabel;
  const double ratio = o.relabel_time_ratio > 0 ? o.relabel_time_ratio : kDefaultRelabelRatio;
  double last_relabel_ms = 0;
  double cyc_est_ms = 0;  // wall time per cycle launch in the last chunk
  // captured cycle chunks, one per length (multiples of kGraphPeriod), cached on the graph for
  // this launch configuration (a recording solve captures its own, released at its end)
  std::vector<hipGraphExec_t> *gexec = nullptr;
  struct LocalGraphs {
    std::vector<hipGraphExec_t> exec;
    ~LocalGraphs() {
      for (hipGraphExec_t x : exec)
        if (x) (void)hipGraphExecDestroy(x);
    }
  } local_graphs;
  if (o.cycle_graphs) {
    const size_t slots = kGraphChunkMax / kGraphPeriod + 1;
    if (o.record_workload) {
      local_graphs.exec.assign(slots, nullptr);
      gexec = &local_graphs.exec;
    } else {
      std::array<int, 9> key = {(int)algo, o.tile_lanes, o.class_lanes[0], o.class_lanes[1],
                                o.class_lanes[2], o.class_lanes[3], o.class_lanes[4],
                                o.forward_hops, grid};
      for (auto &cg : G.cycle_graphs)
        if (cg.key == key) gexec = &cg.exec;
      if (!gexec) {
        G.cycle_graphs.push_back(DevGraph::CycleGraphs{key, std::vector<hipGraphExec_t>(slots, nullptr)});
        gexec = &G.cycle_graphs.back().exec;
      }
    }
  }
  int64_t graph_launches = 0;
  hipEvent_t e_graph[2] = {nullptr, nullptr};
  struct EventPair {
    hipEvent_t *e;
    ~EventPair() {
      for (int q = 0; q < 2; ++q)
        if (e[q]) (void)hipEventDestroy(e[q]);
    }
  } e_graph_guard{e_graph};
  if (o.cycle_graphs && o.profile_kernels > 0) {
    HIP_TRY(hipEventCreate(&e_graph[0]));
    HIP_TRY(hipEventCreate(&e_graph[1]));
  }
  const int64_t max_rounds = o.max_rounds > 0 ? o.max_rounds : (int64_t)100000000;
  const int vgrid = std::max(1, std::min((n + 255) / 256, 16 * ctx->num_cu));
  const int kgrid = std::max(1, std::min((G.k + 255) / 256, 16 * ctx->num_cu));

  hipEvent_t e0 = ctx->ev_solve[0], e1 = ctx->ev_solve[1];    // global relabels
  hipEvent_t ec0 = ctx->ev_solve[2], ec1 = ctx->ev_solve[3];  // cycle phases
  double kernel_ms = 0, relabel_ms = 0, prof_ms = 0;
  int64_t prof_launches = 0, cycles = 0, rounds = 0, bfs_levels = 0, bfs_grid_levels = 0,
          bfs_bu_levels = 0;
  mf_status status = MF_OK;
  if (o.profile_kernels && ctx->prof_ev.size() < (size_t)4 * kCycleChunk) {
    const size_t old = ctx->prof_ev.size();
    ctx->prof_ev.resize((size_t)4 * kCycleChunk);
    for (size_t q = old; q < ctx->prof_ev.size(); ++q) HIP_TRY(hipEventCreate(&ctx->prof_ev[q]));
  }

  const double t0 = now_ms();
  const int rgrid = std::max(1, (int)std::min<int64_t>((G.S + 255) / 256, 16 * ctx->num_cu));
  // residuals back to the capacities, unless no solve has run since the layout was built
  if (!G.fresh && G.dirty) {
    const long long lines = (G.S + 7) / 8;
    const int lgrid = std::max(1, (int)std::min<long long>((lines + 255) / 256, 16 * ctx->num_cu));
    k_reset_dirty<<<lgrid, 256, 0, st>>>(G.layout == MF_LAYOUT_BCSR ? G.dc : G.cr, G.cap0, G.dirty,
                                         lines, G.S, G.layout == MF_LAYOUT_BCSR ? 1 : 0);
  } else if (!G.fresh) {
    if (G.layout == MF_LAYOUT_BCSR) k_reset_pairs<<<rgrid, 256, 0, st>>>(G.dc, G.cap0, G.S, 1);
    else if (G.layout == MF_LAYOUT_BCSR2) k_reset_pairs<<<rgrid, 256, 0, st>>>(G.cr, G.cap0, G.S, 0);
    else k_reset_pairs<<<rgrid, 256, 0, st>>>(G.fdc, G.cap0, G.m, 1);
  }
  G.fresh = false;
  k_reset_vertices<<<rgrid, 256, 0, st>>>(G.h, G.e, G.stamp, n,
                                          G.layout == MF_LAYOUT_RCSR ? G.bcf : nullptr,
                                          G.m);
  HIP_TRY(hipMemsetAsync(G.stats, 0, sizeof(unsigned long long) * kStatSlot * kStatSpread, st));
  HIP_TRY(hipMemsetAsync(G.ctl, 0, sizeof(int) * CTL_SIZE, st));
  const int pgrid = (int)std::min<int64_t>(16 * ctx->num_cu,
                                           std::max<int64_t>(G.k, (G.max_src_deg + 1023) / 1024));
  k_preflow<L><<<pgrid, kBlock, 0, st>>>(lay, G.e, G.src, G.k, G.stats);
  HIP_TRY(hipGetLastError());

  int hstate[CTL_SIZE];
  // elapsed times of a cycle chunk's timed launches (half: which set of profiling events)
  auto collect = [&](int half, int count) -> hipError_t {
    for (int q = 0; q < count; ++q) {
      float kms = 0;
      const size_t pe = (size_t)half * 2 * kCycleChunk + (size_t)2 * q;
      hipError_t err = hipEventElapsedTime(&kms, ctx->prof_ev[pe], ctx->prof_ev[pe + 1]);
      if (err != hipSuccess) return err;
      prof_ms += kms;
    }
    prof_launches += count;
    return hipSuccess;
  };
  // A fixed cadence's round ends with its last chunk of cycles: the next global relabel is
  // enqueued right behind it, without a read-back, and the round's timings are read at the
  // relabel's first synchronisation (finish_cycles).
  struct DeferredRound {
    bool on = false;
    int prev_half = 0, prev = 0, last_half = 0, last = 0, graph = 0;
  } dfr;
  auto finish_cycles = [&]() -> mf_status {
    if (!dfr.on) return MF_OK;
    dfr.on = false;
    if (dfr.prev) HIP_TRY(collect(dfr.prev_half, dfr.prev));
    if (dfr.last) HIP_TRY(collect(dfr.last_half, dfr.last));
    if (dfr.graph && o.profile_kernels > 0) {
      float gms = 0;
      HIP_TRY(hipEventElapsedTime(&gms, e_graph[0], e_graph[1]));
      prof_ms += gms;
      prof_launches += dfr.graph;
    }
    float cms = 0;
    HIP_TRY(hipEventElapsedTime(&cms, ec0, ec1));
    kernel_ms += cms;
    return MF_OK;
  };
  for (;;) {
    if (rounds >= max_rounds) {
      status = MF_ERR_ITERATION_LIMIT;
      break;
    }
    // ---- global relabel ----
    HIP_TRY(hipEventRecord(e0, st));
    bool coop_failed = false;
  relabel_again:
    k_bfs_init<<<vgrid, 256, 0, st>>>(ba, ccnt[0], ccnt[1]);
    k_bfs_seed<<<kgrid, 256, 0, st>>>(ba, G.src, G.snk, G.k);
    int level = 0, levels_total = 0;
    auto frontier_slots = [&](int lv) {
      double sl = 0;
      for (int c = 0; c < bcfg.nclass; ++c)
        sl += (double)hstate[ctl_cnt(CTL_FCNT, lv % 3, c)] * ba.cls_mean[c];
      return sl;
    };
    // the frontier fits the one-workgroup kernel's LDS groups (else it would return at once)
    auto fits_wg = [&](int lv) {
      int grp[3] = {0, 0, 0};
      for (int c = 0; c < bcfg.nclass; ++c)
        grp[wg_group(c)] += hstate[ctl_cnt(CTL_FCNT, lv % 3, c)];
      return grp[0] <= kWgCapA && grp[1] <= kWgCapB && grp[2] <= kWgCapC;
    };
    for (;;) {
      k_bfs_wg<L><<<1, kWgBlock, 0, st>>>(lay, ba, bdisp, level, fmax);
      HIP_TRY(hipGetLastError());
      HIP_TRY(hipMemcpyAsync(hstate, G.ctl, sizeof(int) * CTL_READ, hipMemcpyDeviceToHost, st));
      HIP_TRY(hipStreamSynchronize(st));
      {
        const mf_status fs = finish_cycles();  // the cycles before this relabel have finished
        if (fs != MF_OK) return fs;
      }
      level = hstate[CTL_STATE];
      int size = hstate[CTL_STATE + 1];
      if (size == 0) break;
      bool bottom_up = false, bu_prev = false;
      for (;;) {  // frontier too large for one workgroup: grid launches per level
        // frontier slots estimated from its class counts and the classes' mean degrees
        const double mf = frontier_slots(level);
        long long explored;  // estimated slots of the frontiers the grid-wide levels expanded
        std::memcpy(&explored, &hstate[CTL_STATE + 4], sizeof(explored));
        const double mu = std::max(0.0, (double)G.S - (double)explored);
        if (bu_allowed) {
          if (!bottom_up && mf * kBottomUpAlpha > mu) bottom_up = true;
          else if (bottom_up && (int64_t)size * kBottomUpBeta < G.nonzero) bottom_up = false;
        }
        bottom_up = bottom_up || bu_always;
        int next;
        bool was_bu = false;
        if (bottom_up) {
          // the first bottom-up level of a run walks every vertex; its unreached list, which the
          // next bottom-up level walks, starts empty
          if (!bu_prev && G.bu_rem)
            k_fill_int<<<1, 256, 0, st>>>(ba.rcnt[(level + 1) % 3], 0, kNumClasses * kCtlStride);
          const int stay = bu_always ? 1 : (int)((G.nonzero + kBottomUpBeta - 1) / kBottomUpBeta);
          for (int q = 0; q < kBottomUpBatch; ++q)
            k_bfs_bottom_up<L><<<bugrid, kBlock, 0, st>>>(lay, ba, budisp, G.vo, level + q,
                                                          q == 0 && !bu_prev, level, stay);
          bu_prev = true;
          was_bu = true;
          next = -1;
        } else if (use_coop && mf <= (double)kCoopSlots) {
          // levels inside one cooperative launch until the frontier leaves the mid-size range
          CoopStop cs;
          cs.small = fmax_low;
          cs.small_slots = ba.wg_slots / 2;
          cs.max_slots = std::min<long long>(
              kCoopSlots, (bu_allowed && !bu_always) ? (long long)(mu / kBottomUpAlpha / 2)
                                                     : std::numeric_limits<long long>::max());
          unsigned *bar = reinterpret_cast<unsigned *>(G.ctl + CTL_STATE + 6);
          int *abort_flag = G.ctl + CTL_STATE + 7;
          k_fill_int<<<1, 64, 0, st>>>(G.ctl + CTL_STATE + 6, 0, 2);
          int lv = level;
          void *args[] = {(void *)&lay, (void *)&ba, (void *)&bdisp, (void *)&lv,
                          (void *)&cs,  (void *)&bar, (void *)&abort_flag};
          HIP_TRY(hipLaunchCooperativeKernel((const void *)k_bfs_coop<L>, dim3(kCoopBlocks),
                                             dim3(kBlock), args, 0, st));
          next = -2;
          bu_prev = false;
        } else {
          // top-down levels go kGridLevelBatch at a time between two read-backs; the levels after
          // the first stop by themselves (LevelStop) when the frontier empties, nears half the
          // bottom-up switch point or fits the one-workgroup kernel again, so the host decides
          // those turns as before, at most a batch later, and a read-back (a host round trip)
          // covers several levels of a long-diameter graph
          LevelStop ls;
          ls.first = level;
          ls.small = fmax_low;
          ls.stop_slots = (bu_allowed && !bu_always)
                              ? (long long)(mu / kBottomUpAlpha / 2)
                              : std::numeric_limits<long long>::max();
          for (int q = 0; q < kGridLevelBatch; ++q)
            k_bfs_level<L><<<bgrid, kBlock, 0, st>>>(lay, ba, bdisp, level + q, ls);
          next = -1;
          bu_prev = false;
        }
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(hstate, G.ctl, sizeof(int) * CTL_READ, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        if (next == -2) {
          if (hstate[CTL_STATE + 7] != 0) {  // a grid barrier gave up: redo without them
            coop_failed = true;
            break;
          }
          next = hstate[CTL_STATE];
        }
        if (next < 0) next = std::max(hstate[CTL_STATE + 2], hstate[CTL_STATE + 3]);
        if (was_bu) bfs_bu_levels += next - level;
        bfs_grid_levels += next - level;
        level = next;
        size = 0;
        for (int c = 0; c < bcfg.nclass; ++c) size += hstate[ctl_cnt(CTL_FCNT, level % 3, c)];
        if (size <= fmax_low && frontier_slots(level) <= ba.wg_slots / 2 && fits_wg(level)) break;
      }
      if (coop_failed || size == 0) break;
    }
    if (coop_failed && use_coop) {
      use_coop = false;
      goto relabel_again;
    }
    levels_total = level;
    bfs_levels += level;
    HIP_TRY(hipEventRecord(e1, st));
    HIP_TRY(hipEventSynchronize(e1));
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    relabel_ms += ms;
    last_relabel_ms = ms;
    ++rounds;
    if (hstate[CTL_FLAGS] != 0 || hstate[CTL_FLAGS + 1] != 0) {
      status = MF_ERR_INTERNAL;  // residual path from a source, or a worklist overflow
      break;
    }
    int64_t active = 0;
    for (int c = 0; c < cfg.nclass; ++c) active += hstate[ctl_cnt(CTL_CCNT, 0, c)];
    if (o.verbose)
      printf("round %lld: relabel levels %d, active vertices %lld\n", (long long)rounds,
             levels_total, (long long)active);
    if (active == 0) break;
    // ---- push-relabel cycles, issued in chunks; the round ends early once the worklist is
    //      empty (checked between chunks with one small read-back) ----
    HIP_TRY(hipEventRecord(ec0, st));
    const double tc0 = now_ms();
    int done = 0, nchunk = 0, pending = 0;  // pending: timed launches of the previous chunk
    // (elapsed times of a chunk's timed launches are read while the next chunk runs)
    while (done < K) {
      // chunk length: what the rest of the time budget allows at the last measured cycle time
      // (one read-back per chunk; short chunks when cycles are long, up to kCycleChunk)
      int chunk = kCycleChunk;
      if (adaptive) {
        const double left_ms = ratio * last_relabel_ms - (now_ms() - tc0);
        if (done > 0 && left_ms <= 0) break;
        chunk = cyc_est_ms > 0
                    ? (int)std::min<double>(kCycleChunk, std::max(1.0, left_ms / cyc_est_ms))
                    : kFirstChunk;
      }
      chunk = std::min(chunk, K - done);
      const int half = nchunk & 1;
      const double tk0 = now_ms();
      int timed = 0, graph_chunk = 0;
      // graph mode: the chunk rounded up to whole periods of the list/counter rotation, replayed
      // from a captured graph (captured at its first use in this solve)
      const int gchunk = std::min(kGraphChunkMax, (chunk + kGraphPeriod - 1) / kGraphPeriod * kGraphPeriod);
      if (o.cycle_graphs && done % kGraphPeriod == 0 && gchunk <= K - done) {
        hipGraphExec_t &x = (*gexec)[(size_t)(gchunk / kGraphPeriod)];
        if (!x) {
          hipGraph_t graph = nullptr;
          HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
          for (int q = 0; q < gchunk; ++q) {
            CycleArgs cg = ca;
            cg.in = sets[2 + (q & 1)];
            cg.in.cnt = ccnt[q % 3];
            cg.out = sets[2 + ((q + 1) & 1)];
            cg.out.cnt = ccnt[(q + 1) % 3];
            cg.clear_cnt = ccnt[(q + 2) % 3];
            cg.cycle = q;
            cg.cycle_base = G.cycle_base;
            k_cycle<L><<<grid, kBlock, 0, st>>>(lay, cg, disp);
          }
          const hipError_t cap_err = hipGetLastError();
          HIP_TRY(hipStreamEndCapture(st, &graph));
          HIP_TRY(cap_err);
          const hipError_t inst_err = hipGraphInstantiate(&x, graph, nullptr, nullptr, 0);
          (void)hipGraphDestroy(graph);
          HIP_TRY(inst_err);
        }
        k_set_int<<<1, 1, 0, st>>>(G.cycle_base, (int)(cycles + done));
        if (o.profile_kernels > 0) HIP_TRY(hipEventRecord(e_graph[0], st));  // whole chunk
        HIP_TRY(hipGraphLaunch(x, st));
        if (o.profile_kernels > 0) HIP_TRY(hipEventRecord(e_graph[1], st));
        chunk = gchunk;
        graph_chunk = gchunk;
      }
      for (int q = done; q < done + chunk && !graph_chunk; ++q) {
        ca.in = sets[2 + (q & 1)];
        ca.in.cnt = ccnt[q % 3];
        ca.out = sets[2 + ((q + 1) & 1)];
        ca.out.cnt = ccnt[(q + 1) % 3];
        ca.clear_cnt = ccnt[(q + 2) % 3];
        ca.cycle = (int)(cycles + q);
        if (o.profile_kernels > 0 && q % o.profile_kernels == 0) {
          // every profile_kernels-th launch carries start/end events in its dispatch packet
          const size_t pe = (size_t)half * 2 * kCycleChunk + (size_t)2 * timed++;
          void *args[] = {(void *)&lay, (void *)&ca, (void *)&disp};
          HIP_TRY(hipExtLaunchKernel((const void *)k_cycle<L>, dim3(grid), dim3(kBlock), args, 0,
                                     st, ctx->prof_ev[pe], ctx->prof_ev[pe + 1], 0));
        } else {
          k_cycle<L><<<grid, kBlock, 0, st>>>(lay, ca, disp);
        }
      }
      HIP_TRY(hipGetLastError());
      done += chunk;
      if (!adaptive && done >= K) {  // the round's last chunk: the next relabel follows at once
        dfr.on = true;
        dfr.prev_half = half ^ 1;
        dfr.prev = pending;
        dfr.last_half = half;
        dfr.last = timed;
        dfr.graph = graph_chunk;
        pending = 0;
        ++nchunk;
        graph_launches += graph_chunk;
        break;
      }
      HIP_TRY(hipMemcpyAsync(hstate, G.ctl, sizeof(int) * CTL_READ, hipMemcpyDeviceToHost, st));
      if (pending) HIP_TRY(collect(half ^ 1, pending));
      pending = timed;
      ++nchunk;
      HIP_TRY(hipStreamSynchronize(st));
      cyc_est_ms = (now_ms() - tk0) / chunk;
      if (graph_chunk && o.profile_kernels > 0) {  // replay time over its launches
        float gms = 0;
        HIP_TRY(hipEventElapsedTime(&gms, e_graph[0], e_graph[1]));
        prof_ms += gms;
        prof_launches += graph_chunk;
      }
      graph_launches += graph_chunk;
      int64_t left = 0;
      for (int c = 0; c < cfg.nclass; ++c) left += hstate[ctl_cnt(CTL_CCNT, done % 3, c)];
      if (left == 0) break;
    }
    HIP_TRY(hipEventRecord(ec1, st));
    if (!dfr.on) {
      if (pending) HIP_TRY(collect((nchunk - 1) & 1, pending));
      HIP_TRY(hipEventSynchronize(ec1));
      float cms = 0;
      HIP_TRY(hipEventElapsedTime(&cms, ec0, ec1));
      kernel_ms += cms;
    }
    cycles += done;
  }
  const int ggrid = std::max(1, std::min((G.k + 255) / 256, 4 * ctx->num_cu));
  k_gather_values<<<ggrid, 256, 0, st>>>(G.e, G.snk, G.k, G.values);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(st));
  {
    const mf_status fs = finish_cycles();  // a round the loop left without a further relabel
    if (fs != MF_OK) return fs;
  }
  const double t1 = now_ms();
  unsigned long long hs[kStatSlot];
  long long v0 = 0;
  HIP_TRY(hipMemcpy(&v0, G.values, sizeof(long long), hipMemcpyDeviceToHost));
  HIP_TRY(read_stats(G.stats, hs));
  if (res) {
    memset(res, 0, sizeof(*res));
    res->value = v0;
    res->source_outflow = (int64_t)hs[ST_SOURCE_OUT];
    res->solve_ms = t1 - t0;
    res->kernel_ms = kernel_ms;
    res->relabel_ms = relabel_ms;
    res->rounds = rounds;
    res->cycles = cycles;
    res->active_total = (int64_t)hs[ST_ACTIVE];
    res->slots_scanned = (int64_t)hs[ST_SLOTS];
    res->pushes = (int64_t)hs[ST_PUSHES];
    res->relabels = (int64_t)hs[ST_RELABELS];
    res->bfs_slots = (int64_t)hs[ST_BFS_SLOTS];
    res->bytes_algorithmic = algorithmic_bytes_cycle(G.layout, res->active_total,
                                                     res->slots_scanned, res->pushes, res->relabels);
    res->bytes_bfs = algorithmic_bytes_bfs(G.layout, n, rounds, (int64_t)hs[ST_BFS_VERTS],
                                           res->bfs_slots, (int64_t)hs[ST_BU_VERTS],
                                           (int64_t)hs[ST_BU_SLOTS]);
    res->bfs_bottom_up_levels = (int32_t)bfs_bu_levels;
    res->graph_launches = graph_launches;
    res->profiled_kernel_ms = prof_ms;
    res->profiled_launches = prof_launches;
    res->bfs_levels = bfs_levels;
    res->bfs_grid_levels = bfs_grid_levels;
    res->status = status;
  }
  G.solved = (status == MF_OK);
  return status;
}

mf_status solve_graph(mf_context *ctx, DevGraph &G, mf_algo algo, const mf_options &o,
                      mf_result *res) {
  if (G.layout == MF_LAYOUT_BCSR) return run_solve(ctx, G, G.bcsr(), algo, o, res);
  if (G.layout == MF_LAYOUT_BCSR2) return run_solve(ctx, G, G.bcsr2(), algo, o, res);
  return run_solve(ctx, G, G.rcsr(), algo, o, res);
}

static mf_status check_options(const mf_options &o) {
  auto lanes_ok = [](int v) {
    return v == 1 || v == 2 || v == 4 || v == 8 || v == 16 || v == 32 || v == 64 || v == kBlock;
  };
  if (o.tile_lanes != 0 && !lanes_ok(o.tile_lanes)) return MF_ERR_INVALID_ARGUMENT;
  for (int k = 0; k < kTunedClasses; ++k)
    if (o.class_lanes[k] != 0 && !lanes_ok(o.class_lanes[k])) return MF_ERR_INVALID_ARGUMENT;
  if (o.cycles_per_relabel < 0 || o.max_rounds < 0 || o.bfs_block_frontier < 0)
    return MF_ERR_INVALID_ARGUMENT;
  if (o.bfs_direction < 0 || o.bfs_direction > 2 || o.profile_kernels < 0 ||
      o.forward_hops < -1 || o.bfs_cooperative < 0 || o.bfs_cooperative > 1)
    return MF_ERR_INVALID_ARGUMENT;
  return MF_OK;
}

// Packed-instance solve used by the batch entry point.
// Whole-instance-per-workgroup solve of an uploaded packed graph (BCSR).
static mf_status run_batch_wg(mf_context *ctx, DevGraph &G, const int32_t *vbase, mf_algo algo,
                              const mf_options &o, mf_result *res) {
  hipStream_t st = ctx->stream;
  const int32_t k = G.k;
  const ClassConfig vc = class_config(MF_ALGO_VC, 0);
  const ClassConfig cc = class_config(algo, o.tile_lanes, o.class_lanes);
  // class populations of every instance, from the slot degrees upload() recorded (the same
  // classes it stored on the device)
  std::vector<int32_t> nver((size_t)k), pop((size_t)k * kNumClasses, 0),
      bpop((size_t)k * kNumClasses, 0), order((size_t)k);
  parallel_for(
      k,
      [&](int64_t b, int64_t e) {
        for (int64_t i = b; i < e; ++i) {
          nver[(size_t)i] = vbase[i + 1] - vbase[i];
          for (int32_t v = vbase[i]; v < vbase[i + 1]; ++v) {
            const int64_t deg = G.degree[(size_t)v];
            bpop[(size_t)i * kNumClasses + class_of_degree(vc, deg)]++;
            pop[(size_t)i * kNumClasses + (cc.nclass == 1 ? 0 : class_of_degree(cc, deg))]++;
          }
        }
      },
      1);
  for (int32_t i = 0; i < k; ++i) order[(size_t)i] = i;
  std::stable_sort(order.begin(), order.end(),
                   [&](int32_t x, int32_t y) { return nver[(size_t)x] > nver[(size_t)y]; });
  int *d_vbase = nullptr, *d_nver = nullptr, *d_pop = nullptr, *d_bpop = nullptr,
      *d_order = nullptr, *d_next = nullptr, *d_status = nullptr;
  auto cleanup = [&]() {
    for (void *p : {(void *)d_vbase, (void *)d_nver, (void *)d_pop, (void *)d_bpop,
                    (void *)d_order, (void *)d_next, (void *)d_status})
      if (p) (void)hipFree(p);
  };
  mf_status rc;
  if ((rc = dmalloc(&d_vbase, (size_t)k)) || (rc = dmalloc(&d_nver, (size_t)k)) ||
      (rc = dmalloc(&d_pop, (size_t)k * kNumClasses)) ||
      (rc = dmalloc(&d_bpop, (size_t)k * kNumClasses)) || (rc = dmalloc(&d_order, (size_t)k)) ||
      (rc = dmalloc(&d_next, (size_t)1)) || (rc = dmalloc(&d_status, (size_t)k))) {
    cleanup();
    return rc;
  }
  bool ok =
      hipMemcpyAsync(d_vbase, vbase, sizeof(int) * k, hipMemcpyHostToDevice, st) == hipSuccess &&
      hipMemcpyAsync(d_nver, nver.data(), sizeof(int) * k, hipMemcpyHostToDevice, st) == hipSuccess &&
      hipMemcpyAsync(d_pop, pop.data(), sizeof(int) * pop.size(), hipMemcpyHostToDevice, st) == hipSuccess &&
      hipMemcpyAsync(d_bpop, bpop.data(), sizeof(int) * bpop.size(), hipMemcpyHostToDevice, st) == hipSuccess &&
      hipMemcpyAsync(d_order, order.data(), sizeof(int) * k, hipMemcpyHostToDevice, st) == hipSuccess &&
      hipStreamSynchronize(st) == hipSuccess;
  if (!ok) {
    cleanup();
    return MF_ERR_DEVICE;
  }
  BatchArgs b;
  memset(&b, 0, sizeof(b));
  b.h = G.h;
  b.e = G.e;
  b.lists = G.lists;
  b.N = G.n;
  b.vbase = d_vbase;
  b.nver = d_nver;
  b.src = G.src;
  b.snk = G.snk;
  b.pop = d_pop;
  b.bpop = d_bpop;
  b.vcls = G.vcls;
  b.order = d_order;
  b.k = k;
  b.next = d_next;
  b.values = G.values;
  b.status = d_status;
  b.ovf = G.ctl + CTL_FLAGS + 1;
  b.stats = G.stats;
  b.nclass = cc.nclass;
  for (int c = 0; c < kNumClasses; ++c) {
    b.lanes[c] = cc.lanes[c];
    b.blanes[c] = vc.lanes[c];
  }
  b.use_vcls = cc.nclass > 1;
  b.ratio = o.relabel_time_ratio > 0 ? o.relabel_time_ratio : kDefaultRelabelRatio;
  b.max_rounds = o.max_rounds > 0 ? (int)std::min<int64_t>(o.max_rounds, INT32_MAX) : 1000000;
  b.max_cycles = o.cycles_per_relabel > 0 ? o.cycles_per_relabel : kMaxAdaptiveCycles;
  b.fwd_hops = o.forward_hops < 0 ? 0 : (o.forward_hops > 0 ? o.forward_hops : kDefaultForwardHops);
  b.stamp = G.stamp;
  if (o.cycles_per_relabel > 0) b.ratio = 1e30f;  // fixed cadence: only the cycle cap ends a round
  // persistent workgroups: as many as can be resident at once (the kernel's registers decide),
  // so none waits for another to finish before it starts pulling instances
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_batch_wg<LayoutBCSR>, kBlock, 0) !=
          hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  const int nwg = std::max(1, std::min(k, std::min(per_cu, kBatchGroupsPerCU) * ctx->num_cu));
  const double t0 = now_ms();
  const int rgrid = std::max(1, (int)std::min<int64_t>((G.S + 255) / 256, 16 * ctx->num_cu));
  if (!G.fresh) k_reset_pairs<<<rgrid, 256, 0, st>>>(G.dc, G.cap0, G.S, 1);
  G.fresh = false;
  k_reset_vertices<<<rgrid, 256, 0, st>>>(G.h, G.e, G.stamp, G.n, nullptr, 0);
  if (hipMemsetAsync(G.stats, 0, sizeof(unsigned long long) * kStatSlot * kStatSpread, st) != hipSuccess ||
      hipMemsetAsync(G.ctl, 0, sizeof(int) * CTL_SIZE, st) != hipSuccess ||
      hipMemsetAsync(d_next, 0, sizeof(int), st) != hipSuccess) {
    cleanup();
    return MF_ERR_DEVICE;
  }
  k_batch_wg<LayoutBCSR><<<nwg, kBlock, 0, st>>>(G.bcsr(), b);
  std::vector<int> status((size_t)k);
  unsigned long long hs[kStatSlot];
  int hctl[CTL_SIZE];
  ok = hipGetLastError() == hipSuccess && hipStreamSynchronize(st) == hipSuccess;
  const double t1 = now_ms();
  ok = ok &&
       hipMemcpy(status.data(), d_status, sizeof(int) * k, hipMemcpyDeviceToHost) == hipSuccess &&
       read_stats(G.stats, hs) == hipSuccess &&
       hipMemcpy(hctl, G.ctl, sizeof(hctl), hipMemcpyDeviceToHost) == hipSuccess;
  cleanup();
  if (!ok) return MF_ERR_DEVICE;
  mf_status result = MF_OK;
  for (int32_t i = 0; i < k; ++i) {
    if (status[(size_t)i] == 1 && result == MF_OK) result = MF_ERR_ITERATION_LIMIT;
    if (status[(size_t)i] == 2) result = MF_ERR_INTERNAL;
  }
  if (hctl[CTL_FLAGS + 1]) result = MF_ERR_INTERNAL;  // worklist overflow
  if (res) {
    memset(res, 0, sizeof(*res));
    res->solve_ms = t1 - t0;
    res->kernel_ms = t1 - t0;
    res->active_total = (int64_t)hs[ST_ACTIVE];
    res->slots_scanned = (int64_t)hs[ST_SLOTS];
    res->pushes = (int64_t)hs[ST_PUSHES];
    res->relabels = (int64_t)hs[ST_RELABELS];
    res->bfs_slots = (int64_t)hs[ST_BFS_SLOTS];
    res->bytes_algorithmic = algorithmic_bytes_cycle(MF_LAYOUT_BCSR, res->active_total,
                                                     res->slots_scanned, res->pushes, res->relabels);
    res->status = result;
  }
  G.solved = result == MF_OK;
  return result;
}

mf_status batch_solve_packed(mf_context *ctx, const mf_csr *packed, const int32_t *src,
                             const int32_t *snk, const int32_t *vbase, int32_t k, mf_algo algo,
                             mf_layout layout, const mf_options &o, int64_t *values,
                             mf_result *res) {
  mf_status st = check_options(o);
  if (st != MF_OK) return st;
  HIP_TRY(hipSetDevice(ctx->device));
  DevGraph G;
  const bool per_group = o.batch_mode == 0;
  const double t0 = now_ms();
  st = upload(ctx, packed, src, snk, k, per_group ? MF_LAYOUT_BCSR : layout, G, false);
  const double t1 = now_ms();
  if (st == MF_OK)
    st = per_group ? run_batch_wg(ctx, G, vbase, algo, o, res) : solve_graph(ctx, G, algo, o, res);
  if (res) res->upload_ms = t1 - t0;
  if (st == MF_OK && values) {
    std::vector<long long> v((size_t)k);
    if (hipMemcpy(v.data(), G.values, sizeof(long long) * k, hipMemcpyDeviceToHost) != hipSuccess)
      st = MF_ERR_DEVICE;
    for (int32_t i = 0; i < k; ++i) values[i] = v[(size_t)i];
  }
  free_dev(&G);
  return st;
}

struct BatchResident {
  DevGraph G;
  std::vector<int32_t> vbase;
  int32_t k = 0;
  int32_t batch_mode = 0;
};

mf_status batch_upload(mf_context *ctx, const mf_csr *packed, const int32_t *src,
                       const int32_t *snk, const int32_t *vbase, int32_t k,
                       const mf_options &o, BatchResident **out) {
  if (!out) return MF_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  HIP_TRY(hipSetDevice(ctx->device));
  BatchResident *r = new BatchResident;
  r->vbase.assign(vbase, vbase + (size_t)k + 1);
  r->k = k;
  r->batch_mode = o.batch_mode;
  const mf_status st = upload(ctx, packed, src, snk, k, MF_LAYOUT_BCSR, r->G, false);
  if (st != MF_OK) {
    batch_release(r);
    return st;
  }
  *out = r;
  return MF_OK;
}

mf_status batch_solve_resident(mf_context *ctx, BatchResident *r, const mf_options &o,
                               int64_t *values, mf_result *res) {
  if (!r) return MF_ERR_INVALID_ARGUMENT;
  mf_status st = check_options(o);
  if (st != MF_OK) return st;
  if (o.batch_mode != r->batch_mode) return MF_ERR_INVALID_ARGUMENT;
  HIP_TRY(hipSetDevice(ctx->device));
  const mf_algo algo = o.batch_algo == 1 ? MF_ALGO_VC : MF_ALGO_TC;
  st = r->batch_mode == 0 ? run_batch_wg(ctx, r->G, r->vbase.data(), algo, o, res)
                          : solve_graph(ctx, r->G, algo, o, res);
  if (res) res->upload_ms = 0;
  if (st == MF_OK && values) {
    std::vector<long long> v((size_t)r->k);
    if (hipMemcpy(v.data(), r->G.values, sizeof(long long) * r->k, hipMemcpyDeviceToHost) !=
        hipSuccess)
      return MF_ERR_DEVICE;
    for (int32_t i = 0; i < r->k; ++i) values[i] = v[(size_t)i];
  }
  return st;
}

void batch_release(BatchResident *r) {
  if (!r) return;
  free_dev(&r->G);
  delete r;
}

// ---- serving arena (host side of k_serve_wg) ---------------------------------------------------

constexpr int kServeGroupsPerCU = 1;               // resident serving workgroups per CU
constexpr long long kServeIdleTicks = 200000;      // 2 ms at the 100 MHz wall clock

struct ServeArena {
  mf_context *ctx = nullptr;
  int64_t nv = 0, ns = 0;  // vertex and slot capacity
  int32_t qcap = 0;
  // device arena
  int *off = nullptr, *rev = nullptr, *h = nullptr, *e = nullptr, *stamp = nullptr;
  int *lists = nullptr;
  int2 *dc = nullptr;
  unsigned char *vcls = nullptr;
  int *words = nullptr;  // [0] claimed, [1] published count seen, [2] overflow, [4..5] poll time,
                         // [6..7] last claim or finish time, [8] busy workgroups
  unsigned long long *stats = nullptr;
  // host memory the GPU reads and writes uncached
  int *entries = nullptr;
  ServeReport *reports = nullptr;
  int *pub = nullptr;  // [0] published count, [1] closed
  hipStream_t copy_stream = nullptr, kernel_stream = nullptr;
  bool launched = false;
  int64_t launches = 0;
};

void serve_arena_destroy(ServeArena *a) {
  if (!a) return;
  (void)hipSetDevice(a->ctx->device);
  if (a->pub) __atomic_store_n(&a->pub[1], 1, __ATOMIC_RELEASE);  // idle workgroups leave
  if (a->kernel_stream) (void)hipStreamSynchronize(a->kernel_stream);
  if (a->copy_stream) (void)hipStreamSynchronize(a->copy_stream);
  for (void *p : {(void *)a->off, (void *)a->rev, (void *)a->h, (void *)a->e, (void *)a->stamp,
                  (void *)a->lists, (void *)a->dc, (void *)a->vcls, (void *)a->words,
                  (void *)a->stats})
    if (p) (void)hipFree(p);
  for (void *p : {(void *)a->entries, (void *)a->reports, (void *)a->pub})
    if (p) (void)hipHostFree(p);
  if (a->copy_stream) (void)hipStreamDestroy(a->copy_stream);
  if (a->kernel_stream) (void)hipStreamDestroy(a->kernel_stream);
  delete a;
}

int64_t serve_arena_bytes(int64_t nv, int64_t ns) {
  // off 4, h 4, e 4, stamp 4, class 1, four lists 16 per vertex; {dst, residual} 8 + reverse 4
  // per slot
  return 33 * nv + 12 * ns;
}

mf_status serve_arena_create(mf_context *ctx, int64_t nv, int64_t ns, int32_t qcap,
                             ServeArena **out) {
  if (!ctx || !out || nv <= 0 || ns <= 0 || qcap <= 0 || nv >= INT32_MAX / 4 || ns >= INT32_MAX)
    return MF_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  HIP_TRY(hipSetDevice(ctx->device));
  ServeArena *a = new ServeArena();
  a->ctx = ctx;
  a->nv = nv;
  a->ns = ns;
  a->qcap = qcap;
  mf_status rc = MF_OK;
  auto fail = [&](mf_status r) {
    serve_arena_destroy(a);
    return r;
  };
  if ((rc = dmalloc(&a->off, (size_t)nv + 1)) || (rc = dmalloc(&a->rev, (size_t)ns)) ||
      (rc = dmalloc(&a->h, (size_t)nv)) || (rc = dmalloc(&a->e, (size_t)nv)) ||
      (rc = dmalloc(&a->stamp, (size_t)nv)) || (rc = dmalloc(&a->lists, (size_t)4 * nv)) ||
      (rc = dmalloc(&a->dc, (size_t)ns)) || (rc = dmalloc(&a->vcls, (size_t)nv)) ||
      (rc = dmalloc(&a->words, (size_t)64)) ||
      (rc = dmalloc(&a->stats, (size_t)kStatSlot * kStatSpread)))
    return fail(rc);
  if (hipHostMalloc((void **)&a->entries, sizeof(int) * kServeEntryInts * (size_t)qcap,
                    hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess ||
      hipHostMalloc((void **)&a->reports, sizeof(ServeReport) * (size_t)qcap,
                    hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess ||
      hipHostMalloc((void **)&a->pub, sizeof(int) * 16, hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess)
    return fail(MF_ERR_OUT_OF_MEMORY);
  memset(a->entries, 0, sizeof(int) * kServeEntryInts * (size_t)qcap);
  memset((void *)a->reports, 0, sizeof(ServeReport) * (size_t)qcap);
  memset(a->pub, 0, sizeof(int) * 16);
  if (hipStreamCreateWithFlags(&a->copy_stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&a->kernel_stream, hipStreamNonBlocking) != hipSuccess ||
      hipMemsetAsync(a->words, 0, sizeof(int) * 64, a->kernel_stream) != hipSuccess ||
      hipMemsetAsync(a->stats, 0, sizeof(unsigned long long) * kStatSlot * kStatSpread,
                     a->kernel_stream) != hipSuccess ||
      hipStreamSynchronize(a->kernel_stream) != hipSuccess)
    return fail(MF_ERR_DEVICE);
  *out = a;
  return MF_OK;
}

mf_status serve_prepare(const mf_csr *g, ServeInstance &I) {
  mf_status st = validate_csr(g);
  if (st != MF_OK) return st;
  SlotLayout L;
  if ((st = build_slot_layout(g, L)) != MF_OK) return st;
  const ClassConfig vc = class_config(MF_ALGO_VC, 0);
  I.n = g->n;
  I.S = L.slots;
  I.s = g->s;
  I.t = g->t;
  I.off = std::move(L.off);
  I.rev = std::move(L.rev);
  I.dc.resize((size_t)(2 * I.S));
  for (int64_t j = 0; j < I.S; ++j) {
    I.dc[(size_t)(2 * j)] = L.dst[(size_t)j];
    I.dc[(size_t)(2 * j + 1)] = L.cap0[(size_t)j];
  }
  I.vcls.resize((size_t)I.n);
  for (int q = 0; q < kNumClasses; ++q) I.pop[q] = I.bpop[q] = 0;
  for (int32_t v = 0; v < I.n; ++v) {
    const int c = class_of_degree(vc, (int64_t)(I.off[(size_t)v + 1] - I.off[(size_t)v]));
    I.vcls[(size_t)v] = (unsigned char)c;
    I.bpop[c]++;
  }
  I.pop[0] = I.n;  // thread-centric cycles: one class
  return MF_OK;
}

mf_status serve_copy(ServeArena *a, ServeInstance &I, int32_t vb, int64_t sb, void **done) {
  if (vb < 0 || (int64_t)vb + I.n + 1 > a->nv || sb < 0 || sb + I.S > a->ns)
    return MF_ERR_INVALID_ARGUMENT;
  HIP_TRY(hipSetDevice(a->ctx->device));
  // arena ids: vertices from vb, slots from sb
  I.a_off.resize((size_t)I.n + 1);
  for (int32_t v = 0; v <= I.n; ++v) I.a_off[(size_t)v] = (int32_t)(sb + I.off[(size_t)v]);
  I.a_dc.resize(I.dc.size());
  I.a_rev.resize(I.rev.size());
  for (int64_t j = 0; j < I.S; ++j) {
    I.a_dc[(size_t)(2 * j)] = vb + I.dc[(size_t)(2 * j)];
    I.a_dc[(size_t)(2 * j + 1)] = I.dc[(size_t)(2 * j + 1)];
    I.a_rev[(size_t)j] = (int32_t)(sb + I.rev[(size_t)j]);
  }
  hipStream_t st = a->copy_stream;
  HIP_TRY(hipMemcpyAsync(a->off + vb, I.a_off.data(), sizeof(int) * ((size_t)I.n + 1),
                         hipMemcpyHostToDevice, st));
  if (I.S) {
    HIP_TRY(hipMemcpyAsync(a->dc + sb, I.a_dc.data(), sizeof(int2) * (size_t)I.S,
                           hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(a->rev + sb, I.a_rev.data(), sizeof(int) * (size_t)I.S,
                           hipMemcpyHostToDevice, st));
  }
  HIP_TRY(hipMemcpyAsync(a->vcls + vb, I.vcls.data(), (size_t)I.n, hipMemcpyHostToDevice, st));
  hipEvent_t ev = nullptr;
  HIP_TRY(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  if (hipEventRecord(ev, st) != hipSuccess) {
    (void)hipEventDestroy(ev);
    return MF_ERR_DEVICE;
  }
  *done = ev;
  return MF_OK;
}

// 1: the copies behind the ticket have finished, 0: not yet, -1: the copy failed.
int serve_copy_state(void *done) {
  const hipError_t e = hipEventQuery((hipEvent_t)done);
  return e == hipSuccess ? 1 : (e == hipErrorNotReady ? 0 : -1);
}

void serve_copy_release(void *done) {
  if (done) (void)hipEventDestroy((hipEvent_t)done);
}

void serve_publish(ServeArena *a, int64_t q, const ServeInstance &I, int32_t vb) {
  int *en = a->entries + (size_t)(q % a->qcap) * kServeEntryInts;
  const int vals[kServeEntryInts] = {vb, I.n, vb + I.s, vb + I.t,
                                     I.pop[0], I.pop[1], I.pop[2], I.pop[3], I.pop[4], I.pop[5],
                                     I.bpop[0], I.bpop[1], I.bpop[2], I.bpop[3], I.bpop[4],
                                     I.bpop[5]};
  for (int i = 0; i < kServeEntryInts; ++i) __atomic_store_n(&en[i], vals[i], __ATOMIC_RELAXED);
  __atomic_store_n(&a->pub[0], (int)(q + 1), __ATOMIC_RELEASE);
}

bool serve_poll(ServeArena *a, int64_t q, int64_t *value, int32_t *status, double *device_ms) {
  ServeReport *r = a->reports + (size_t)(q % a->qcap);
  if (__atomic_load_n(&r->seq, __ATOMIC_ACQUIRE) != (int)(q + 1)) return false;
  *value = __atomic_load_n(&r->value, __ATOMIC_RELAXED);
  *status = __atomic_load_n(&r->status, __ATOMIC_RELAXED);
  const long long t0 = __atomic_load_n(&r->t_start, __ATOMIC_RELAXED);
  const long long t1 = __atomic_load_n(&r->t_end, __ATOMIC_RELAXED);
  *device_ms = (double)(t1 - t0) * 1e-5;  // 100 MHz wall clock
  return true;
}

bool serve_running(ServeArena *a) {
  return a->launched && hipStreamQuery(a->kernel_stream) == hipErrorNotReady;
}

int64_t serve_launches(const ServeArena *a) { return a->launches; }

int context_device(const mf_context *ctx) { return ctx->device; }

void *context_staging(mf_context *ctx, size_t bytes) {
  if (ctx->staging_bytes >= bytes) return ctx->staging;
  if (ctx->staging) (void)hipHostFree(ctx->staging);
  ctx->staging = nullptr;
  ctx->staging_bytes = 0;
  if (hipSetDevice(ctx->device) != hipSuccess ||
      hipHostMalloc(&ctx->staging, bytes, hipHostMallocDefault) != hipSuccess) {
    ctx->staging = nullptr;
    return nullptr;
  }
  ctx->staging_bytes = bytes;
  return ctx->staging;
}

mf_status serve_launch(ServeArena *a) {
  HIP_TRY(hipSetDevice(a->ctx->device));
  const ClassConfig vc = class_config(MF_ALGO_VC, 0);
  const ClassConfig cc = class_config(MF_ALGO_TC, 0);
  ServeArgs sa;
  memset(&sa, 0, sizeof(sa));
  BatchArgs &b = sa.b;
  b.h = a->h;
  b.e = a->e;
  b.lists = a->lists;
  b.N = (int)a->nv;
  b.vcls = a->vcls;
  b.ovf = a->words + 2;
  b.stats = a->stats;
  b.nclass = cc.nclass;
  for (int c = 0; c < kNumClasses; ++c) {
    b.lanes[c] = cc.lanes[c];
    b.blanes[c] = vc.lanes[c];
  }
  b.use_vcls = 0;
  b.ratio = 1.0f;  // cycles take about as long as the global relabel before them
  b.max_rounds = 1000000;
  b.max_cycles = kMaxAdaptiveCycles;
  b.fwd_hops = kDefaultForwardHops;
  b.stamp = a->stamp;
  int *host_entries = nullptr, *host_pub = nullptr;
  ServeReport *host_reports = nullptr;
  HIP_TRY(hipHostGetDevicePointer((void **)&host_entries, a->entries, 0));
  HIP_TRY(hipHostGetDevicePointer((void **)&host_reports, a->reports, 0));
  HIP_TRY(hipHostGetDevicePointer((void **)&host_pub, a->pub, 0));
  sa.entries = host_entries;
  sa.reports = host_reports;
  sa.published = host_pub;
  sa.closed = host_pub + 1;
  sa.claimed = a->words;
  sa.pub_seen = a->words + 1;
  sa.poll_at = (long long *)(a->words + 4);
  sa.last_claim = (long long *)(a->words + 6);
  sa.busy = a->words + 8;
  sa.qcap = a->qcap;
  sa.idle_ticks = kServeIdleTicks;
  const int nwg = std::max(1, kServeGroupsPerCU * a->ctx->num_cu);
  k_serve_wg<LayoutBCSR><<<nwg, kBlock, 0, a->kernel_stream>>>(LayoutBCSR{a->off, a->dc, a->rev},
                                                               sa);
  HIP_TRY(hipGetLastError());
  a->launched = true;
  a->launches++;
  return MF_OK;
}

}  // namespace mf

// ---- solving without allocating: device memory from a static per-device array -------------------
//
// mf_solve_noalloc keeps every device array of a solve (layout, vertex state, worklists, control
// words) in `scratch`, a static device array of kScratchBytes that every GPU the library runs on
// holds until the process ends.  Host code never passes `scratch` itself to a kernel or a copy:
// its device address comes from hipGetSymbolAddress, taken once per device when a context is
// created on it, and kept in g_scratch_ptr.

constexpr size_t kScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace mf {

constexpr int kScratchDevices = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[kScratchDevices];  // device address of `scratch`, per device

static mf_status static_scratch_setup(int device) {
  if (device < 0 || device >= kScratchDevices) return MF_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return MF_OK;
  void *p = nullptr;
  HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));  // the context's device is current
  if (!p) return MF_ERR_DEVICE;
  g_scratch_ptr[device] = p;
  return MF_OK;
}

// Places the BCSR arrays of an (n, m) graph and its solver state in the scratch array, every
// region aligned to 256 bytes (base null: only sizes).  Returns the bytes used, or -1 when they
// exceed kScratchBytes.
template <class T>
static void carve(char *base, size_t &at, bool &fits, T **p, size_t count) {
  at = (at + 255) & ~(size_t)255;
  const size_t bytes = sizeof(T) * std::max<size_t>(count, 1);
  if (at + bytes > kScratchBytes) fits = false;
  *p = (fits && base) ? reinterpret_cast<T *>(base + at) : nullptr;
  at += bytes;
}

static int64_t carve_static(char *base, int32_t n, int64_t m, int64_t giants, DevGraph &G) {
  size_t at = 0;
  bool fits = true;
  const size_t S = (size_t)(2 * m);
  carve(base, at, fits, &G.off, (size_t)n + 1);
  carve(base, at, fits, &G.dc, S);
  carve(base, at, fits, &G.rev, S);
  carve(base, at, fits, &G.cap0, S);
  carve(base, at, fits, &G.fslot, (size_t)m);
  carve(base, at, fits, &G.h, (size_t)n);
  carve(base, at, fits, &G.e, (size_t)n);
  carve(base, at, fits, &G.stamp, (size_t)n);
  carve(base, at, fits, &G.lists, (size_t)4 * n);
  carve(base, at, fits, &G.ctl, (size_t)CTL_SIZE);
  carve(base, at, fits, &G.stats, (size_t)kStatSlot * kStatSpread);
  carve(base, at, fits, &G.vcls, (size_t)n);
  carve(base, at, fits, &G.vorder, (size_t)n);
  carve(base, at, fits, &G.giant, (size_t)std::max<int64_t>(giants, 1) * kGiantWords);
  carve(base, at, fits, &G.src, 1);
  carve(base, at, fits, &G.snk, 1);
  carve(base, at, fits, &G.values, 1);
  carve(base, at, fits, &G.cycle_base, 1);
  carve(base, at, fits, &G.cls_tab, (size_t)4 * kNumClasses);
  return fits ? (int64_t)at : -1;
}

// Single-instance BCSR upload into the scratch array: the layout is built on the host (O(m)) and
// copied; nothing is allocated on the device.
static mf_status upload_static(mf_context *ctx, const mf_csr *c, DevGraph &G) {
  SlotLayout L;
  mf_status rc = build_slot_layout(c, L);
  if (rc != MF_OK) return rc;
  const int32_t n = c->n;
  std::vector<int64_t> degree((size_t)n);
  int64_t giants = 0;
  const ClassConfig vc = class_config(MF_ALGO_VC, 0);
  for (int32_t v = 0; v < n; ++v) {
    degree[(size_t)v] = L.off[(size_t)v + 1] - L.off[(size_t)v];
    if (vc.lanes[class_of_degree(vc, degree[(size_t)v])] == kSplitLanes) ++giants;
  }
  if (carve_static(nullptr, n, c->m, giants, G) < 0) return MF_ERR_OVERFLOW;
  char *base = static_cast<char *>(g_scratch_ptr[ctx->device]);
  if (!base) return MF_ERR_DEVICE;
  carve_static(base, n, c->m, giants, G);
  G.layout = MF_LAYOUT_BCSR;
  G.n = n;
  G.m = c->m;
  G.S = L.slots;
  G.k = 1;
  G.h_off.assign(c->offsets, c->offsets + n + 1);
  G.h_dst.assign(c->dst, c->dst + c->m);
  G.h_cap.resize((size_t)c->m);
  for (int64_t i = 0; i < c->m; ++i) G.h_cap[(size_t)i] = c->cap ? c->cap[i] : 1;
  G.h_src.assign(1, c->s);
  G.h_snk.assign(1, c->t);
  G.h_fslot = L.fslot;
  int64_t sc = 0;
  for (int32_t a = c->offsets[c->s]; a < c->offsets[c->s + 1]; ++a)
    if (c->dst[a] != c->s) sc += G.h_cap[(size_t)a];
  if (sc > INT32_MAX) return MF_ERR_OVERFLOW;  // every excess must fit int32
  std::vector<int2> dc((size_t)G.S);
  for (int64_t j = 0; j < G.S; ++j) dc[(size_t)j] = make_int2(L.dst[(size_t)j], L.cap0[(size_t)j]);
  hipStream_t st = ctx->stream;
  HIP_TRY(hipMemcpyAsync(G.off, L.off.data(), sizeof(int) * ((size_t)n + 1),
                         hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(G.dc, dc.data(), sizeof(int2) * (size_t)G.S, hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(G.rev, L.rev.data(), sizeof(int) * (size_t)G.S, hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(G.cap0, L.cap0.data(), sizeof(int) * (size_t)G.S, hipMemcpyHostToDevice,
                         st));
  HIP_TRY(hipMemcpyAsync(G.fslot, L.fslot.data(), sizeof(int) * (size_t)c->m,
                         hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(G.src, &c->s, sizeof(int), hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(G.snk, &c->t, sizeof(int), hipMemcpyHostToDevice, st));
  // control words start at zero (the previous call on this device left its own): a fill kernel
  k_fill_int<<<1, 256, 0, st>>>(G.ctl, 0, CTL_SIZE);
  HIP_TRY(hipGetLastError());
  if ((rc = set_vertex_classes(G, degree, st)) != MF_OK) return rc;
  G.max_src_deg = degree[(size_t)c->s];
  HIP_TRY(hipStreamSynchronize(st));
  G.fresh = true;  // the residuals were written as the capacities
  return MF_OK;
}

}  // namespace mf

struct mf_graph {
  mf::DevGraph g;
};

using namespace mf;

extern "C" {

mf_status mf_context_create(int device, mf_context **out) {
  if (!out) return MF_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return MF_ERR_NO_DEVICE;
  if (device < 0 || device >= count) return MF_ERR_INVALID_ARGUMENT;
  HIP_TRY(hipSetDevice(device));
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, device));
  mf_context *ctx = new mf_context();
  ctx->device = device;
  ctx->num_cu = prop.multiProcessorCount;
  int coop = 0;
  ctx->cooperative = hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, device) ==
                         hipSuccess &&
                     coop != 0;
  snprintf(ctx->name, sizeof(ctx->name), "%s", prop.name);
  snprintf(ctx->arch, sizeof(ctx->arch), "%s", prop.gcnArchName);
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&ctx->ev_solve[0]) != hipSuccess ||
      hipEventCreate(&ctx->ev_solve[1]) != hipSuccess ||
      hipEventCreate(&ctx->ev_solve[2]) != hipSuccess ||
      hipEventCreate(&ctx->ev_solve[3]) != hipSuccess) {
    mf_context_destroy(ctx);
    return MF_ERR_DEVICE;
  }
  const mf_status sc = static_scratch_setup(device);
  if (sc != MF_OK) {
    mf_context_destroy(ctx);
    return sc;
  }
  *out = ctx;
  return MF_OK;
}

mf_status mf_context_destroy(mf_context *ctx) {
  if (!ctx) return MF_ERR_INVALID_ARGUMENT;
  (void)hipSetDevice(ctx->device);
  for (hipEvent_t e : ctx->prof_ev) (void)hipEventDestroy(e);
  for (hipEvent_t e : ctx->ev_solve)
    if (e) (void)hipEventDestroy(e);
  if (ctx->staging) (void)hipHostFree(ctx->staging);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return MF_OK;
}

mf_status mf_device_synchronize(mf_context *ctx) {
  if (!ctx) return MF_ERR_INVALID_ARGUMENT;
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return MF_OK;
}

mf_status mf_device_info(mf_context *ctx, int32_t *num_cu, char *name, size_t name_len) {
  if (!ctx) return MF_ERR_INVALID_ARGUMENT;
  if (num_cu) *num_cu = ctx->num_cu;
  if (name && name_len) snprintf(name, name_len, "%s (%s)", ctx->name, ctx->arch);
  return MF_OK;
}

mf_status mf_graph_upload(mf_context *ctx, const mf_csr *g, mf_layout layout, mf_graph **out) {
  if (!ctx || !out) return MF_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  if (layout != MF_LAYOUT_RCSR && layout != MF_LAYOUT_BCSR && layout != MF_LAYOUT_BCSR2)
    return MF_ERR_INVALID_ARGUMENT;
  mf_status st = validate_csr(g);
  if (st != MF_OK) return st;
  HIP_TRY(hipSetDevice(ctx->device));
  mf_graph *mg = new mf_graph();
  const int32_t s = g->s, t = g->t;
  st = upload(ctx, g, &s, &t, 1, layout, mg->g);
  if (st != MF_OK) {
    free_dev(&mg->g);
    delete mg;
    return st;
  }
  *out = mg;
  return MF_OK;
}

mf_status mf_graph_solve(mf_context *ctx, mf_graph *g, mf_algo algo, const mf_options *opts,
                         mf_result *res) {
  if (!ctx || !g || (algo != MF_ALGO_TC && algo != MF_ALGO_VC)) return MF_ERR_INVALID_ARGUMENT;
  mf_options o;
  mf_options_default(&o);
  if (opts) o = *opts;
  mf_status st = check_options(o);
  if (st != MF_OK) return st;
  HIP_TRY(hipSetDevice(ctx->device));
  return solve_graph(ctx, g->g, algo, o, res);
}

mf_status mf_graph_read_flow(mf_context *ctx, mf_graph *mg, int32_t *flow_out,
                             int32_t return_excess) {
  if (!ctx || !mg || (!flow_out && mg->g.m > 0)) return MF_ERR_INVALID_ARGUMENT;
  DevGraph &G = mg->g;
  if (!G.solved) return MF_ERR_INVALID_ARGUMENT;
  HIP_TRY(hipSetDevice(ctx->device));
  if (G.layout != MF_LAYOUT_RCSR && G.h_fslot.empty() && G.m > 0) {
    G.h_fslot.resize((size_t)G.m);
    HIP_TRY(hipMemcpy(G.h_fslot.data(), G.fslot, sizeof(int) * G.m, hipMemcpyDeviceToHost));
  }
  if (G.layout == MF_LAYOUT_RCSR) {
    std::vector<int2> fdc((size_t)G.m);
    HIP_TRY(hipMemcpy(fdc.data(), G.fdc, sizeof(int2) * G.m, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < G.m; ++i) flow_out[i] = G.h_cap[(size_t)i] - fdc[(size_t)i].y;
  } else if (G.layout == MF_LAYOUT_BCSR) {
    std::vector<int2> dc((size_t)G.S);
    HIP_TRY(hipMemcpy(dc.data(), G.dc, sizeof(int2) * G.S, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < G.m; ++i)
      flow_out[i] = G.h_cap[(size_t)i] - dc[(size_t)G.h_fslot[(size_t)i]].y;
  } else {
    std::vector<int2> cr((size_t)G.S);
    HIP_TRY(hipMemcpy(cr.data(), G.cr, sizeof(int2) * G.S, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < G.m; ++i)
      flow_out[i] = G.h_cap[(size_t)i] - cr[(size_t)G.h_fslot[(size_t)i]].x;
  }
  if (return_excess) {
    mf_csr c;
    c.n = G.n;
    c.m = G.m;
    c.offsets = G.h_off.data();
    c.dst = G.h_dst.data();
    c.cap = G.h_cap.data();
    c.s = G.h_src[0];
    c.t = G.h_snk[0];
    return_excess_to_source(&c, flow_out);
  }
  return MF_OK;
}

mf_status mf_graph_read_workload(mf_context *ctx, mf_graph *mg, uint64_t *cycles,
                                 uint8_t *wave_class, int64_t capacity, int64_t *num_waves) {
  if (!ctx || !mg || !num_waves) return MF_ERR_INVALID_ARGUMENT;
  DevGraph &G = mg->g;
  *num_waves = G.work_waves;
  if (!G.work || (!cycles && !wave_class)) return MF_OK;
  if (capacity < G.work_waves) return MF_ERR_INVALID_ARGUMENT;
  HIP_TRY(hipSetDevice(ctx->device));
  if (cycles)
    HIP_TRY(hipMemcpy(cycles, G.work, sizeof(uint64_t) * (size_t)G.work_waves,
                      hipMemcpyDeviceToHost));
  if (wave_class)
    HIP_TRY(hipMemcpy(wave_class, G.work_cls, (size_t)G.work_waves, hipMemcpyDeviceToHost));
  return MF_OK;
}

mf_status mf_graph_read_layout(mf_context *ctx, mf_graph *mg, mf_layout_arrays *out) {
  if (!ctx || !mg || !out) return MF_ERR_INVALID_ARGUMENT;
  memset(out, 0, sizeof(*out));
  DevGraph &G = mg->g;
  HIP_TRY(hipSetDevice(ctx->device));
  auto alloc = [](int64_t k) { return (int32_t *)malloc(sizeof(int32_t) * (size_t)(k > 0 ? k : 1)); };
  auto pairs = [&](const int2 *d, int64_t k, int32_t *x, int32_t *y) -> hipError_t {
    std::vector<int2> h((size_t)k);
    hipError_t err = hipMemcpy(h.data(), d, sizeof(int2) * (size_t)k, hipMemcpyDeviceToHost);
    for (int64_t j = 0; j < k; ++j) {
      if (x) x[j] = h[(size_t)j].x;
      if (y) y[j] = h[(size_t)j].y;
    }
    return err;
  };
  out->n = G.n;
  if (G.layout == MF_LAYOUT_RCSR) {
    out->slots = G.m;
    out->a = alloc((int64_t)G.n + 1);
    out->b = alloc(G.m);
    out->c = alloc(G.m);
    if (!out->a || !out->b || !out->c) {
      mf_layout_free(out);
      return MF_ERR_OUT_OF_MEMORY;
    }
    HIP_TRY(hipMemcpy(out->a, G.roff, sizeof(int) * (G.n + 1), hipMemcpyDeviceToHost));
    HIP_TRY(pairs(G.rdi, G.m, out->b, out->c));
    return MF_OK;
  }
  out->slots = G.S;
  out->a = alloc((int64_t)G.n + 1);
  out->b = alloc(G.S);
  out->c = alloc(G.S);
  out->d = alloc(G.S);
  out->e = alloc(G.m);
  if (!out->a || !out->b || !out->c || !out->d || !out->e) {
    mf_layout_free(out);
    return MF_ERR_OUT_OF_MEMORY;
  }
  HIP_TRY(hipMemcpy(out->a, G.off, sizeof(int) * (G.n + 1), hipMemcpyDeviceToHost));
  if (G.layout == MF_LAYOUT_BCSR) {
    HIP_TRY(pairs(G.dc, G.S, out->b, nullptr));
    HIP_TRY(hipMemcpy(out->c, G.rev, sizeof(int) * G.S, hipMemcpyDeviceToHost));
  } else {
    HIP_TRY(hipMemcpy(out->b, G.dst, sizeof(int) * G.S, hipMemcpyDeviceToHost));
    HIP_TRY(pairs(G.cr, G.S, nullptr, out->c));
  }
  HIP_TRY(hipMemcpy(out->d, G.cap0, sizeof(int) * G.S, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(out->e, G.fslot, sizeof(int) * G.m, hipMemcpyDeviceToHost));
  return MF_OK;
}

mf_status mf_graph_read_cut(mf_context *ctx, mf_graph *mg, uint8_t *sink_side) {
  if (!ctx || !mg || !sink_side) return MF_ERR_INVALID_ARGUMENT;
  DevGraph &G = mg->g;
  if (!G.solved) return MF_ERR_INVALID_ARGUMENT;
  HIP_TRY(hipSetDevice(ctx->device));
  std::vector<int> h((size_t)G.n);
  HIP_TRY(hipMemcpy(h.data(), G.h, sizeof(int) * G.n, hipMemcpyDeviceToHost));
  // After the final global relabel, h < H exactly on the vertices that reach the sink.
  for (int32_t v = 0; v < G.n; ++v) sink_side[v] = h[(size_t)v] < G.n ? 1 : 0;
  return MF_OK;
}

mf_status mf_graph_read_state(mf_context *ctx, mf_graph *mg, int32_t *height, int32_t *excess) {
  if (!ctx || !mg) return MF_ERR_INVALID_ARGUMENT;
  DevGraph &G = mg->g;
  HIP_TRY(hipSetDevice(ctx->device));
  if (height) HIP_TRY(hipMemcpy(height, G.h, sizeof(int) * G.n, hipMemcpyDeviceToHost));
  if (excess) HIP_TRY(hipMemcpy(excess, G.e, sizeof(int) * G.n, hipMemcpyDeviceToHost));
  return MF_OK;
}

mf_status mf_graph_destroy(mf_context *ctx, mf_graph *mg) {
  if (!ctx || !mg) return MF_ERR_INVALID_ARGUMENT;
  (void)hipSetDevice(ctx->device);
  free_dev(&mg->g);
  delete mg;
  return MF_OK;
}

mf_status mf_solve_noalloc(mf_context *ctx, const mf_csr *g, mf_algo algo, mf_layout layout,
                           const mf_options *opts, mf_result *res, int32_t *flow_out) {
  if (!ctx || (algo != MF_ALGO_TC && algo != MF_ALGO_VC) || layout != MF_LAYOUT_BCSR)
    return MF_ERR_INVALID_ARGUMENT;
  mf_status st = validate_csr(g);
  if (st != MF_OK) return st;
  mf_options o;
  mf_options_default(&o);
  if (opts) o = *opts;
  if ((st = check_options(o)) != MF_OK) return st;
  // the options whose launches need memory of their own are off: captured cycle graphs,
  // per-wave workload records and sampled launch timing
  o.cycle_graphs = 0;
  o.record_workload = 0;
  o.profile_kernels = 0;
  HIP_TRY(hipSetDevice(ctx->device));
  const double t0 = now_ms();
  mf_graph view;  // every device pointer of view.g points into the scratch array: none is freed
  st = upload_static(ctx, g, view.g);
  const double t1 = now_ms();
  if (st == MF_OK) st = solve_graph(ctx, view.g, algo, o, res);
  if (res) res->upload_ms = t1 - t0;
  if (st == MF_OK && flow_out) st = mf_graph_read_flow(ctx, &view, flow_out, 1);
  return st;
}

mf_status mf_solve(mf_context *ctx, const mf_csr *g, mf_algo algo, mf_layout layout,
                   const mf_options *opts, mf_result *res, int32_t *flow_out) {

  const double t0 = now_ms();
  mf_graph *mg = nullptr;
  mf_status st = mf_graph_upload(ctx, g, layout, &mg);
  if (st != MF_OK) return st;
  const double t1 = now_ms();
  st = mf_graph_solve(ctx, mg, algo, opts, res);
  if (res) res->upload_ms = t1 - t0;
  if (st == MF_OK && flow_out) st = mf_graph_read_flow(ctx, mg, flow_out, 1);
  mf_graph_destroy(ctx, mg);
  return st;
}

// Reference-shaped entry point on host RCSR arrays (see mf.h).
mf_status mf_rcsr_push_relabel(mf_context *ctx, int algo_type, int V, int E, int source, int sink,
                               int *cpu_height, int *cpu_excess_flow, const int *cpu_offsets,
                               const int *cpu_destinations, const int *cpu_capacities,
                               int *cpu_fflows, int *cpu_bflows, const int *cpu_roffsets,
                               const int *cpu_rdestinations, const int *cpu_flow_idx,
                               int *excess_total, const mf_options *opts, mf_result *res) {
  // the per-arc arrays may be NULL when there are no arcs
  if (!ctx || !cpu_height || !cpu_excess_flow || !cpu_offsets || !cpu_roffsets || !excess_total ||
      (E > 0 && (!cpu_destinations || !cpu_capacities || !cpu_fflows || !cpu_bflows ||
                 !cpu_rdestinations || !cpu_flow_idx)) ||
      (algo_type != 0 && algo_type != 1))
    return MF_ERR_INVALID_ARGUMENT;
  mf_csr c;
  c.n = V;
  c.m = E;
  c.offsets = cpu_offsets;
  c.dst = cpu_destinations;
  c.cap = cpu_capacities;
  c.s = source;
  c.t = sink;
  mf_status st = validate_csr(&c);
  if (st != MF_OK) return st;
  if (cpu_roffsets[0] != 0 || cpu_roffsets[V] != E) return MF_ERR_INVALID_ARGUMENT;
  for (int v = 0; v < V; ++v)
    if (cpu_roffsets[v + 1] < cpu_roffsets[v]) return MF_ERR_INVALID_ARGUMENT;
  for (int q = 0; q < E; ++q)
    if (cpu_flow_idx[q] < 0 || cpu_flow_idx[q] >= E || cpu_rdestinations[q] < 0 ||
        cpu_rdestinations[q] >= V)
      return MF_ERR_INVALID_ARGUMENT;
  mf_options o;
  mf_options_default(&o);
  if (opts) o = *opts;
  if ((st = check_options(o)) != MF_OK) return st;
  HIP_TRY(hipSetDevice(ctx->device));
  DevGraph G;
  hipStream_t s = ctx->stream;
  G.layout = MF_LAYOUT_RCSR;
  G.n = V;
  G.m = E;
  G.S = E;
  G.k = 1;
  G.h_off.assign(cpu_offsets, cpu_offsets + V + 1);
  G.h_dst.assign(cpu_destinations, cpu_destinations + E);
  G.h_cap.assign(cpu_capacities, cpu_capacities + E);
  G.h_src.assign(1, source);
  G.h_snk.assign(1, sink);
  int64_t sc = 0;
  for (int32_t i = cpu_offsets[source]; i < cpu_offsets[source + 1]; ++i)
    if (cpu_destinations[i] != source) sc += cpu_capacities[i];
  auto fail = [&](mf_status code) {
    free_dev(&G);
    return code;
  };
  if (sc > INT32_MAX) return fail(MF_ERR_OVERFLOW);
  std::vector<int2> fdc((size_t)E), rdi((size_t)E);
  for (int i = 0; i < E; ++i) fdc[(size_t)i] = make_int2(cpu_destinations[i], cpu_capacities[i]);
  for (int q = 0; q < E; ++q) rdi[(size_t)q] = make_int2(cpu_rdestinations[q], cpu_flow_idx[q]);
  if ((st = dmalloc(&G.off, (size_t)V + 1)) || (st = dmalloc(&G.fdc, (size_t)E)) ||
      (st = dmalloc(&G.bcf, (size_t)E)) || (st = dmalloc(&G.roff, (size_t)V + 1)) ||
      (st = dmalloc(&G.rdi, (size_t)E)) || (st = dmalloc(&G.cap0, (size_t)E)) ||
      (st = dmalloc(&G.h, (size_t)V)) || (st = dmalloc(&G.e, (size_t)V)) ||
      (st = dmalloc(&G.stamp, (size_t)V)) || (st = dmalloc(&G.cycle_base, 1)) ||
      (st = dmalloc(&G.cls_tab, (size_t)4 * kNumClasses)) ||
      (st = dmalloc(&G.lists, (size_t)4 * V)) || (st = dmalloc(&G.ctl, (size_t)CTL_SIZE)) ||
      (st = dmalloc(&G.stats, (size_t)kStatSlot * kStatSpread)) || (st = dmalloc(&G.src, (size_t)1)) ||
      (st = dmalloc(&G.snk, (size_t)1)) || (st = dmalloc(&G.values, (size_t)1)))
    return fail(st);
  std::vector<int64_t> degree((size_t)V);
  for (int32_t v = 0; v < V; ++v)
    degree[(size_t)v] = (int64_t)(cpu_offsets[v + 1] - cpu_offsets[v]) +
                        (cpu_roffsets[v + 1] - cpu_roffsets[v]);
  if ((st = set_vertex_classes(G, degree, s)) != MF_OK) return fail(st);
  G.max_src_deg = degree[(size_t)source];
  // copies of the per-arc arrays only when there are arcs (their pointers may be NULL then)
  auto put = [&](void *dst, const void *src, size_t bytes) {
    return bytes == 0 || hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s) == hipSuccess;
  };
  auto get = [&](void *dst, const void *src, size_t bytes) {
    return bytes == 0 || hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) == hipSuccess;
  };
  bool ok = put(G.off, cpu_offsets, sizeof(int) * (V + 1)) &&
            put(G.fdc, fdc.data(), sizeof(int2) * E) &&
            put(G.roff, cpu_roffsets, sizeof(int) * (V + 1)) &&
            put(G.rdi, rdi.data(), sizeof(int2) * E) && put(G.cap0, cpu_capacities, sizeof(int) * E) &&
            hipMemcpyAsync(G.src, &source, sizeof(int), hipMemcpyHostToDevice, s) == hipSuccess &&
            hipMemcpyAsync(G.snk, &sink, sizeof(int), hipMemcpyHostToDevice, s) == hipSuccess &&
            hipStreamSynchronize(s) == hipSuccess;
  if (!ok) return fail(MF_ERR_DEVICE);
  mf_result r;
  st = solve_graph(ctx, G, algo_type == 0 ? MF_ALGO_TC : MF_ALGO_VC, o, &r);
  if (res) *res = r;
  if (st == MF_OK) {
    std::vector<int2> fd((size_t)E);
    ok = get(fd.data(), G.fdc, sizeof(int2) * E) && get(cpu_bflows, G.bcf, sizeof(int) * E) &&
         get(cpu_height, G.h, sizeof(int) * V) && get(cpu_excess_flow, G.e, sizeof(int) * V);
    if (!ok) return fail(MF_ERR_DEVICE);
    for (int i = 0; i < E; ++i) cpu_fflows[i] = fd[(size_t)i].y;
    *excess_total = cpu_excess_flow[sink];
  }
  return fail(st);
}

}  // extern "C"
