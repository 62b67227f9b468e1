// Device-side building blocks of the push-relabel solver for gfx950 (wave64).
//
// Layout accessors: every kernel is written once against the small interface below and
// instantiated for the three residual layouts.  A "segment" is a contiguous slot range of a
// vertex: BCSR and BCSR2 have one segment per vertex (forward and reverse slots mixed), RCSR
// has two (forward arcs, then reverse arcs).  A "key" names one residual direction and is
// what a push acts on.
#pragma once

#include <hip/hip_runtime.h>
#include <climits>
#include <cstdint>

namespace mfk {

constexpr int kNumClasses = 6;   // degree classes of the worklists
constexpr int kBlock = 256;      // threads per block of the cycle / BFS-level kernels
// Giant vertices (more than kGiantSlots residual slots) form the last degree class.  A cycle
// splits each giant into parts of kPartSlots slots, one block per part: the largest hubs of a
// power-law graph hold hundreds of thousands of slots, and a single block scanning one would
// run alone for most of the launch.  kSplitLanes marks the split class in the lanes tables.
constexpr int kGiantSlots = 16384;
constexpr int kPartSlots = 4096;
constexpr int kSplitLanes = 4096;
constexpr int kWgBlock = 1024;   // threads of the one-workgroup BFS kernel
constexpr int kMultiPushLanes = 64;  // tiles this wide (and whole blocks) push along many arcs

__device__ __forceinline__ int ld_acq(const int *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_rel(int *p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---- residual layouts ------------------------------------------------------------------------

// BCSR re-laid for the scan: slot j holds {dst, residual} in one 8-byte word so the
// neighbour scan is a single coalesced 64-bit load per slot; the reverse slot index is a
// separate array read only by pushes and the BFS.
//
// dirty (optional): one byte per 64-byte line of slot pairs (slots 8i..8i+7) that a push changed,
// so the next solve restores only those lines to the capacities.  Null: not tracked.
struct LayoutBCSR {
  static constexpr int kSeg = 1;
  const int *off;
  int2 *dc;
  const int *rev;
  unsigned char *dirty;
  __device__ __forceinline__ int2 range(int u, int) const { return make_int2(off[u], off[u + 1]); }
  __device__ __forceinline__ void scan(int, int j, int &w, int &r, int &key) const {
    int2 x = dc[j];
    w = x.x;
    r = x.y;
    key = j;
  }
  __device__ __forceinline__ int target(int key) const { return dc[key].x; }
  __device__ __forceinline__ int residual(int key) const { return dc[key].y; }
  __device__ __forceinline__ int *res_ptr(int key) const { return &dc[key].y; }
  __device__ __forceinline__ void push(int key, int d) const {
    const int rk = rev[key];
    atomicSub(&dc[key].y, d);
    atomicAdd(&dc[rk].y, d);
    if (dirty) {
      dirty[key >> 3] = 1;
      dirty[rk >> 3] = 1;
    }
  }
  // BFS, two stages so a lane can keep several slots in flight:
  //   bfs_slot(j)   -> neighbour w of slot j and the index of the residual of the arc w -> u
  //   bfs_rin(idx)  -> that residual
  __device__ __forceinline__ void bfs_slot(int, int j, int &w, int &idx) const {
    w = dc[j].x;
    idx = rev[j];
  }
  __device__ __forceinline__ int bfs_rin(int, int idx) const { return dc[idx].y; }
  __device__ __forceinline__ int degree(int u) const { return off[u + 1] - off[u]; }
};

// BCSR2 (maxflow-bcsr2 layout): dst array plus interleaved {residual, reverse slot} pairs.
struct LayoutBCSR2 {
  static constexpr int kSeg = 1;
  const int *off;
  const int *dst;
  int2 *cr;
  unsigned char *dirty;  // as in LayoutBCSR, over the {residual, reverse slot} pairs
  __device__ __forceinline__ int2 range(int u, int) const { return make_int2(off[u], off[u + 1]); }
  __device__ __forceinline__ void scan(int, int j, int &w, int &r, int &key) const {
    w = dst[j];
    r = cr[j].x;
    key = j;
  }
  __device__ __forceinline__ int target(int key) const { return dst[key]; }
  __device__ __forceinline__ int residual(int key) const { return cr[key].x; }
  __device__ __forceinline__ int *res_ptr(int key) const { return &cr[key].x; }
  __device__ __forceinline__ void push(int key, int d) const {
    const int rv = cr[key].y;
    atomicSub(&cr[key].x, d);
    atomicAdd(&cr[rv].x, d);
    if (dirty) {
      dirty[key >> 3] = 1;
      dirty[rv >> 3] = 1;
    }
  }
  __device__ __forceinline__ void bfs_slot(int, int j, int &w, int &idx) const {
    w = dst[j];
    idx = cr[j].y;
  }
  __device__ __forceinline__ int bfs_rin(int, int idx) const { return cr[idx].x; }
  __device__ __forceinline__ int degree(int u) const { return off[u + 1] - off[u]; }
};

// RCSR (maxflow-cuda layout): forward arcs {dst, forward residual} per arc, backward residual
// per arc, reverse CSR {tail, arc index} per in-arc.  Key: arc a (< m) for the forward
// direction, m + k for reverse position k.
struct LayoutRCSR {
  static constexpr int kSeg = 2;
  const int *off;
  int2 *fdc;
  int *bcf;
  const int *roff;
  const int2 *rdi;
  int m;
  __device__ __forceinline__ int2 range(int u, int seg) const {
    return seg == 0 ? make_int2(off[u], off[u + 1]) : make_int2(roff[u], roff[u + 1]);
  }
  __device__ __forceinline__ void scan(int seg, int j, int &w, int &r, int &key) const {
    if (seg == 0) {
      int2 x = fdc[j];
      w = x.x;
      r = x.y;
      key = j;
    } else {
      int2 y = rdi[j];
      w = y.x;
      r = bcf[y.y];
      key = m + j;
    }
  }
  __device__ __forceinline__ int target(int key) const {
    return key < m ? fdc[key].x : rdi[key - m].x;
  }
  __device__ __forceinline__ int residual(int key) const {
    return key < m ? fdc[key].y : bcf[rdi[key - m].y];
  }
  __device__ __forceinline__ int *res_ptr(int key) const {
    return key < m ? &fdc[key].y : &bcf[rdi[key - m].y];
  }
  __device__ __forceinline__ void push(int key, int d) const {
    if (key < m) {
      atomicSub(&fdc[key].y, d);
      atomicAdd(&bcf[key], d);
    } else {
      const int a = rdi[key - m].y;
      atomicSub(&bcf[a], d);
      atomicAdd(&fdc[a].y, d);
    }
  }
  __device__ __forceinline__ void bfs_slot(int seg, int j, int &w, int &idx) const {
    if (seg == 0) {  // forward arc u -> w: the arc w -> u is its backward direction
      w = fdc[j].x;
      idx = j;
    } else {  // in-arc a = w -> u: its forward residual
      const int2 y = rdi[j];
      w = y.x;
      idx = y.y;
    }
  }
  __device__ __forceinline__ int bfs_rin(int seg, int idx) const {
    return seg == 0 ? bcf[idx] : fdc[idx].y;
  }
  __device__ __forceinline__ int degree(int u) const {
    return off[u + 1] - off[u] + roff[u + 1] - roff[u];
  }
};

// ---- worklists -------------------------------------------------------------------------------

// One list set: per degree class a region of the item array and a counter.  An append past
// the region's capacity is dropped and raises *ovf (the host turns that into an error), so a
// logic error can never write outside the allocation.
struct ListSet {
  int *items;
  // region offset of class c at cls[c], its capacity at cls[kNumClasses + c]; the table lives in
  // device memory (or LDS), not in the descriptor: a descriptor copied to a local variable and
  // indexed by class at run time would otherwise be placed in scratch memory
  const int *cls;
  int *cnt;     // kNumClasses counters, cnt[c * cstride]
  int cstride;  // kCtlStride for counters in device memory, 1 for counters in LDS
  int *ovf;
};

// Counters in device memory sit one per 128-byte line: atomics on one line serialise (about
// 11 ns each on gfx950), so classes must not share a line.
constexpr int kCtlStride = 32;

// a[i] for a run-time i without addressing the array: a chain of selects keeps a by-value
// argument struct in registers (an indexed access would copy it to scratch memory).
template <typename V, int N>
__device__ __forceinline__ V pick(const V (&a)[N], int i) {
  V v = a[0];
#pragma unroll
  for (int q = 1; q < N; ++q) v = (i == q) ? a[q] : v;
  return v;
}

// Value of lane `src` of the wave (src the same in every lane, e.g. from a ballot).
__device__ __forceinline__ int wave_lane(int v, int src) {
  return __builtin_amdgcn_readlane(v, src);
}

// Number of items of class c, clamped to the region capacity.  The counter was completed by an
// earlier kernel or, inside the one-workgroup BFS, lives in LDS behind a workgroup barrier.
__device__ __forceinline__ int list_count(const ListSet &L, int c) {
  const int k = L.cnt[c * L.cstride];
  const int cap = L.cls[kNumClasses + c];
  return k < cap ? k : cap;
}

struct Dispatch {
  int nclass;
  int lanes[kNumClasses];
  int blk_begin[kNumClasses + 1];
};

// Wave-aggregated append: one atomic per wave and class, lanes get consecutive slots
// (ballot + popcount of lower lanes).
__device__ __forceinline__ void append(const ListSet &L, int item, bool want, int cls,
                                       int nclass) {
  const int lane = __lane_id();
  if (__ballot(want) == 0) return;
#pragma unroll
  for (int c = 0; c < kNumClasses; ++c) {
    if (c >= nclass) break;
    const bool mine = want && cls == c;
    const unsigned long long mask = __ballot(mine);
    if (mask == 0) continue;
    const int leader = __ffsll((long long)mask) - 1;
    const int rank = __popcll(mask & ((1ull << lane) - 1ull));
    int base = 0;
    if (lane == leader) base = atomicAdd(&L.cnt[c * L.cstride], __popcll(mask));
    base = wave_lane(base, leader);
    if (mine) {
      if (base + rank < L.cls[kNumClasses + c]) L.items[L.cls[c] + base + rank] = item;
      else *L.ovf = 1;
    }
  }
}

// ---- appends staged in LDS ------------------------------------------------------------------
//
// Grid-wide kernels stage their appends in LDS and reserve space in the device list once per
// block and class at the end (stage_flush), instead of once per wave and class: a 10,000-block
// launch would otherwise queue tens of thousands of atomics on the counter lines.  Appends that
// do not fit the stage go to the device list directly.

constexpr int kStageCap = 256;  // items per class a block stages

struct Stage {
  int cnt[kNumClasses];
  int base[kNumClasses];
  int items[kNumClasses][kStageCap];
};

__device__ __forceinline__ void stage_init(Stage &S) {
  if (threadIdx.x < (unsigned)kNumClasses) S.cnt[threadIdx.x] = 0;
  __syncthreads();
}

__device__ __forceinline__ void staged_append(Stage &S, const ListSet &L, int item, bool want,
                                              int cls, int nclass) {
  const int lane = __lane_id();
  if (__ballot(want) == 0) return;
#pragma unroll
  for (int c = 0; c < kNumClasses; ++c) {
    if (c >= nclass) break;
    const bool mine = want && cls == c;
    const unsigned long long mask = __ballot(mine);
    if (mask == 0) continue;
    const int leader = __ffsll((long long)mask) - 1;
    const int rank = __popcll(mask & ((1ull << lane) - 1ull));
    int base = 0;
    if (lane == leader) base = atomicAdd(&S.cnt[c], __popcll(mask));
    base = wave_lane(base, leader);
    const bool fits = base + rank < kStageCap;
    if (mine && fits) S.items[c][base + rank] = item;
    const unsigned long long spill = __ballot(mine && !fits);
    if (spill == 0) continue;
    const int sl = __ffsll((long long)spill) - 1;
    const int srank = __popcll(spill & ((1ull << lane) - 1ull));
    int gbase = 0;
    if (lane == sl) gbase = atomicAdd(&L.cnt[c * L.cstride], __popcll(spill));
    gbase = wave_lane(gbase, sl);
    if (mine && !fits) {
      if (gbase + srank < L.cls[kNumClasses + c]) L.items[L.cls[c] + gbase + srank] = item;
      else *L.ovf = 1;
    }
  }
}

// Every thread of the block calls this once, after its last staged_append.
__device__ __forceinline__ void stage_flush(Stage &S, const ListSet &L, int nclass) {
  __syncthreads();
  if (threadIdx.x < (unsigned)nclass) {
    const int c = threadIdx.x;
    const int k = min(S.cnt[c], kStageCap);
    S.base[c] = k ? atomicAdd(&L.cnt[c * L.cstride], k) : 0;
  }
  __syncthreads();
#pragma unroll
  for (int c = 0; c < kNumClasses; ++c) {
    if (c >= nclass) break;
    const int k = min(S.cnt[c], kStageCap);
    const int base = S.base[c], cap = L.cls[kNumClasses + c], off = L.cls[c];
    for (int i = (int)threadIdx.x; i < k; i += (int)blockDim.x) {
      if (base + i < cap) L.items[off + base + i] = S.items[c][i];
      else *L.ovf = 1;
    }
  }
}

__device__ __forceinline__ unsigned long long pack_hk(int h, int key) {
  return ((unsigned long long)(unsigned)h << 32) | (unsigned)key;
}

// ---- cross-lane steps without the LDS crossbar ----------------------------------------------
//
// Lanes of a tile of T lanes (T a power of two dividing 64, tiles aligned in the wave) exchange
// values through DPP lane modifiers and gfx950's permlane swaps, which are VALU operations,
// instead of ds_bpermute / ds_swizzle, which go through the LDS unit and wait for it.
//
// Step OFF of a reduction gives every lane the value of one lane in the other half of its
// aligned group of 2*OFF lanes:
//   OFF 1, 2   quad permutations [1,0,3,2] and [2,3,0,1]            (DPP quad_perm)
//   OFF 4      lane i <-> 7 - i in each group of 8                   (DPP row_half_mirror)
//   OFF 8      lane i <-> 15 - i in each row of 16                   (DPP row_mirror)
//   OFF 16     rows 0 <-> 1 and 2 <-> 3                              (v_permlane16_swap)
//   OFF 32     lanes 0..31 <-> 32..63                                (v_permlane32_swap)
// The mirrors are not lane ^ OFF, but after steps 1 .. OFF/2 every lane of a group of OFF
// holds the group's reduction, so any partner in the other half completes the next level.
// The permlane swaps return two registers that hold the lane's own value and its partner's;
// the reductions below only combine them with commutative operations.
enum : int {
  kDppQuadXor1 = 0xB1,  // quad_perm [1,0,3,2]
  kDppQuadXor2 = 0x4E,  // quad_perm [2,3,0,1]
  kDppRowShr = 0x110,   // + n: row_shr:n (n = 1..15)
  kDppRowMirror = 0x140,
  kDppRowHalfMirror = 0x141,
  kDppRowBcast15 = 0x142,
  kDppRowBcast31 = 0x143,
};

template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ int dpp(int x) {
  // lanes whose row is masked off, or whose source lane is outside the row, receive 0
  return __builtin_amdgcn_update_dpp(0, x, CTRL, ROW_MASK, 0xF, true);
}

// The two registers {own, partner} of reduction step OFF (see above), for 32-bit x.
template <int OFF>
__device__ __forceinline__ void xpair(int x, int &p0, int &p1) {
  if constexpr (OFF == 1) {
    p0 = x;
    p1 = dpp<kDppQuadXor1>(x);
  } else if constexpr (OFF == 2) {
    p0 = x;
    p1 = dpp<kDppQuadXor2>(x);
  } else if constexpr (OFF == 4) {
    p0 = x;
    p1 = dpp<kDppRowHalfMirror>(x);
  } else if constexpr (OFF == 8) {
    p0 = x;
    p1 = dpp<kDppRowMirror>(x);
  } else if constexpr (OFF == 16) {
    const auto r = __builtin_amdgcn_permlane16_swap(x, x, false, false);
    p0 = (int)r[0];
    p1 = (int)r[1];
  } else {
    static_assert(OFF == 32, "reduction steps are 1, 2, 4, 8, 16, 32 lanes");
    const auto r = __builtin_amdgcn_permlane32_swap(x, x, false, false);
    p0 = (int)r[0];
    p1 = (int)r[1];
  }
}

// 64-bit values: both halves take the same route, so each returned pair is one lane's value.
template <int OFF>
__device__ __forceinline__ void xpair64(unsigned long long x, unsigned long long &p0,
                                        unsigned long long &p1) {
  int l0, l1, h0, h1;
  xpair<OFF>((int)(unsigned)x, l0, l1);
  xpair<OFF>((int)(unsigned)(x >> 32), h0, h1);
  p0 = ((unsigned long long)(unsigned)h0 << 32) | (unsigned)l0;
  p1 = ((unsigned long long)(unsigned)h1 << 32) | (unsigned)l1;
}

template <int T>
__device__ __forceinline__ unsigned long long tile_min(unsigned long long v) {
  static_assert(T >= 1 && T <= 64 && (T & (T - 1)) == 0, "tile of 1..64 lanes");
  if constexpr (T > 1) {
    unsigned long long a, b;
    xpair64<T / 2>(tile_min<T / 2>(v), a, b);
    return a < b ? a : b;
  }
  return v;
}

template <int T>
__device__ __forceinline__ int tile_max(int v) {
  static_assert(T >= 1 && T <= 64 && (T & (T - 1)) == 0, "tile of 1..64 lanes");
  if constexpr (T > 1) {
    int a, b;
    xpair<T / 2>(tile_max<T / 2>(v), a, b);
    return max(a, b);
  }
  return v;
}

template <int T>
__device__ __forceinline__ int tile_sum32(int v) {
  static_assert(T >= 1 && T <= 64 && (T & (T - 1)) == 0, "tile of 1..64 lanes");
  if constexpr (T > 1) {
    int a, b;
    xpair<T / 2>(tile_sum32<T / 2>(v), a, b);
    return a + b;
  }
  return v;
}

template <int T>
__device__ __forceinline__ long long tile_sum(long long v) {
  static_assert(T >= 1 && T <= 64 && (T & (T - 1)) == 0, "tile of 1..64 lanes");
  if constexpr (T > 1) {
    unsigned long long a, b;
    xpair64<T / 2>((unsigned long long)tile_sum<T / 2>(v), a, b);
    return (long long)(a + b);
  }
  return v;
}

// Value of the tile's first lane, in every lane of the tile.
template <int T>
__device__ __forceinline__ int tile_first(int v) {
  if constexpr (T == 1) {
    return v;
  } else if constexpr (T == 64) {
    return __builtin_amdgcn_readlane(v, 0);
  } else {
    return tile_sum32<T>((__lane_id() & (T - 1)) == 0 ? v : 0);
  }
}

// One step of an inclusive scan: lanes whose position in their row of 16 is >= N add the value
// N lanes below them.
template <int N>
__device__ __forceinline__ unsigned long long scan_row_step(unsigned long long v) {
  const int lo = dpp<kDppRowShr + N>((int)(unsigned)v);
  const int hi = dpp<kDppRowShr + N>((int)(unsigned)(v >> 32));
  const unsigned long long y = ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo;
  return (__lane_id() & 15) >= N ? v + y : v;
}

// Inclusive prefix sum over the lanes of each tile of T lanes, in lane order: row shifts inside
// each row of 16, then the total of row 0 (2) broadcast into row 1 (3), then the total of lanes
// 0..31 into lanes 32..63 (DPP row_bcast).
template <int T>
__device__ __forceinline__ long long tile_scan(long long sv) {
  static_assert(T >= 1 && T <= 64 && (T & (T - 1)) == 0, "tile of 1..64 lanes");
  unsigned long long v = (unsigned long long)sv;
  if constexpr (T >= 2) v = scan_row_step<1>(v);
  if constexpr (T >= 4) v = scan_row_step<2>(v);
  if constexpr (T >= 8) v = scan_row_step<4>(v);
  if constexpr (T >= 16) v = scan_row_step<8>(v);
  if constexpr (T >= 32) {
    const int lo = dpp<kDppRowBcast15, 0xA>((int)(unsigned)v);
    const int hi = dpp<kDppRowBcast15, 0xA>((int)(unsigned)(v >> 32));
    v += ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo;
  }
  if constexpr (T >= 64) {
    const int lo = dpp<kDppRowBcast31, 0xC>((int)(unsigned)v);
    const int hi = dpp<kDppRowBcast31, 0xC>((int)(unsigned)(v >> 32));
    v += ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo;
  }
  return (long long)v;
}

// Statistics accumulated per thread and flushed once per block.  32-bit per lane (a lane's counts
// within one launch, or one instance of a persistent kernel, which flushes after each, stay far
// below 2^32): 64-bit counters cost k_cycle the registers of a sixth wave per SIMD.
struct Stats {
  unsigned active, slots, pushes, relabels;
};

enum StatIdx {
  ST_ACTIVE = 0,
  ST_SLOTS = 1,
  ST_PUSHES = 2,
  ST_RELABELS = 3,
  ST_BFS_VERTS = 4,
  ST_BFS_SLOTS = 5,
  ST_SOURCE_OUT = 6,
  ST_BU_VERTS = 8,  // bottom-up BFS levels: unreached vertices examined
  ST_BU_SLOTS = 9,  //   and slots they scanned
  ST_COUNT = 12
};

// Blocks [0, nblk) split into groups of max(1, nblk / cnt) blocks; group q takes items q,
// q + ngroups, ... and its width = blocks * blockDim threads share each item.  Leftover blocks
// get group >= cnt and no items.
struct BlockGroups {
  int group, ngroups, lane, width;
};

__device__ __forceinline__ BlockGroups block_groups(int b, int nblk, int cnt) {
  const int gsz = max(1, nblk / max(cnt, 1));
  BlockGroups bg;
  bg.ngroups = max(1, nblk / gsz);
  bg.group = b / gsz;
  bg.lane = (b % gsz) * (int)blockDim.x + (int)threadIdx.x;
  bg.width = gsz * (int)blockDim.x;
  return bg;
}

// Statistics live in kStatSpread copies of kStatSlot counters, one 128-byte line per copy;
// block b adds into copy b % kStatSpread and the host sums the copies.  (Atomics on one line
// serialise at about 11 ns each: one copy for every block of a 10,000-block launch would cost
// tenths of a millisecond.)
constexpr int kStatSlot = 16;
constexpr int kStatSpread = 64;

__device__ __forceinline__ void block_flush(long long a, long long b, long long c, long long d,
                                            unsigned long long *out, int i0) {
  __shared__ long long red[4][kWgBlock / 64];
  __syncthreads();  // a previous flush in the same kernel may still be reading red
  a = tile_sum<64>(a);
  b = tile_sum<64>(b);
  c = tile_sum<64>(c);
  d = tile_sum<64>(d);
  const int w = threadIdx.x / 64;
  if (__lane_id() == 0) {
    red[0][w] = a;
    red[1][w] = b;
    red[2][w] = c;
    red[3][w] = d;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    long long s = 0;
    for (int k = 0; k < (int)(blockDim.x / 64); ++k) s += red[threadIdx.x][k];
    if (s)
      atomicAdd(&out[(size_t)(blockIdx.x % kStatSpread) * kStatSlot + i0 + threadIdx.x],
                (unsigned long long)s);
  }
}

}  // namespace mfk
