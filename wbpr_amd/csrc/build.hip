// Residual layouts built on the device from a CSR already in HBM.
//
// The slot layouts (BCSR, BCSR2) order the slots of a vertex x by arc index over every arc
// incident to x: out-arc i of x contributes its forward slot, in-arc i its reverse slot, a
// self-loop both (forward first).  That is the order of the host builder, which appends arcs in
// CSR order (maxflow-bcsr/src/graph.cpp:322-365 appends them in file order).  With the arcs
// sorted by head (a stable radix sort, so arc order within a head), every slot position is a
// closed form:
//
//   forward slot of arc i = u -> v:  soff[u] + (i - off[u]) + #{in-arcs j of u : j < i}
//   reverse slot of arc i:           soff[v] + #{out-arcs j of v : j < i} + #{in-arcs j of v : j < i}
//                                    (+1 for a self-loop: its forward slot comes first)
//
// with soff[x] = off[x] + in_off[x].  The RCSR reverse CSR is the sorted in-arc list itself.
//
// Both terms need no search per arc.  The in-arcs of u other than self-loops have arc indices
// outside u's out-arc range [off[u], off[u+1]), so #{in-arcs j of u : j < i} for an out-arc i of u
// is the same count before off[u] for every out-arc (one search per vertex, k_in_before), plus,
// only for a vertex with self-loops, those of its self-loops before i.  Likewise the out-arcs of
// v before an in-arc i from another tail are none (i < off[v]) or all of them (i >= off[v+1]),
// and soff[v] - in_off[v] = off[v], so the reverse slot is off[v] + that count + the arc's rank
// in the sorted in-arc list (pos_in[i] - in_off[v] + in_off[v]).
// Replaces the O(m) host pass (and the reference's O(sum of squared degrees) linear searches,
// maxflow-cuda/src/graph.cpp:386-399, maxflow-bcsr/src/graph.cpp:351-362).
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>

#include "build.h"

namespace mf {
namespace {

constexpr int kThreads = 256;

#define BUILD_TRY(expr)                       \
  do {                                        \
    if ((expr) != hipSuccess) {               \
      (void)hipGetLastError();                \
      return MF_ERR_DEVICE;                   \
    }                                         \
  } while (0)

// first index in a[0, len) with a[k] >= x (a sorted ascending)
__device__ __forceinline__ int64_t lower_bound(const int *a, int64_t len, int x) {
  int64_t lo = 0, hi = len;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (a[mid] < x) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

__global__ void k_tail_and_iota(const int *off, int n, int64_t m, int *tail, int *iota) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m; i += stride) {
    // tail = last u with off[u] <= i
    tail[i] = (int)lower_bound(off, (int64_t)n + 1, (int)i + 1) - 1;
    iota[i] = (int)i;
  }
}

__global__ void k_in_offsets(const int *sorted_heads, int64_t m, int n, int *in_off) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v <= n; v += stride)
    in_off[v] = (int)lower_bound(sorted_heads, m, (int)v);
}

__global__ void k_in_position(const int *in_list, int64_t m, int *pos_in) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < m; p += stride)
    pos_in[in_list[p]] = (int)p;
}

__global__ void k_slot_offsets(const int *off, const int *in_off, int n, int *soff) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v <= n; v += stride)
    soff[v] = off[v] + in_off[v];
}

// Per vertex u: the number of in-arcs of u whose arc index is below off[u], negated minus one
// when u has a self-loop (whose in-arc indices fall inside u's own out-arc range).
__global__ void k_in_before(const int *off, const int *in_off, const int *in_list, int n,
                            int *before) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t uu = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; uu < n; uu += stride) {
    const int u = (int)uu;
    const int lo = in_off[u], len = in_off[u + 1] - lo;
    const int cb = (int)lower_bound(in_list + lo, len, off[u]);
    const bool self = cb < len && in_list[lo + cb] < off[u + 1];
    before[u] = self ? -cb - 1 : cb;
  }
}

// Slot pair of every arc; layout 1 = BCSR ({dst, residual} + rev), 2 = BCSR2 (dst + {residual, rev}).
__global__ void k_place_slots(const int *off, const int *dst, const int *cap, const int *tail,
                              const int *in_off, const int *in_list, const int *pos_in,
                              const int *in_before, int64_t m, int layout, int2 *dc, int *rev,
                              int *dst2, int2 *cr, int *cap0, int *fslot) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t ii = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; ii < m; ii += stride) {
    const int i = (int)ii;
    const int u = tail[i], v = dst[i], c = cap ? cap[i] : 1;
    const int cb = in_before[u];
    int earlier = cb;  // in-arcs of u with an index below i
    if (cb < 0) {      // u has self-loops: count its in-arcs below i by search (rare)
      const int lo = in_off[u];
      earlier = (int)lower_bound(in_list + lo, in_off[u + 1] - lo, i);
    }
    const int a = in_off[u] + i + earlier;
    const int ov = off[v], ov1 = off[v + 1];
    const int before = u == v ? i - ov : (i >= ov1 ? ov1 - ov : 0);
    const int b = ov + before + pos_in[i] + (u == v ? 1 : 0);
    if (layout == 1) {
      dc[a] = make_int2(v, c);
      dc[b] = make_int2(u, 0);
      rev[a] = b;
      rev[b] = a;
    } else {
      dst2[a] = v;
      dst2[b] = u;
      cr[a] = make_int2(c, b);
      cr[b] = make_int2(0, a);
    }
    cap0[a] = c;
    cap0[b] = 0;
    fslot[i] = a;
  }
}

// RCSR: forward arcs in CSR order; the reverse CSR lists the in-arcs of every head in arc order
// as {tail, arc}.
__global__ void k_place_rcsr(const int *dst, const int *cap, const int *tail, const int *in_list,
                             int64_t m, int2 *fdc, int *cap0, int2 *rdi) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m; i += stride) {
    const int c = cap ? cap[i] : 1;
    fdc[i] = make_int2(dst[i], c);
    cap0[i] = c;
    const int a = in_list[i];
    rdi[i] = make_int2(tail[a], a);
  }
}

int grid_for(int64_t work) {
  return (int)std::min<int64_t>(std::max<int64_t>((work + kThreads - 1) / kThreads, 1), 65536);
}

template <class T>
hipError_t dalloc(T **p, int64_t count) {
  return hipMalloc((void **)p, sizeof(T) * (size_t)std::max<int64_t>(count, 1));
}

}  // namespace

mf_status build_layout_device(hipStream_t st, int32_t n, int64_t m, const int *off,
                              const int *dst, const int *cap, mf_layout layout,
                              const DeviceLayout &out, int *in_off_out) {
  int *tail = nullptr, *iota = nullptr, *heads = nullptr, *in_list = nullptr, *pos_in = nullptr;
  int *in_off = nullptr, *in_before = nullptr;
  void *tmp = nullptr;
  mf_status result = MF_ERR_DEVICE;
  auto cleanup = [&]() {
    void *ptrs[] = {tail, iota, heads, in_list, pos_in, in_off, in_before, tmp};
    for (void *p : ptrs)
      if (p) (void)hipFree(p);
  };
  do {
    if (dalloc(&tail, m) != hipSuccess || dalloc(&iota, m) != hipSuccess ||
        dalloc(&heads, m) != hipSuccess || dalloc(&in_list, m) != hipSuccess ||
        dalloc(&pos_in, m) != hipSuccess || dalloc(&in_off, (int64_t)n + 1) != hipSuccess ||
        dalloc(&in_before, (int64_t)n) != hipSuccess) {
      result = MF_ERR_OUT_OF_MEMORY;
      break;
    }
    const int ga = grid_for(m), gv = grid_for((int64_t)n + 1);
    k_tail_and_iota<<<ga, kThreads, 0, st>>>(off, n, m, tail, iota);
    if (hipGetLastError() != hipSuccess) break;
    // stable sort of the arcs by head: sorted heads and the arc list grouped by head
    unsigned bits = 1;
    while (bits < 31 && (1u << bits) <= (unsigned)n) ++bits;
    size_t tmp_bytes = 0;
    const unsigned *keys_in = reinterpret_cast<const unsigned *>(dst);
    unsigned *keys_out = reinterpret_cast<unsigned *>(heads);
    if (rocprim::radix_sort_pairs(nullptr, tmp_bytes, keys_in, keys_out, iota, in_list,
                                  (size_t)m, 0, bits, st) != hipSuccess)
      break;
    if (hipMalloc(&tmp, std::max<size_t>(tmp_bytes, 1)) != hipSuccess) {
      result = MF_ERR_OUT_OF_MEMORY;
      break;
    }
    if (rocprim::radix_sort_pairs(tmp, tmp_bytes, keys_in, keys_out, iota, in_list, (size_t)m, 0,
                                  bits, st) != hipSuccess)
      break;
    k_in_offsets<<<gv, kThreads, 0, st>>>(heads, m, n, in_off);
    if (hipGetLastError() != hipSuccess) break;
    if (layout == MF_LAYOUT_RCSR) {
      k_place_rcsr<<<ga, kThreads, 0, st>>>(dst, cap, tail, in_list, m, out.fdc, out.cap0,
                                           out.rdi);
      if (hipGetLastError() != hipSuccess) break;
      if (hipMemcpyAsync(out.roff, in_off, sizeof(int) * ((size_t)n + 1),
                         hipMemcpyDeviceToDevice, st) != hipSuccess)
        break;
    } else {
      k_in_position<<<ga, kThreads, 0, st>>>(in_list, m, pos_in);
      k_slot_offsets<<<gv, kThreads, 0, st>>>(off, in_off, n, out.soff);
      k_in_before<<<gv, kThreads, 0, st>>>(off, in_off, in_list, n, in_before);
      k_place_slots<<<ga, kThreads, 0, st>>>(off, dst, cap, tail, in_off, in_list, pos_in,
                                            in_before, m, layout == MF_LAYOUT_BCSR ? 1 : 2, out.dc,
                                            out.rev, out.dst2, out.cr, out.cap0, out.fslot);
      if (hipGetLastError() != hipSuccess) break;
    }
    if (in_off_out && hipMemcpyAsync(in_off_out, in_off, sizeof(int) * ((size_t)n + 1),
                                     hipMemcpyDeviceToHost, st) != hipSuccess)
      break;
    if (hipStreamSynchronize(st) != hipSuccess) break;
    result = MF_OK;
  } while (false);
  if (result != MF_OK) (void)hipStreamSynchronize(st);
  cleanup();
  return result;
}

}  // namespace mf
