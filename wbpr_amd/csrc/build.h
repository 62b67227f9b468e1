// Device-side construction of the residual layouts (build.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/mf.h"

namespace mf {

// Output arrays, allocated by the caller in device memory (the ones of the chosen layout).
struct DeviceLayout {
  int *soff = nullptr;   // BCSR, BCSR2: n+1 slot offsets
  int2 *dc = nullptr;    // BCSR: S {dst, residual}
  int *rev = nullptr;    // BCSR: S reverse slot
  int *dst2 = nullptr;   // BCSR2: S dst
  int2 *cr = nullptr;    // BCSR2: S {residual, reverse slot}
  int *fslot = nullptr;  // BCSR, BCSR2: m, forward slot of every arc
  int *cap0 = nullptr;   // S initial residual (m for RCSR)
  int2 *fdc = nullptr;   // RCSR: m {dst, forward residual}
  int *roff = nullptr;   // RCSR: n+1 reverse offsets
  int2 *rdi = nullptr;   // RCSR: m {tail, arc}
};

// off (n+1), dst (m) and cap (m, or null for unit capacities) are device arrays of a CSR graph.
// in_off_out (host, n+1, optional) receives the in-degree offsets.  Synchronises st.
mf_status build_layout_device(hipStream_t st, int32_t n, int64_t m, const int *off,
                              const int *dst, const int *cap, mf_layout layout,
                              const DeviceLayout &out, int *in_off_out);

}  // namespace mf
