// Device helpers of the immersed-boundary stencil claims: the Peskin kernel
// and a wavefront-aggregated counter update.
#pragma once
#include "mc_common.h"
#include "mc_internal.h"

namespace mc {

// fluid.py:31-38 -- Peskin 4-point cosine kernel (float32 cosine).
__device__ __forceinline__ float peskin(float r, float h) {
  float ra = fabsf(r) / h;
  return (ra <= 2.0f) ? 0.25f * (1.0f + fcos((MC_PI_F * ra) / 2.0f)) : 0.0f;
}

// Wavefront-aggregated atomic add: one device-scope atomic per wavefront
// instead of one per lane (a single counter hit by every lane of a large
// launch serialises).  Every lane of the wavefront must call it; returns this
// lane's offset, lanes in lane order.
__device__ __forceinline__ int wave_fetch_add(int* ctr, int v) {
  const int lane = threadIdx.x & 63;
  int inc = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int u = __shfl_up(inc, d, 64);
    if (lane >= d) inc += u;
  }
  const int total = __shfl(inc, 63, 64);
  int base = 0;
  if (lane == 63 && total > 0)
    base = __hip_atomic_fetch_add(ctr, total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  base = __shfl(base, 63, 64);
  return base + inc - v;
}

// Lanes per node in the stencil kernel: one lane per tap, the node's group a
// power of two within a wavefront.
template <int K>
constexpr int stencil_lanes() {
  return K * K <= 1 ? 1 : (K * K <= 4 ? 4 : (K * K <= 16 ? 16 : (K * K <= 32 ? 32 : 64)));
}

}  // namespace mc
