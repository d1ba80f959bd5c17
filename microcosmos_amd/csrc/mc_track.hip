// Swimmer centre of mass and per-step displacement tracking
// (experiments/swim_sweep/swimmers/swim_base.py:167-209; the rollout fitness
// of examples/01_es_locomotion.ipynb).
//
// One workgroup per node group.  The mean is a fixed-order float32 sum:
// thread i folds the group's nodes i, i + 256, i + 512, ... in order, then
// the 256 partials are halved pairwise (p[i] += p[i + s], s = 128 .. 1), then
// divided by the node count.  The oracle restates exactly this order.
#include "mc_common.h"
#include "mc_internal.h"

namespace mc {
namespace {

constexpr int COM_THREADS = 256;

// Group centre: periodic != 0 -> reference-node minimum-image mean, wrapped
// (swim_base.py:167-177); periodic == 0 -> plain mean (the notebook's
// position[s].mean(axis=0)).  Result in (cx, cy) of thread 0.
__device__ __forceinline__ void group_com(const float* __restrict__ pos, int lo, int hi, float W,
                                          float H, int periodic, float* sx, float* sy, float& cx,
                                          float& cy) {
  const int n = hi - lo;
  float rx = 0.0f, ry = 0.0f;
  if (periodic && n > 0) {
    rx = pos[2 * lo];
    ry = pos[2 * lo + 1];
  }
  float ax = 0.0f, ay = 0.0f;
  for (int j = threadIdx.x; j < n; j += COM_THREADS) {
    float dx = pos[2 * (lo + j)], dy = pos[2 * (lo + j) + 1];
    if (periodic) {
      dx = min_image(dx, rx, W);
      dy = min_image(dy, ry, H);
    }
    ax = ax + dx;
    ay = ay + dy;
  }
  sx[threadIdx.x] = ax;
  sy[threadIdx.x] = ay;
  __syncthreads();
  for (int s = COM_THREADS / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      sx[threadIdx.x] = sx[threadIdx.x] + sx[threadIdx.x + s];
      sy[threadIdx.x] = sy[threadIdx.x] + sy[threadIdx.x + s];
    }
    __syncthreads();
  }
  const float fn = (float)n;
  const float mx = sx[0] / fn, my = sy[0] / fn;
  if (periodic) {
    cx = n > 0 ? fmodf_floor(rx + mx, W) : mx;
    cy = n > 0 ? fmodf_floor(ry + my, H) : my;
  } else {
    cx = mx;
    cy = my;
  }
}

__global__ void __launch_bounds__(COM_THREADS) k_com(const float* __restrict__ pos,
                                                     const int* __restrict__ off, float W, float H,
                                                     int periodic, float* __restrict__ com) {
  __shared__ float sx[COM_THREADS], sy[COM_THREADS];
  const int g = blockIdx.x;
  float cx, cy;
  group_com(pos, off[g], off[g + 1], W, H, periodic, sx, sy, cx, cy);
  if (threadIdx.x == 0) {
    com[2 * g] = cx;
    com[2 * g + 1] = cy;
  }
}

// swim_base.py:179-209: reset (mode 0) sets prev = centre, cumulative = 0 and
// the group's step count to 0; resume (mode 2) does the same but keeps
// cumulative; update (mode 1) advances the count and, past
// start_step, adds the minimum-image step displacement of the centre to
// cumulative; prev = centre.
__global__ void __launch_bounds__(COM_THREADS) k_track(const float* __restrict__ pos,
                                                       const int* __restrict__ off,
                                                       const int2* __restrict__ dims, int mode,
                                                       int start_step,
                                                       float* __restrict__ prev,
                                                       float* __restrict__ cum,
                                                       long long* __restrict__ count) {
  __shared__ float sx[COM_THREADS], sy[COM_THREADS];
  const int g = blockIdx.x;
  const float H = (float)dims[g].x, W = (float)dims[g].y;
  float cx, cy;
  group_com(pos, off[g], off[g + 1], W, H, 1, sx, sy, cx, cy);
  if (threadIdx.x != 0) return;
  if (mode == 0 || mode == 2) {
    if (mode == 0) {
      cum[2 * g] = 0.0f;
      cum[2 * g + 1] = 0.0f;
    }
    count[g] = 0;
  } else {
    const long long k = count[g] + 1;  // this step's number since the reset
    count[g] = k;
    if (k > start_step) {
      cum[2 * g] = cum[2 * g] + min_image(cx, prev[2 * g], W);
      cum[2 * g + 1] = cum[2 * g + 1] + min_image(cy, prev[2 * g + 1], H);
    }
  }
  prev[2 * g] = cx;
  prev[2 * g + 1] = cy;
}

}  // namespace

hipError_t launch_com(int n_groups, const int* off, const float* pos, int H, int W, int periodic,
                      float* com, hipStream_t s) {
  if (n_groups <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_com, dim3(n_groups), dim3(COM_THREADS), 0, s, pos, off, (float)W, (float)H,
                     periodic, com);
  return hipGetLastError();
}

hipError_t launch_track(const TrackWork& t, const float* pos, int mode, hipStream_t s) {
  if (t.n_groups <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_track, dim3(t.n_groups), dim3(COM_THREADS), 0, s, pos, t.off, t.dims, mode,
                     t.start_step, t.prev, t.cum, t.count);
  return hipGetLastError();
}

}  // namespace mc
