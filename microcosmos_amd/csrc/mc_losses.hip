// Folding losses of the filament-folding experiment and their gradients with
// respect to the particle positions (experiments/filament_folding/
// experiment.py): Chamfer (:57-63), density-aware Chamfer (:105-128) and the
// smooth maximum of the nearest-neighbour distances (:139-149), and the
// stability term over the frames after the shape is scored (:878-893).
//
// Three launches:
//   k_nn         NN_G lanes per particle and per target: the nearest point of
//                the other cloud (squared distance, first index, number of
//                tied minima), each lane over every NN_G-th point of the
//                other cloud staged in LDS, the lanes' results combined by
//                shuffles (a combination that gives the sequential scan's
//                answer: smaller wins, equal minima add their ties and keep
//                the smaller index, the first NaN wins);
//   k_loss       one workgroup: the density counts (DCD), the loss in a fixed
//                order, and per particle the weight of its own term;
//   k_loss_grad  NN_G lanes per particle: 2 w (p - t) over the targets tied at
//                its minimum, plus the targets whose tied minima include it
//                (jax's min rule: a tied minimum's derivative is split
//                equally), in target order.
// Squared distances are dx * dx + dy * dy in float32 (contraction off), the
// same expression in every pass, so the tie tests are exact.
#include <mutex>

#include "mc_common.h"
#include "mc_internal.h"
#include "../../include/microcosmos_hip.h"

namespace mc {
namespace {

constexpr int NN_THREADS = 256;
constexpr int NN_G = 16;                    // lanes per point (a wave holds 4 points)
constexpr int NN_PTS = NN_THREADS / NN_G;   // points per workgroup
constexpr int LOSS_THREADS = 1024;

struct LossWs {
  float *dmin_p, *dmin_t, *cnt_p, *cnt_t, *wown, *scal;
  int *arg_p, *tie_p, *arg_t, *tie_t;
};

__host__ __device__ inline LossWs loss_ws(void* base, long long n, long long m) {
  LossWs w;
  float* f = reinterpret_cast<float*>(base);
  w.dmin_p = f;
  w.cnt_t = w.dmin_p + n;
  w.wown = w.cnt_t + n;
  w.dmin_t = w.wown + n;
  w.cnt_p = w.dmin_t + m;
  w.scal = w.cnt_p + m;
  int* q = reinterpret_cast<int*>(w.scal + 4);
  w.arg_p = q;
  w.tie_p = w.arg_p + n;
  w.arg_t = w.tie_p + n;
  w.tie_t = w.arg_t + m;
  return w;
}

__device__ __forceinline__ float sqdist(float ax, float ay, float bx, float by) {
  const float dx = ax - bx, dy = ay - by;
  return dx * dx + dy * dy;
}

// Nearest-point partial results of two lanes (each over its own subset of
// the other cloud), combined as one sequential scan over both subsets would
// have left them: the first NaN wins (with one tie), a smaller minimum
// wins, equal minima keep the smaller index and add their ties.
__device__ __forceinline__ void nn_combine(float& b, int& a, int& t, float b1, int a1, int t1) {
  const bool n0 = b != b, n1 = b1 != b1;
  if (n0 && n1) {
    a = min(a, a1);
    t = 1;
  } else if (n1 || (!n0 && b1 < b)) {
    b = b1;
    a = a1;
    t = t1;
  } else if (!n0 && b1 == b) {
    a = min(a, a1);
    t = t + t1;
  }
}

// blocks [0, bp): particles against the targets; [bp, grid): targets against
// the particles.  The other cloud is read in LDS tiles of NN_THREADS points;
// lane r of a point's NN_G lanes takes tile entries r, r + NN_G, ...
__global__ void __launch_bounds__(NN_THREADS) k_nn(const float2* __restrict__ p, int n,
                                                   const float2* __restrict__ t, int m, int bp,
                                                   LossWs w) {
  __shared__ float2 tile[NN_THREADS];
  const bool rows = (int)blockIdx.x < bp;
  const int lane = threadIdx.x % NN_G;
  const int i = (rows ? blockIdx.x : blockIdx.x - bp) * NN_PTS + threadIdx.x / NN_G;
  const float2* self = rows ? p : t;
  const float2* other = rows ? t : p;
  const int ns = rows ? n : m, no = rows ? m : n;
  const float2 a = i < ns ? self[i] : make_float2(0.0f, 0.0f);
  float best = __builtin_inff();
  int arg = 0, ties = 0;
  for (int j0 = 0; j0 < no; j0 += NN_THREADS) {
    __syncthreads();
    if (j0 + (int)threadIdx.x < no) tile[threadIdx.x] = other[j0 + threadIdx.x];
    __syncthreads();
    const int cnt = min(NN_THREADS, no - j0);
    for (int k = lane; k < cnt; k += NN_G) {
      const float2 b = tile[k];
      const float d = rows ? sqdist(a.x, a.y, b.x, b.y) : sqdist(b.x, b.y, a.x, a.y);
      if (d < best || (d != d && best == best)) {
        best = d;
        arg = j0 + k;
        ties = 1;
      } else if (d == best) {
        ++ties;
      }
    }
  }
  for (int off = NN_G / 2; off > 0; off >>= 1) {
    const float b1 = __shfl_xor(best, off, NN_G);
    const int a1 = __shfl_xor(arg, off, NN_G);
    const int t1 = __shfl_xor(ties, off, NN_G);
    nn_combine(best, arg, ties, b1, a1, t1);
  }
  if (i < ns && lane == 0) {
    if (rows) {
      w.dmin_p[i] = best;
      w.arg_p[i] = arg;
      w.tie_p[i] = ties;
    } else {
      w.dmin_t[i] = best;
      w.arg_t[i] = arg;
      w.tie_t[i] = ties;
    }
  }
}

// Fixed-order block sum of per-thread partials (stride-ordered, then a tree).
__device__ float block_sum(float v, float* red) {
  red[threadIdx.x] = v;
  __syncthreads();
  for (int s = LOSS_THREADS / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + s];
    __syncthreads();
  }
  const float r = red[0];
  __syncthreads();
  return r;
}

__device__ float block_max(float v, float* red) {
  red[threadIdx.x] = v;
  __syncthreads();
  for (int s = LOSS_THREADS / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = fmax_nan(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  const float r = red[0];
  __syncthreads();
  return r;
}

__global__ void __launch_bounds__(LOSS_THREADS) k_loss(int kind, int n, int m, float alpha,
                                                       LossWs w, float* __restrict__ loss) {
  __shared__ float red[LOSS_THREADS];
  const int tid = threadIdx.x;
  if (kind == MC_LOSS_DCD) {
    // counts: particles per target (argmin rows), targets per particle
    for (int j = tid; j < m; j += LOSS_THREADS) {
      int c = 0;
      for (int i = 0; i < n; ++i) c += w.arg_p[i] == j ? 1 : 0;
      w.cnt_p[j] = (float)c;
    }
    for (int i = tid; i < n; i += LOSS_THREADS) {
      int c = 0;
      for (int j = 0; j < m; ++j) c += w.arg_t[j] == i ? 1 : 0;
      w.cnt_t[i] = (float)c;
    }
    __threadfence_block();
    __syncthreads();
  }
  if (kind == MC_LOSS_SMOOTH_MAX) {
    float zm = -__builtin_inff();
    for (int i = tid; i < n; i += LOSS_THREADS)
      zm = fmax_nan(zm, alpha * sqrtf(w.dmin_p[i] + 1e-8f));
    const float zmax = block_max(zm, red);
    float s = 0.0f;
    for (int i = tid; i < n; i += LOSS_THREADS)
      s = s + expf(alpha * sqrtf(w.dmin_p[i] + 1e-8f) - zmax);
    const float tot = block_sum(s, red);
    for (int i = tid; i < n; i += LOSS_THREADS) {
      const float r = sqrtf(w.dmin_p[i] + 1e-8f);
      w.wown[i] = (expf(alpha * r - zmax) / tot) / (2.0f * r);
    }
    if (tid == 0) *loss = (zmax + logf(tot)) / alpha;
    return;
  }
  const bool dcd = kind == MC_LOSS_DCD;
  float sp = 0.0f, st = 0.0f;
  for (int i = tid; i < n; i += LOSS_THREADS) {
    const float c = dcd ? w.cnt_p[w.arg_p[i]] : 1.0f;
    sp = sp + w.dmin_p[i] / c;
    w.wown[i] = 1.0f / ((float)n * c);
  }
  for (int j = tid; j < m; j += LOSS_THREADS) {
    const float c = dcd ? w.cnt_t[w.arg_t[j]] : 1.0f;
    st = st + w.dmin_t[j] / c;
  }
  const float a = block_sum(sp, red), b = block_sum(st, red);
  if (tid == 0) *loss = a / (float)n + b / (float)m;
}

__global__ void __launch_bounds__(NN_THREADS) k_loss_grad(int kind, const float2* __restrict__ p,
                                                          int n, const float2* __restrict__ t,
                                                          int m, LossWs w,
                                                          float* __restrict__ grad) {
  __shared__ float2 tile[NN_THREADS];
  const int lane = threadIdx.x % NN_G;
  const int i = blockIdx.x * NN_PTS + threadIdx.x / NN_G;
  const float2 a = i < n ? p[i] : make_float2(0.0f, 0.0f);
  const float own = i < n ? w.wown[i] / (float)w.tie_p[i] : 0.0f;
  const float best = i < n ? w.dmin_p[i] : 0.0f;
  const bool targets = kind != MC_LOSS_SMOOTH_MAX;
  const bool dcd = kind == MC_LOSS_DCD;
  float gx = 0.0f, gy = 0.0f;
  for (int j0 = 0; j0 < m; j0 += NN_THREADS) {
    __syncthreads();
    if (j0 + (int)threadIdx.x < m) tile[threadIdx.x] = t[j0 + threadIdx.x];
    __syncthreads();
    const int cnt = min(NN_THREADS, m - j0);
    if (i >= n) continue;
    for (int k = lane; k < cnt; k += NN_G) {
      const int j = j0 + k;
      const float2 b = tile[k];
      const float d = sqdist(a.x, a.y, b.x, b.y);
      float wt = 0.0f;
      if (d == best) wt = own;
      if (targets && d == w.dmin_t[j]) {
        const float c = dcd ? w.cnt_t[w.arg_t[j]] : 1.0f;
        wt = wt + 1.0f / ((float)m * c) / (float)w.tie_t[j];
      }
      if (wt != 0.0f) {
        gx = gx + wt * (a.x - b.x);
        gy = gy + wt * (a.y - b.y);
      }
    }
  }
  // the lanes' partial sums, in a fixed butterfly order
  for (int off = NN_G / 2; off > 0; off >>= 1) {
    gx = gx + __shfl_xor(gx, off, NN_G);
    gy = gy + __shfl_xor(gy, off, NN_G);
  }
  if (i < n && lane == 0) {
    grad[2 * i] = 2.0f * gx;
    grad[2 * i + 1] = 2.0f * gy;
  }
}

// experiment.py:878-893: per frame s, mean(d(x_s, main)^2) (+ the same
// against the anchor frame), d the minimum-image displacement (utils.py:18-23)
// over the (n, 2) array; the loss is the mean over the frames.  Each frame's
// sums are fixed-order block sums, so the result does not depend on timing.
// The gradient of the floor-mod is 1 with respect to its first argument
// (jax's rem rule), so d/dx_s = 2 d / (2 n S) and the main and anchor frames
// take minus the sum over s.
__global__ void __launch_bounds__(LOSS_THREADS) k_stability(const float* __restrict__ frames,
                                                            int nf, int n,
                                                            const float* __restrict__ main_pos,
                                                            const float* __restrict__ anchor_pos,
                                                            float sx, float sy,
                                                            float* __restrict__ loss,
                                                            float* __restrict__ g_frames,
                                                            float* __restrict__ g_main,
                                                            float* __restrict__ g_anchor) {
  __shared__ float red[LOSS_THREADS];
  const int tid = threadIdx.x, ne = 2 * n;
  const float inv = 1.0f / (float)ne, fs = (float)nf;
  float total = 0.0f;
  for (int s = 0; s < nf; ++s) {
    const float* x = frames + (size_t)s * ne;
    float pm = 0.0f, pa = 0.0f;
    for (int e = tid; e < ne; e += LOSS_THREADS) {
      const float L = (e & 1) ? sy : sx;
      const float dm = min_image(x[e], main_pos[e], L);
      pm = pm + dm * dm;
      float ga = 0.0f;
      if (anchor_pos) {
        const float da = min_image(x[e], anchor_pos[e], L);
        pa = pa + da * da;
        ga = 2.0f * da * inv / fs;
      }
      if (g_frames) g_frames[(size_t)s * ne + e] = 2.0f * dm * inv / fs + ga;
    }
    float d = block_sum(pm, red) * inv;
    if (anchor_pos) d = d + block_sum(pa, red) * inv;
    total = total + d;
  }
  if (tid == 0) *loss = total / fs;
  if (!g_frames) return;
  for (int e = tid; e < ne; e += LOSS_THREADS) {
    const float L = (e & 1) ? sy : sx;
    float gm = 0.0f, ga = 0.0f;
    for (int s = 0; s < nf; ++s) {
      const float xv = frames[(size_t)s * ne + e];
      gm = gm + 2.0f * min_image(xv, main_pos[e], L) * inv / fs;
      if (anchor_pos) ga = ga + 2.0f * min_image(xv, anchor_pos[e], L) * inv / fs;
    }
    g_main[e] = -gm;
    if (anchor_pos) g_anchor[e] = -ga;
  }
}

}  // namespace

hipError_t launch_stability(const float* frames, int nf, int n, const float* main_pos,
                            const float* anchor_pos, int height, int width, float* loss,
                            float* g_frames, float* g_main, float* g_anchor, hipStream_t s) {
  hipLaunchKernelGGL(k_stability, dim3(1), dim3(LOSS_THREADS), 0, s, frames, nf, n, main_pos,
                     anchor_pos, (float)width, (float)height, loss, g_frames, g_main, g_anchor);
  return hipGetLastError();
}

size_t point_loss_workspace(long long n, long long m) {
  return (size_t)(5 * n + 4 * m + 4) * sizeof(float);
}

// The workspace of mc_point_loss_static: one array per device, large enough
// for MC_POINT_LOSS_STATIC_MAX particles and as many targets (2.25 MiB).  The
// kernels write every word they read, so nothing needs clearing between
// calls.  Host code uses only the device address that hipGetSymbolAddress
// returns for it (the name itself is a host-side stand-in).
constexpr long long SCRATCH_POINTS = MC_POINT_LOSS_STATIC_MAX;
static __device__ __align__(256) float scratch[5 * SCRATCH_POINTS + 4 * SCRATCH_POINTS + 4];

hipError_t point_loss_scratch(void** out) {
  constexpr int MAX_DEVICES = 64;
  static std::mutex mu;
  static void* addr[MAX_DEVICES] = {};
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= MAX_DEVICES) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[dev]) {
    // once per device, with that device current: each device holds its own
    // copy of the array at its own address
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    addr[dev] = p;
  }
  *out = addr[dev];
  return hipSuccess;
}

hipError_t launch_point_loss(int kind, const float* particles, int n, const float* targets, int m,
                             float alpha, float* loss, float* grad, void* workspace,
                             hipStream_t s) {
  const LossWs w = loss_ws(workspace, n, m);
  const int bp = (n + NN_PTS - 1) / NN_PTS, bt = (m + NN_PTS - 1) / NN_PTS;
  const float2* p = reinterpret_cast<const float2*>(particles);
  const float2* t = reinterpret_cast<const float2*>(targets);
  hipLaunchKernelGGL(k_nn, dim3(bp + bt), dim3(NN_THREADS), 0, s, p, n, t, m, bp, w);
  hipLaunchKernelGGL(k_loss, dim3(1), dim3(LOSS_THREADS), 0, s, kind, n, m, alpha, w, loss);
  if (grad)
    hipLaunchKernelGGL(k_loss_grad, dim3(bp), dim3(NN_THREADS), 0, s, kind, p, n, t, m, w, grad);
  return hipGetLastError();
}

}  // namespace mc
