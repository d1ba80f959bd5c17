// Sparse immersed-boundary transfer (fluid.py:202-384) with deterministic,
// atomic-free floating-point accumulation.
//
// The reference scatters node-major, tap-minor (fluid.py:111-118,130-132).  A
// node's K x K stencil touches each cell at most once (grid >= K per axis), so
// the value a cell accumulates is a left fold over the covering nodes in
// ascending node order.  The device reproduces that fold exactly:
//   1. stencil: per node, wrapped start cell and normalised K*K weights
//      (fluid.py:40-87); count node-tile incidences (integer atomics);
//   2. scan: tile offsets and the list of active tiles;
//   3. fill: node ids into their tiles' bins (order unspecified);
//   4. tile init: sort each bin ascending, then every cell thread folds the
//      kernel weights of its covering nodes in that order; covered cells (the
//      support U) get rho, u, u_corrected = u, 1/max(kw, 1) and a U bit;
//   5. per IB iteration: node gather (interpolation + clamped correction),
//      then the per-cell ordered fold of the spread (fluid.py:292-316);
//      the last cell pass also post-clamps and forms F = (u_c - u) rho;
//   6. drag: node gather of F (fluid.py:364-382) and the final position wrap
//      (simulate.py:97).
// Only U (about 6 % of cells at 16k nodes on 1024^2) is ever read or written.
#include "mc_bin.h"
#include "mc_common.h"
#include "mc_internal.h"

namespace mc {
namespace {

constexpr int CH = 128;  // nodes staged in LDS per chunk

__device__ __forceinline__ float peskin(float r, float h) {
  float ra = fabsf(r) / h;
  return (ra <= 2.0f) ? 0.25f * (1.0f + dcos((MC_PI_F * ra) / 2.0f)) : 0.0f;
}

// Distinct tile indices covered along one axis by cells s..s+K-1 (mod n).
__device__ __forceinline__ int axis_tiles(int s, int K, int n, int* out) {
  int cnt = 0;
  for (int d = 0; d < K; ++d) {
    int t = wrap1(s + d, n) / IB_TILE;
    bool seen = false;
    for (int q = 0; q < cnt; ++q) seen = seen || (out[q] == t);
    if (!seen) out[cnt++] = t;
  }
  return cnt;
}

template <int K>
__global__ void k_ib_stencil(Consts c, int N, const float* __restrict__ pos,
                             const float* __restrict__ vel, int* __restrict__ st_sx,
                             int* __restrict__ st_sy, float* __restrict__ st_w,
                             float* __restrict__ vib, int* __restrict__ tile_cnt, int tiles_x) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  constexpr int K2 = K * K;
  const float gx = pos[2 * n] * c.scale_x;
  const float gy = pos[2 * n + 1] * c.scale_y;
  vib[2 * n] = (vel[2 * n] * c.scale_x) * c.dt;
  vib[2 * n + 1] = (vel[2 * n + 1] * c.scale_y) * c.dt;
  const float sxf = floorf(gx) - (float)c.half;
  const float syf = floorf(gy) - (float)c.half;
  // 1-D weights: tap t uses wx[t / K] and wy[t % K] (fluid.py:53-69)
  float wx[K], wy[K];
#pragma unroll
  for (int d = 0; d < K; ++d) {
    wx[d] = peskin((sxf + (float)d) - gx, c.half_width);
    wy[d] = peskin((syf + (float)d) - gy, c.half_width);
  }
  float total = 0.0f;
#pragma unroll
  for (int t = 0; t < K2; ++t) total = total + wx[t / K] * wy[t % K];
  const float den = total + MC_EPSIL;
#pragma unroll
  for (int t = 0; t < K2; ++t) st_w[(size_t)n * K2 + t] = (wx[t / K] * wy[t % K]) / den;
  const int sx = imod(f2i(sxf), c.fl_w);
  const int sy = imod(f2i(syf), c.fl_h);
  st_sx[n] = sx;
  st_sy[n] = sy;
  int tx[K], ty[K];
  const int ntx = axis_tiles(sx, K, c.fl_w, tx);
  const int nty = axis_tiles(sy, K, c.fl_h, ty);
  for (int a = 0; a < nty; ++a)
    for (int b = 0; b < ntx; ++b) atomicAdd(&tile_cnt[ty[a] * tiles_x + tx[b]], 1);
}

__global__ void k_ib_fill(Consts c, int N, const int* __restrict__ st_sx,
                          const int* __restrict__ st_sy, const int* __restrict__ tile_off,
                          int* __restrict__ tile_fill, int* __restrict__ tile_list, int tiles_x) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  int tx[IB_KMAX], ty[IB_KMAX];
  const int ntx = axis_tiles(st_sx[n], c.K, c.fl_w, tx);
  const int nty = axis_tiles(st_sy[n], c.K, c.fl_h, ty);
  for (int a = 0; a < nty; ++a)
    for (int b = 0; b < ntx; ++b) {
      const int t = ty[a] * tiles_x + tx[b];
      const int p = atomicAdd(&tile_fill[t], 1);
      tile_list[tile_off[t] + p] = n;
    }
}

struct TileStage {
  int id[CH];
  int sx[CH];
  int sy[CH];
  float cx[CH];
  float cy[CH];
  float w[CH * IB_TAPS_MAX];
};

// Stage nodes list[j0 .. j0+cnt) (cnt <= CH <= blockDim) into LDS: one node
// per thread, all of its global loads issued together.
__device__ __forceinline__ void stage_nodes(TileStage& S, const int* list, int j0, int cnt, int K2,
                                            const int* st_sx, const int* st_sy, const float* st_w,
                                            const float* corr) {
  const int j = threadIdx.x;
  if (j < cnt) {
    const int n = list[j0 + j];
    S.id[j] = n;
    S.sx[j] = st_sx[n];
    S.sy[j] = st_sy[n];
    if (corr) {
      S.cx[j] = corr[2 * n];
      S.cy[j] = corr[2 * n + 1];
    }
    const float* wn = st_w + (size_t)n * K2;
    float wv[IB_TAPS_MAX];
#pragma unroll
    for (int t = 0; t < IB_TAPS_MAX; ++t)
      if (t < K2) wv[t] = wn[t];
#pragma unroll
    for (int t = 0; t < IB_TAPS_MAX; ++t)
      if (t < K2) S.w[j * K2 + t] = wv[t];
  }
  __syncthreads();
}

// Tap index of cell (cx, cy) in the stencil starting at (sx, sy), or -1.
__device__ __forceinline__ int tap_of(int cx, int cy, int sx, int sy, int K, int W, int H) {
  const int dx = wrap1(cx - sx, W);
  const int dy = wrap1(cy - sy, H);
  return (dx < K && dy < K) ? dx * K + dy : -1;
}

__global__ void __launch_bounds__(256) k_ib_tile_init(
    Consts c, const int* __restrict__ tile_off, const int* __restrict__ active,
    const int* __restrict__ n_active, int* __restrict__ tile_list, int* __restrict__ sort_tmp,
    const int* __restrict__ st_sx, const int* __restrict__ st_sy, const float* __restrict__ st_w,
    const float* __restrict__ f_in, float* __restrict__ rho_s, float* __restrict__ u0,
    float* __restrict__ ucorr, float* __restrict__ inv, uint32_t* __restrict__ mask, int mwpr,
    int tiles_x, int2* __restrict__ cell_span, int2* __restrict__ entries) {
  __shared__ TileStage S;
  __shared__ int sort_lds[SORT_LDS_CAP];
  __shared__ int wsum[16];
  const int W = c.fl_w, H = c.fl_h, K = c.K, K2 = K * K;
  const size_t HW = (size_t)H * W;
  const int na = *n_active;
  for (int b = blockIdx.x; b < na; b += gridDim.x) {
    const int t = active[b];
    const int off = tile_off[t];
    const int m = tile_off[t + 1] - off;
    sort_unique_ids(tile_list + off, m, sort_lds, sort_tmp + off);
    const int cx = (t % tiles_x) * IB_TILE + (threadIdx.x & (IB_TILE - 1));
    const int cy = (t / tiles_x) * IB_TILE + (threadIdx.x / IB_TILE);
    const bool valid = cx < W && cy < H;
    // pass 1: per cell, fold of the covering nodes' weights and their count
    float kw = 0.0f;
    int cnt = 0;
    for (int j0 = 0; j0 < m; j0 += CH) {
      const int nc = min(CH, m - j0);
      stage_nodes(S, tile_list + off, j0, nc, K2, st_sx, st_sy, st_w, nullptr);
      if (valid) {
        for (int j = 0; j < nc; ++j) {
          const int tap = tap_of(cx, cy, S.sx[j], S.sy[j], K, W, H);
          if (tap >= 0) {
            ++cnt;
            kw = kw + S.w[j * K2 + tap];
          }
        }
      }
      __syncthreads();
    }
    // the tile's entry region holds at most K*K entries per node-tile incidence
    int tot;
    const int incl = block_inclusive_scan(cnt, wsum, &tot);
    const int start = K2 * off + (incl - cnt);
    // pass 2: the ordered (node, weight) list of every covered cell
    if (m > CH) {
      int e = start;
      for (int j0 = 0; j0 < m; j0 += CH) {
        const int nc = min(CH, m - j0);
        stage_nodes(S, tile_list + off, j0, nc, K2, st_sx, st_sy, st_w, nullptr);
        if (valid && cnt > 0) {
          for (int j = 0; j < nc; ++j) {
            const int tap = tap_of(cx, cy, S.sx[j], S.sy[j], K, W, H);
            if (tap >= 0) entries[e++] = make_int2(S.id[j], __float_as_int(S.w[j * K2 + tap]));
          }
        }
        __syncthreads();
      }
    } else if (valid && cnt > 0) {
      // the single chunk is still staged
      int e = start;
      for (int j = 0; j < m; ++j) {
        const int tap = tap_of(cx, cy, S.sx[j], S.sy[j], K, W, H);
        if (tap >= 0) entries[e++] = make_int2(S.id[j], __float_as_int(S.w[j * K2 + tap]));
      }
    }
    if (valid && cnt > 0) {
      const size_t cell = (size_t)cy * W + cx;
      cell_span[cell] = make_int2(start, cnt);
      float f[9];
#pragma unroll
      for (int i = 0; i < 9; ++i) f[i] = fmax_nan(f_in[i * HW + cell], MC_EPSIL);
      float rho, ux, uy;
      lbm_macro(f, c.umax, rho, ux, uy);
      rho_s[cell] = rho;
      u0[cell] = ux;
      u0[HW + cell] = uy;
      ucorr[cell] = ux;
      ucorr[HW + cell] = uy;
      inv[cell] = 1.0f / fmax_nan(kw, 1.0f);
      atomicOr(&mask[(size_t)cy * mwpr + (cx >> 5)], 1u << (cx & 31));
    }
    __syncthreads();
  }
}

// Node phase: gather of a (2, H, W) field over the stencil, left fold over taps.
template <int K>
__device__ __forceinline__ void gather2(const Consts& c, int n, const int* st_sx, const int* st_sy,
                                        const float* st_w, const float* fld, float& ax, float& ay) {
  constexpr int K2 = K * K;
  const int W = c.fl_w, H = c.fl_h;
  const size_t HW = (size_t)H * W;
  const int sx = st_sx[n], sy = st_sy[n];
  int row[K], col[K];
#pragma unroll
  for (int d = 0; d < K; ++d) {
    col[d] = wrap1(sx + d, W);
    row[d] = wrap1(sy + d, H) * W;
  }
  ax = 0.0f;
  ay = 0.0f;
#pragma unroll
  for (int t = 0; t < K2; ++t) {
    const size_t cell = (size_t)row[t % K] + col[t / K];
    const float w = st_w[(size_t)n * K2 + t];
    ax = ax + fld[cell] * w;
    ay = ay + fld[HW + cell] * w;
  }
}

template <int K>
__global__ void k_ib_interp(Consts c, int N, const int* __restrict__ st_sx,
                            const int* __restrict__ st_sy, const float* __restrict__ st_w,
                            const float* __restrict__ vib, const float* __restrict__ ucorr,
                            float* __restrict__ corr) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  float ux, uy;
  gather2<K>(c, n, st_sx, st_sy, st_w, ucorr, ux, uy);
  float cx = (vib[2 * n] - ux) * c.relax;
  float cy = (vib[2 * n + 1] - uy) * c.relax;
  const float mag = sqrtf((cx * cx + cy * cy) + MC_EPSIL);
  const float cf = fmin_nan(1.0f, c.umax / (mag + MC_EPSIL));
  corr[2 * n] = cx * cf;
  corr[2 * n + 1] = cy * cf;
}

// Cell phase: every covered cell folds c_n * w over its ordered entry list
// (fluid.py:109-120), scales by 1/max(kw, 1) and adds to u_corrected
// (fluid.py:307,310); the last pass also post-clamps and forms the force.
__global__ void __launch_bounds__(256) k_ib_spread(
    Consts c, const int* __restrict__ active, const int* __restrict__ n_active,
    const int2* __restrict__ cell_span, const int2* __restrict__ entries,
    const float* __restrict__ corr, const float* __restrict__ rho_s, const float* __restrict__ u0,
    float* __restrict__ ucorr, const float* __restrict__ inv, const uint32_t* __restrict__ mask,
    int mwpr, float* __restrict__ force, int do_spread, int do_final, int tiles_x) {
  const int W = c.fl_w, H = c.fl_h;
  const size_t HW = (size_t)H * W;
  const int na = *n_active;
  for (int b = blockIdx.x; b < na; b += gridDim.x) {
    const int t = active[b];
    const int cx = (t % tiles_x) * IB_TILE + (threadIdx.x & (IB_TILE - 1));
    const int cy = (t / tiles_x) * IB_TILE + (threadIdx.x / IB_TILE);
    if (cx >= W || cy >= H) continue;
    if (!((mask[(size_t)cy * mwpr + (cx >> 5)] >> (cx & 31)) & 1u)) continue;
    const size_t cell = (size_t)cy * W + cx;
    float ucx = ucorr[cell], ucy = ucorr[HW + cell];
    if (do_spread) {
      const int2 sp = cell_span[cell];
      float ax = 0.0f, ay = 0.0f;
      for (int e = sp.x; e < sp.x + sp.y; ++e) {
        const int2 en = entries[e];
        const float w = __int_as_float(en.y);
        ax = ax + corr[2 * en.x] * w;
        ay = ay + corr[2 * en.x + 1] * w;
      }
      const float iv = inv[cell];
      ucx = ucx + ax * iv;
      ucy = ucy + ay * iv;
      ucorr[cell] = ucx;
      ucorr[HW + cell] = ucy;
    }
    if (do_final) {
      const float mag = sqrtf((ucx * ucx + ucy * ucy) + MC_EPSIL);
      const float cl = fmin_nan(1.0f, c.umax / (mag + MC_EPSIL));
      ucx = ucx * cl;
      ucy = ucy * cl;
      const float rho = rho_s[cell];
      float fx = (ucx - u0[cell]) * rho;
      float fy = (ucy - u0[HW + cell]) * rho;
      if (c.force_clamp) {
        const float fm = sqrtf((fx * fx + fy * fy) + MC_EPSIL);
        const float fc = fmin_nan(1.0f, c.umax / (fm + MC_EPSIL));
        fx = fx * fc;
        fy = fy * fc;
      }
      force[cell] = fx;
      force[HW + cell] = fy;
    }
  }
}

template <int K>
__global__ void k_ib_drag(Consts c, int N, const int* __restrict__ st_sx,
                          const int* __restrict__ st_sy, const float* __restrict__ st_w,
                          const float* __restrict__ force, const float* __restrict__ pos,
                          float* __restrict__ vel_io, float* __restrict__ pos_out,
                          long long* __restrict__ tick) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  // advance the actuation step counter (the rod kernel of this step, which
  // read it, has completed before this kernel starts)
  if (tick && n == 0) *tick = *tick + 1;
  if (n >= N) return;
  float fx, fy;
  gather2<K>(c, n, st_sx, st_sy, st_w, force, fx, fy);
  const float dvx = (fx / c.node_mass) / c.sdt_x;
  const float dvy = (fy / c.node_mass) / c.sdt_y;
  vel_io[2 * n] = vel_io[2 * n] - dvx;
  vel_io[2 * n + 1] = vel_io[2 * n + 1] - dvy;
  const float px = pos[2 * n], py = pos[2 * n + 1];
  pos_out[2 * n] = fmodf_floor(px, c.dom_wf);
  pos_out[2 * n + 1] = fmodf_floor(py, c.dom_hf);
}

__global__ void __launch_bounds__(1024) k_bin_scan(int T, int* __restrict__ cnt,
                                                   int* __restrict__ off, int* __restrict__ fill,
                                                   int* __restrict__ active,
                                                   int* __restrict__ n_active) {
  __shared__ int wsum[16];
  __shared__ int wsum2[16];
  int carry = 0, acarry = 0;
  for (int base = 0; base < T; base += blockDim.x) {
    const int i = base + threadIdx.x;
    const int v = (i < T) ? cnt[i] : 0;
    const int a = v > 0 ? 1 : 0;
    int tot, atot;
    const int inc = block_inclusive_scan(v, wsum, &tot);
    const int ainc = block_inclusive_scan(a, wsum2, &atot);
    if (i < T) {
      off[i] = carry + inc - v;
      fill[i] = 0;
      cnt[i] = 0;
      if (a) active[acarry + ainc - 1] = i;
    }
    carry += tot;
    acarry += atot;
  }
  if (threadIdx.x == 0) {
    off[T] = carry;
    *n_active = acarry;
  }
}

}  // namespace

hipError_t launch_bin_scan(int T, const int* cnt, int* off, int* fill, int* active, int* n_active,
                           hipStream_t s) {
  hipLaunchKernelGGL(k_bin_scan, dim3(1), dim3(1024), 0, s, T, const_cast<int*>(cnt), off, fill,
                     active, n_active);
  return hipGetLastError();
}

#define MC_KSWITCH(LAUNCH)                          \
  switch (c.K) {                                    \
    case 1: LAUNCH(1); break;                       \
    case 2: LAUNCH(2); break;                       \
    case 3: LAUNCH(3); break;                       \
    case 4: LAUNCH(4); break;                       \
    case 5: LAUNCH(5); break;                       \
    default: LAUNCH(6); break;                      \
  }

hipError_t launch_ib_bin(const Consts& c, IBWork& w, const float* pos, float* vel_io,
                         hipStream_t s) {
  const int N = w.N;
  if (N == 0) return hipSuccess;
  const int nb = (N + 255) / 256;
#define L_STENCIL(KK)                                                                            \
  hipLaunchKernelGGL(k_ib_stencil<KK>, dim3(nb), dim3(256), 0, s, c, N, pos, vel_io, w.st_sx,    \
                     w.st_sy, w.st_w, w.vib, w.tile_cnt, w.tiles_x)
  MC_KSWITCH(L_STENCIL);
#undef L_STENCIL
  hipLaunchKernelGGL(k_bin_scan, dim3(1), dim3(1024), 0, s, w.T, w.tile_cnt, w.tile_off,
                     w.tile_fill, w.tile_active, w.n_active);
  hipLaunchKernelGGL(k_ib_fill, dim3(nb), dim3(256), 0, s, c, N, w.st_sx, w.st_sy, w.tile_off,
                     w.tile_fill, w.tile_list, w.tiles_x);
  return hipGetLastError();
}

hipError_t launch_ib_solve(const Consts& c, IBWork& w, const float* f_in, hipStream_t s) {
  const int N = w.N;
  // no nodes: no support; the LBM pass takes the clamp-only force everywhere
  if (N == 0) return hipSuccess;
  const int nb = (N + 255) / 256;
  const int tg = w.T;  // one workgroup per tile: inactive ones exit at once
  hipLaunchKernelGGL(k_ib_tile_init, dim3(tg), dim3(256), 0, s, c, w.tile_off, w.tile_active,
                     w.n_active, w.tile_list, w.sort_scratch, w.st_sx, w.st_sy, w.st_w, f_in, w.rho,
                     w.u0, w.ucorr, w.inv, w.mask, w.mask_words_per_row, w.tiles_x, w.cell_span,
                     w.entries);
  if (c.iters == 0) {
    hipLaunchKernelGGL(k_ib_spread, dim3(tg), dim3(256), 0, s, c, w.tile_active, w.n_active,
                       w.cell_span, w.entries, w.corr, w.rho, w.u0, w.ucorr, w.inv, w.mask,
                       w.mask_words_per_row, w.force, 0, 1, w.tiles_x);
  }
#define L_INTERP(KK)                                                                             \
  hipLaunchKernelGGL(k_ib_interp<KK>, dim3(nb), dim3(256), 0, s, c, N, w.st_sx, w.st_sy, w.st_w, \
                     w.vib, w.ucorr, w.corr)
  for (int it = 0; it < c.iters; ++it) {
    MC_KSWITCH(L_INTERP);
    hipLaunchKernelGGL(k_ib_spread, dim3(tg), dim3(256), 0, s, c, w.tile_active, w.n_active,
                       w.cell_span, w.entries, w.corr, w.rho, w.u0, w.ucorr, w.inv, w.mask,
                       w.mask_words_per_row, w.force, 1, it == c.iters - 1 ? 1 : 0, w.tiles_x);
  }
#undef L_INTERP
  return hipGetLastError();
}

hipError_t launch_ib_drag(const Consts& c, IBWork& w, const float* pos, float* vel_io,
                          float* pos_out, long long* tick, hipStream_t s) {
  const int N = w.N;
  if (N == 0) {
    if (tick) return launch_tick(tick, s);
    return hipSuccess;
  }
  const int nb = (N + 255) / 256;
#define L_DRAG(KK)                                                                               \
  hipLaunchKernelGGL(k_ib_drag<KK>, dim3(nb), dim3(256), 0, s, c, N, w.st_sx, w.st_sy, w.st_w,   \
                     w.force, pos, vel_io, pos_out, tick)
  MC_KSWITCH(L_DRAG);
#undef L_DRAG
  return hipGetLastError();
}

}  // namespace mc
