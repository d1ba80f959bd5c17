// Deterministic binning primitives shared by the immersed-boundary and steric
// kernels: a single-workgroup exclusive scan with active-bin compaction, and
// an in-workgroup sort of unique ids (bitonic in LDS, rank-merge in global
// memory beyond the LDS capacity).  Bins are filled with integer atomics (order
// unspecified); sorting each bin's unique node ids restores the reference's
// update order before any floating-point accumulation happens.
#pragma once
#include "mc_common.h"
#include "mc_internal.h"

namespace mc {

// Inclusive scan of v across a workgroup of up to 1024 threads (wave64).
// `wsum` must hold blockDim.x / 64 ints of LDS.  Returns the inclusive prefix;
// *total receives the workgroup sum.
__device__ __forceinline__ int block_inclusive_scan(int v, int* wsum, int* total) {
  const int lane = threadIdx.x & 63;
  const int wid = threadIdx.x >> 6;
  const int nw = (blockDim.x + 63) >> 6;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    int t = __shfl_up(v, d, 64);
    if (lane >= d) v += t;
  }
  if (lane == 63) wsum[wid] = v;
  __syncthreads();
  if (wid == 0) {
    int s = (lane < nw) ? wsum[lane] : 0;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      int t = __shfl_up(s, d, 64);
      if (lane >= d) s += t;
    }
    if (lane < nw) wsum[lane] = s;
  }
  __syncthreads();
  int pre = (wid > 0) ? wsum[wid - 1] : 0;
  *total = wsum[nw - 1];
  __syncthreads();
  return v + pre;
}

// Inclusive scans of two values at once across a workgroup of up to 1024
// threads; `wsum` holds 32 ints of LDS.  Returns both prefixes; *total
// receives both workgroup sums.
__device__ __forceinline__ int2 block_inclusive_scan2(int2 v, int* wsum, int2* total) {
  const int lane = threadIdx.x & 63;
  const int wid = threadIdx.x >> 6;
  const int nw = (blockDim.x + 63) >> 6;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int tx = __shfl_up(v.x, d, 64), ty = __shfl_up(v.y, d, 64);
    if (lane >= d) {
      v.x += tx;
      v.y += ty;
    }
  }
  if (lane == 63) {
    wsum[wid] = v.x;
    wsum[16 + wid] = v.y;
  }
  __syncthreads();
  if (wid == 0) {
    int sx = (lane < nw) ? wsum[lane] : 0, sy = (lane < nw) ? wsum[16 + lane] : 0;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const int tx = __shfl_up(sx, d, 64), ty = __shfl_up(sy, d, 64);
      if (lane >= d) {
        sx += tx;
        sy += ty;
      }
    }
    if (lane < nw) {
      wsum[lane] = sx;
      wsum[16 + lane] = sy;
    }
  }
  __syncthreads();
  const int px = (wid > 0) ? wsum[wid - 1] : 0, py = (wid > 0) ? wsum[16 + wid - 1] : 0;
  *total = make_int2(wsum[nw - 1], wsum[16 + nw - 1]);
  __syncthreads();
  return make_int2(v.x + px, v.y + py);
}

// Number of elements of sorted array a[0..n) that are < v.
template <class T>
__device__ __forceinline__ int lower_bound_t(const T* a, int n, T v) {
  int lo = 0, hi = n;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (a[mid] < v) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// Sort m unique keys in place (ascending).  All threads of the workgroup call
// it.  lds: `cap` keys of LDS; gtmp: m keys of global scratch (used beyond cap).
template <class T>
__device__ inline void sort_unique(T* list, int m, T* lds, int cap, T* gtmp, T pad) {
  if (m <= 1) return;
  if (m <= cap) {
    int P = 1;
    while (P < m) P <<= 1;
    for (int i = threadIdx.x; i < P; i += blockDim.x) lds[i] = (i < m) ? list[i] : pad;
    __syncthreads();
    for (int k = 2; k <= P; k <<= 1) {
      for (int j = k >> 1; j > 0; j >>= 1) {
        for (int i = threadIdx.x; i < P; i += blockDim.x) {
          int ixj = i ^ j;
          if (ixj > i) {
            T a = lds[i], b = lds[ixj];
            bool up = ((i & k) == 0);
            if ((a > b) == up) { lds[i] = b; lds[ixj] = a; }
          }
        }
        __syncthreads();
      }
    }
    for (int i = threadIdx.x; i < m; i += blockDim.x) list[i] = lds[i];
    __syncthreads();
    return;
  }
  // bottom-up merge by rank in the partner run (keys are unique)
  T* src = list;
  T* dst = gtmp;
  for (int width = 1; width < m; width <<= 1) {
    for (int i = threadIdx.x; i < m; i += blockDim.x) {
      int start = (i / (2 * width)) * (2 * width);
      int mid = min(start + width, m);
      int end = min(start + 2 * width, m);
      T v = src[i];
      int pos;
      if (i < mid) pos = (i - start) + lower_bound_t(src + mid, end - mid, v);
      else pos = (i - mid) + lower_bound_t(src + start, mid - start, v);
      dst[start + pos] = v;
    }
    __threadfence_block();
    __syncthreads();
    T* t = src; src = dst; dst = t;
  }
  if (src != list) {
    for (int i = threadIdx.x; i < m; i += blockDim.x) list[i] = src[i];
    __threadfence_block();
    __syncthreads();
  }
}

__device__ inline void sort_unique_ids(int* list, int m, int* lds, int* gtmp) {
  sort_unique<int>(list, m, lds, SORT_LDS_CAP, gtmp, 0x7fffffff);
}

}  // namespace mc
