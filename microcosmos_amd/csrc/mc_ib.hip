// Sparse immersed-boundary transfer (fluid.py:202-384), solved per interaction
// cluster with deterministic, atomic-free floating-point accumulation.
//
// Two nodes interact through the IB only when their K x K stencils share a
// lattice cell.  Filaments (connected components of the edge graph) whose
// stencils share cells are merged into clusters; no cell is touched by two
// clusters, so one workgroup can run a whole cluster's IB -- kernel-weight
// fold, every iteration, post-clamp, force and node drag -- with workgroup
// barriers only.  Launches per step:
//   1. claim (k_ib_claim): per node, wrapped start cell, normalised K*K
//      weights (fluid.py:40-87) and v_ib; per tap, a compare-and-swap that
//      makes the first arriving component the cell's owner.  The node records
//      which of its taps won; a tap that finds another component's claim
//      emits a link into its instance's link bucket;
//   2. solve (k_ib_solve, one workgroup per component): the workgroup of a
//      cluster's smallest component unions the instance's links in LDS, lists
//      the cluster's members, ranks its nodes by id, builds its covered cells
//      ("segments", one per won tap) and each cell's entry list in ascending
//      node id, then runs the reference's arithmetic; the other members'
//      workgroups return.  Instances with more components than the LDS
//      union-find holds take a one-workgroup cluster build
//      (k_ib_cluster_build) between the two launches instead;
//   3. wide solve (k_ib_wide): clusters above the per-workgroup size, all of
//      them together, with grid barriers (returns at once when there is none).
//
// The reference scatters node-major, tap-minor (fluid.py:111-118,130-132).  A
// node's stencil covers each cell at most once (lattice >= K per axis), so a
// cell's sum is a left fold over its covering nodes in ascending node id; a
// node's gather is a left fold over its taps in tap order (fluid.py:90-97).
// Both folds are reproduced exactly.  Segment numbering comes from LDS
// atomics and is arbitrary; no floating-point result depends on it.
//
// Idle state between calls: owner == -1 on every lattice cell and no pending
// links.  The solves reset the owners of the cells they covered and clear
// the link counts.
#include <algorithm>

#include "mc_bin.h"
#include "mc_claim.h"
#include "mc_common.h"
#include "mc_internal.h"
#include "mc_predict.h"

namespace mc {
namespace {

// Relaxed load at device scope: reads a value produced by atomics at L2
// rather than a possibly stale L1 line.
__device__ __forceinline__ int ld_l2(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Union-find on a generic pointer (LDS or global memory).
__device__ __forceinline__ int uf_find(int* p, int x) {
  while (true) {
    const int q = __hip_atomic_load(&p[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (q == x) return x;
    x = q;
  }
}

// Link the larger root under the smaller one, so a tree's root is its
// smallest component id; a failed CAS means the root moved: retry.
__device__ __forceinline__ void uf_union(int* p, int a, int b) {
  while (true) {
    a = uf_find(p, a);
    b = uf_find(p, b);
    if (a == b) return;
    if (a < b) {
      const int t = a;
      a = b;
      b = t;
    }
    if (atomicCAS(&p[a], a, b) == a) return;
  }
}

// The lattice position of synthetic IB point `nd` +/- (fluid.py:246-263):
// the node's incident edge vectors (tgt - src, minimum image) summed from
// zero in the reference's scatter order (every edge at its source in edge
// order, then at its target in edge order), normalised with the EPSIL under
// the square root, turned by +90 degrees, and scaled by the width and the
// mean lattice scale; (gx, gy) holds the node's lattice position on entry.
__device__ __forceinline__ void synthetic_point(const Consts& c, const InstGeo& gi,
                                                const float* __restrict__ pos,
                                                const int* __restrict__ inc_off,
                                                const int2* __restrict__ inc, int nd, bool minus,
                                                float& gx, float& gy) {
  float tx = 0.0f, ty = 0.0f;
  for (int i = inc_off[nd], e = inc_off[nd + 1]; i < e; ++i) {
    const int2 st = inc[i];
    const float2 a = *reinterpret_cast<const float2*>(pos + 2 * st.x);
    const float2 b = *reinterpret_cast<const float2*>(pos + 2 * st.y);
    tx = tx + min_image(b.x, a.x, gi.dom_wf);
    ty = ty + min_image(b.y, a.y, gi.dom_hf);
  }
  const float nrm = sqrtf((tx * tx + ty * ty) + MC_EPSIL);
  const float ux = tx / nrm, uy = ty / nrm;
  const float ox = (c.syn_d * -uy) * gi.mean_scale;
  const float oy = (c.syn_d * ux) * gi.mean_scale;
  gx = minus ? gx - ox : gx + ox;
  gy = minus ? gy - oy : gy + oy;
}

// One lane per (point, tap).  Every lane of a node group evaluates the node's
// geometry; lane t owns tap t (dx = t / K, dy = t % K): its weight, the claim
// of its cell (compare-and-swap: the first arriving component owns it; the
// node records which of its taps won) and a link when another component holds
// the cell.  Links go to the node's instance bucket (bucketed) or to one list.
// Block 0 also clears the wide solve's counters and the solve's scratch
// allocator (the previous step's launches have completed).
// The N points are the nodes, or with inc_off the 2 x nodes synthetic points
// (bd.node_inst and comp_of_node then per point; pos and vel per node).
template <int K>
__global__ void __launch_bounds__(256) k_ib_claim(
    Consts c, BatchDev bd, int N, const float* __restrict__ pos, const float* __restrict__ vel,
    const int* __restrict__ comp_of_node, int* __restrict__ st_sx, int* __restrict__ st_sy,
    float* __restrict__ st_w, unsigned long long* __restrict__ st_won, float* __restrict__ vib,
    int* __restrict__ owner, int2* __restrict__ links, int* __restrict__ nlinks,
    const int* __restrict__ link_off, int bucketed, WideCtl* __restrict__ wctl,
    IbCtl* __restrict__ ictl, int nodes, const int* __restrict__ inc_off,
    const int2* __restrict__ inc) {
  constexpr int K2 = K * K;
  constexpr int L = stencil_lanes<K>();
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = gid / L, t = gid - n * L;
  const int lane = threadIdx.x & 63, base = lane - t;  // the group's first lane
  if (gid == 0) {
    // the previous solve launch has completed: its scratch is free
    if (ictl) __hip_atomic_store(&ictl->scratch_top, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (wctl) {
      // atomic stores: the counters are later updated by device-scope atomics
      __hip_atomic_store(&wctl->bar, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&wctl->wide, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&wctl->wide_nodes, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&wctl->nseg, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&wctl->nq, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
  const bool node_ok = n < N;
  const bool tap = node_ok && t < K2;
  float prod = 0.0f;
  int prev = -1, comp = -1, inst = 0;
  int sx = 0, sy = 0, W = 1, H = 1, g = -1;
  const int dx = t / K, dy = t - dx * K;
  float gx = 0.0f, gy = 0.0f, sxf = 0.0f, syf = 0.0f;
  const InstGeo* gi = nullptr;
  if (node_ok) {
    inst = bd.node_inst[n];
    gi = &bd.geo[inst];
    W = gi->Wf;
    H = gi->Hf;
    const bool minus = inc_off && n >= nodes;
    const int nd = minus ? n - nodes : n;  // the point's node
    gx = pos[2 * nd] * gi->scale_x;
    gy = pos[2 * nd + 1] * gi->scale_y;
    if (inc_off) synthetic_point(c, *gi, pos, inc_off, inc, nd, minus, gx, gy);
    sxf = floorf(gx) - (float)c.half;
    syf = floorf(gy) - (float)c.half;
    sx = imod(f2i(sxf), W);
    sy = imod(f2i(syf), H);
    comp = comp_of_node[n];
    if (tap) g = wrap1(sy + dy, H) * W + wrap1(sx + dx, W);
    if (t == 0) {
      st_sx[n] = sx;
      st_sy[n] = sy;
      vib[2 * n] = (vel[2 * nd] * gi->scale_x) * c.dt;
      vib[2 * n + 1] = (vel[2 * nd + 1] * gi->scale_y) * c.dt;
    }
  }
  // Neighbouring nodes of a filament cover many of the same cells.  A tap
  // whose cell is also covered by a tap of the same component among the three
  // node groups before it in the workgroup leaves the claim to that tap
  // (which claims it, or leaves it to an earlier one in turn): the cell's
  // owner and links are the same, with fewer contended atomics.  The match is
  // checked on the cell index itself.
  constexpr int NPB = 256 / L;  // node groups per workgroup
  __shared__ int4 grp_geo[NPB];
  const int grp = threadIdx.x / L;
  if (t == 0) grp_geo[grp] = make_int4(sx, sy, comp, 0);
  __syncthreads();
  bool dup = false;
#pragma unroll
  for (int m = 1; m <= 3 && m < NPB; ++m) {
    const int k = grp - m;
    if (k < 0 || !tap) break;
    const int4 q = grp_geo[k];
    int ddx = sx - q.x, ddy = sy - q.y;  // the group's offset, wrapped to the nearer image
    ddx = ddx > W / 2 ? ddx - W : (ddx < -(W / 2) ? ddx + W : ddx);
    ddy = ddy > H / 2 ? ddy - H : (ddy < -(H / 2) ? ddy + H : ddy);
    const int dxk = dx + ddx, dyk = dy + ddy;
    if (q.z == comp && dxk >= 0 && dxk < K && dyk >= 0 && dyk < K &&
        wrap1(q.y + dyk, H) * W + wrap1(q.x + dxk, W) == g)
      dup = true;
  }
  if (tap) {
    prev = dup ? -2 : atomicCAS(&owner[gi->fl_off + g], -1, comp);
    // 1-D weights: tap t uses wx[t / K] and wy[t % K] (fluid.py:53-69)
    const float wxv = peskin((sxf + (float)dx) - gx, c.half_width);
    const float wyv = peskin((syf + (float)dy) - gy, c.half_width);
    prod = wxv * wyv;
  }
  // the normalising sum over taps 0..K2-1 in order (every lane of the group)
  float total = 0.0f;
#pragma unroll
  for (int k = 0; k < K2; ++k) total = total + __shfl(prod, base + k, 64);
  const unsigned long long won_lanes = __ballot(tap && prev == -1);
  const int prev_left = __shfl(prev, t > 0 ? lane - 1 : lane, 64);
  // a cell claimed by another component links the two (a repeat of the left
  // neighbour tap's claimant is dropped; at most one link per tap; a tap that
  // left the claim to another (prev -2) neither wins nor links)
  const bool emit = tap && prev >= 0 && prev != comp && (t == 0 || prev != prev_left);
  // one counter atomic per wavefront when its nodes share a bucket
  const int key = bucketed ? inst : 0;
  const int k0 = __shfl(key, 0, 64);
  int li;
  if (__all(!emit || key == k0)) li = wave_fetch_add(&nlinks[k0], emit ? 1 : 0);
  else li = emit ? atomicAdd(&nlinks[key], 1) : 0;
  if (!tap) return;
  const float den = total + MC_EPSIL;
  st_w[(size_t)n * K2 + t] = prod / den;
  if (t == 0) st_won[n] = (won_lanes >> base) & (K2 == 64 ? ~0ull : ((1ull << K2) - 1ull));
  if (emit) links[(bucketed ? link_off[key] : 0) + li] = make_int2(prev, comp);
}

// Density and clamped velocity of lattice cell `cell` (flat index within its
// instance) of f_all (fluid.py:268-276): the cell state the forcing starts from.
__device__ __forceinline__ void cell_state(const float* __restrict__ f_all, const InstGeo& gi,
                                           int cell, float& rho, float2& u) {
  const size_t HW = (size_t)gi.Hf * gi.Wf;
  const float* fin = f_all + 9 * gi.fl_off;
  float f[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) f[i] = fmax_nan(fin[i * HW + cell], MC_EPSIL);
  float ux, uy;
  lbm_macro(f, gi.umax, rho, ux, uy);
  u = make_float2(ux, uy);
}

// Single workgroup.  Unions the linked components (union-find in LDS when the
// components fit, else in global memory), then labels every component with
// its cluster root (the smallest component id of the cluster), counts and
// lists each cluster's members.  Also advances the actuation step counter:
// the rod kernel of this step, which read it, has completed.
__global__ void __launch_bounds__(1024) k_ib_cluster_build(
    int C, int lds_ints, const int* __restrict__ comp_node_off, const int2* __restrict__ links,
    int* __restrict__ nlinks, int* __restrict__ gparent, int* __restrict__ label,
    int* __restrict__ cl_cnt, int* __restrict__ cl_node_off, int* __restrict__ cl_mem_cnt,
    int* __restrict__ cl_mem_off, int* __restrict__ cl_members, long long* __restrict__ tick,
    WideCtl* __restrict__ wctl, int wide_min) {
  extern __shared__ __align__(16) int sm[];
  __shared__ int wsum[16];
  __shared__ int wsum2[16];
  const int tid = threadIdx.x, bs = blockDim.x;
  const bool lds = 3 * (long long)C <= lds_ints;
  int* par = lds ? sm : gparent;  // global: singletons between calls
  int* cnt = lds ? sm + C : cl_cnt;
  int* mcnt = lds ? sm + 2 * C : cl_mem_cnt;
  if (tick && tid == 0) *tick = *tick + 1;
  if (wctl && tid == 0) {
    // atomic stores: the counters are later updated by device-scope atomics
    __hip_atomic_store(&wctl->bar, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&wctl->wide, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&wctl->wide_nodes, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&wctl->nseg, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&wctl->nq, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  for (int c = tid; c < C; c += bs) {
    if (lds) par[c] = c;
    cnt[c] = 0;
    mcnt[c] = 0;
  }
  __syncthreads();
  const int nl = ld_l2(nlinks);
  for (int i = tid; i < nl; i += bs) {
    const int2 l = links[i];
    uf_union(par, l.x, l.y);
  }
  __syncthreads();
  for (int c = tid; c < C; c += bs) label[c] = uf_find(par, c);
  __syncthreads();
  for (int c = tid; c < C; c += bs) {
    if (!lds) par[c] = c;
    const int r = label[c];
    atomicAdd(&cnt[r], comp_node_off[c + 1] - comp_node_off[c]);
    atomicAdd(&mcnt[r], 1);
  }
  __syncthreads();
  int* fill = lds ? par : cl_mem_cnt;  // fill cursors
  int carry = 0, carry2 = 0;
  for (int base = 0; base < C; base += bs) {
    const int c = base + tid;
    const bool root = c < C && label[c] == c;
    // global counters were written by atomics (performed at L2): read there
    const int v = root ? (lds ? cnt[c] : ld_l2(&cnt[c])) : 0;
    const int v2 = root ? (lds ? mcnt[c] : ld_l2(&mcnt[c])) : 0;
    int tot, tot2;
    const int inc = block_inclusive_scan(v, wsum, &tot);
    const int inc2 = block_inclusive_scan(v2, wsum2, &tot2);
    if (wctl && root && v > wide_min) {
      atomicAdd(&wctl->wide, 1);
      atomicAdd(&wctl->wide_nodes, v);
    }
    if (c < C) {
      cl_node_off[c] = carry + inc - v;
      cl_mem_off[c] = carry2 + inc2 - v2;
      if (lds) {
        cl_cnt[c] = v;
        cl_mem_cnt[c] = v2;
      }
      fill[c] = 0;  // global: counts back up to the member count
    }
    carry += tot;
    carry2 += tot2;
  }
  __syncthreads();
  for (int c = tid; c < C; c += bs) {
    const int r = label[c];
    cl_members[cl_mem_off[r] + atomicAdd(&fill[r], 1)] = c;
  }
  if (tid == 0) *nlinks = 0;
}

__device__ __forceinline__ size_t al16(size_t b) { return (b + 15) & ~(size_t)15; }

struct ClusterArgs {
  Consts c;
  const int *comp_node_off, *comp_nodes, *cl_node_off, *cl_mem_cnt, *cl_mem_off, *cl_members;
  int *label, *cl_cnt;  // per component: cluster root; per root: node count
  const CompMeta* meta;
  const int *st_sx, *st_sy;
  const unsigned long long* st_won;
  const float *st_w, *vib, *f_in, *pos;
  float *vel_io, *pos_out, *force;
  float* syn_force;  // synthetic points: per-point drag out (no node update), else NULL
  int *cell_seg, *owner;
  uint8_t* mask;
  const int* comp_order;  // workgroup -> component (NULL: identity)
  BatchDev bd;
  unsigned char* gscratch;
  size_t bpn;        // global scratch bytes per node
  size_t lds_bytes;  // dynamic LDS of the launch (a multiple of 16)
  // membership: per-instance link buckets and a union-find per workgroup
  // (bucketed), else the cluster build's tables
  int bucketed;
  const int2* links;
  int* nlinks;
  const int* link_off;
  IbCtl* ictl;
  long long* tick;  // the actuation step counter, advanced once (may be NULL)
  // wide solve
  int N, C, wide_wg, wide_min;
  const int* comp_of_node;
  unsigned char* wscratch;
  int* wcount;  // per lattice cell entry counter (0 between calls)
  WideCtl* wctl;
};

// Arrays of one cluster solve.  e = r * K2 + t indexes tap t of the node of
// rank r; s indexes segments.  Placed in LDS when they fit, else in the
// cluster's global scratch; scell / srho / su0 always live in global scratch.
struct CLayout {
  int *nodes, *sstart, *eidx, *tapseg, *tmp;
  float *wt, *sinv, *gw;  // gw: the entries' weights in segment order
  float2 *corr, *vib, *suc;
};

struct GExtra {
  int* scell;
  float* srho;
  float2* su0;
};

// Global scratch of a cluster, sized for ns <= ne.
__device__ __forceinline__ void glayout(unsigned char* base, int nq, int ne, CLayout& L, GExtra& X) {
  size_t off = 0;
  auto take = [&](size_t b) {
    unsigned char* p = base + off;
    off += al16(b);
    return p;
  };
  X.scell = (int*)take(4 * (size_t)ne);
  X.srho = (float*)take(4 * (size_t)ne);
  X.su0 = (float2*)take(8 * (size_t)ne);
  L.nodes = (int*)take(4 * (size_t)nq);
  L.tmp = (int*)take(4 * (size_t)nq);
  L.sstart = (int*)take(4 * ((size_t)ne + 1));
  L.eidx = (int*)take(4 * (size_t)ne);
  L.tapseg = (int*)take(4 * (size_t)ne);
  L.wt = (float*)take(4 * (size_t)ne);
  L.gw = (float*)take(4 * (size_t)ne);
  L.corr = (float2*)take(8 * (size_t)nq);
  L.vib = (float2*)take(8 * (size_t)nq);
  L.sinv = (float*)take(4 * (size_t)ne);
  L.suc = (float2*)take(8 * (size_t)ne);
}

// The same arrays in LDS for ns segments; returns the bytes used.
__device__ __forceinline__ size_t llayout(unsigned char* base, int nq, int ne, int ns, CLayout& L) {
  size_t off = 0;
  auto take = [&](size_t b) {
    unsigned char* p = base + off;
    off += al16(b);
    return p;
  };
  L.nodes = (int*)take(4 * (size_t)nq);
  L.sstart = (int*)take(4 * ((size_t)ns + 1));
  L.eidx = (int*)take(4 * (size_t)ne);
  L.tapseg = (int*)take(4 * (size_t)ne);
  L.wt = (float*)take(4 * (size_t)ne);
  L.gw = (float*)take(4 * (size_t)ne);
  L.corr = (float2*)take(8 * (size_t)nq);
  L.vib = (float2*)take(8 * (size_t)nq);
  L.sinv = (float*)take(4 * (size_t)ns);
  L.suc = (float2*)take(8 * (size_t)ns);
  L.tmp = nullptr;
  return off;
}

// Workgroup size of the cluster solve: the unrolled K*K tap arrays of K >= 5
// need more than the 128 VGPRs a 1024-thread workgroup leaves per lane.
template <int K>
constexpr int cluster_threads() { return K <= 4 ? 1024 : 512; }

constexpr int SPT = 4;  // segments per thread whose loads are issued together

// The K2 per-tap values of one node, 16 bytes per load when the rows allow.
template <int K2>
__device__ __forceinline__ void load_taps(const float* __restrict__ p, float (&v)[K2]) {
  if constexpr (K2 % 4 == 0) {
#pragma unroll
    for (int t = 0; t < K2; t += 4) {
      const float4 q = *reinterpret_cast<const float4*>(p + t);
      v[t] = q.x;
      v[t + 1] = q.y;
      v[t + 2] = q.z;
      v[t + 3] = q.w;
    }
  } else {
#pragma unroll
    for (int t = 0; t < K2; ++t) v[t] = p[t];
  }
}

// Ascending sort of M ints in registers (odd-even transposition: M rounds).
template <int M>
__device__ __forceinline__ void sort_regs(int (&v)[M]) {
#pragma unroll
  for (int r = 0; r < M; ++r)
#pragma unroll
    for (int i = r & 1; i + 1 < M; i += 2) {
      const int lo = min(v[i], v[i + 1]), hi = max(v[i], v[i + 1]);
      v[i] = lo;
      v[i + 1] = hi;
    }
}

// Node data of the small solve, loaded by the thread of the node's rank.
template <int K2>
struct NodeIn {
  int n, sx, sy;
  unsigned long long won;
  float w[K2];
  float2 vb, p0, v0;
};

template <int K2>
__device__ __forceinline__ void load_node(const ClusterArgs& a, int n, NodeIn<K2>& d) {
  d.n = n;
  d.sx = a.st_sx[n];
  d.sy = a.st_sy[n];
  d.won = a.st_won[n];
  load_taps<K2>(a.st_w + (size_t)n * K2, d.w);
  d.vb = *reinterpret_cast<const float2*>(a.vib + 2 * n);
  // the drag's inputs, read now (only this thread writes them, at the end);
  // synthetic points hand their drag to the node pass instead
  if (!a.syn_force) {
    d.p0 = *reinterpret_cast<const float2*>(a.pos + 2 * n);
    d.v0 = *reinterpret_cast<const float2*>(a.vel_io + 2 * n);
  }
}

// LDS bytes of the small solve: per entry the grouped entry list (4 B); per
// node its stencil start, segment base and won mask (16 B) and correction
// (8 B); per segment its entry start, cursor, cell, 1/kw, rho (4 B each) and
// u0, u (8 B each); per entry its weight and segment (4 B each, tap-major:
// entry (r, t) at t nq + r; the segments' slots later hold the grouped
// entries' weights).
__device__ __forceinline__ size_t small_bytes(int nq, int ne, int ns) {
  return al16(4 * (size_t)ne) + al16(16 * (size_t)nq) + al16(8 * (size_t)nq) +
         al16(4 * ((size_t)ns + 1)) + 4 * al16(4 * (size_t)ns) + 2 * al16(8 * (size_t)ns) +
         2 * al16(4 * (size_t)ne);
}

// Small-cluster solve: one node per thread for the node phases (thread r
// holds the node of rank r and its taps' segments in registers), one entry
// (node, tap) per thread for the entry phases, every array in LDS.  The
// segments are the cells the cluster's won taps claimed, one per cell, node
// r's won taps numbered from seg0 (the exclusive prefix of the won counts) in
// tap order; the winning tap writes its segment into the lattice's cell ->
// segment map, where every tap of the cluster reads it (a cell's taps all
// belong to the cluster that claimed it).
// A segment's entries (entry id e = r K2 + t) are placed by arrival, then
// sorted by e -- by node rank, i.e. ascending node id -- by the segment's own
// thread, which folds the kernel weight in that order.  Global memory is
// touched in two places only: the cells' populations (one gather, in flight
// while the taps look their segments up) and the results at the end.
template <int K>
__device__ __forceinline__ void solve_small(const ClusterArgs& a, const InstGeo& gi,
                                            unsigned char* base, const NodeIn<K * K>& d, int nq,
                                            int ns, int seg0, int* wsum) {
  constexpr int K2 = K * K;
  const Consts& c = a.c;
  const int ne = nq * K2;
  const int W = gi.Wf, H = gi.Hf;
  const size_t HW = (size_t)H * W;
  const float umax = gi.umax;
  float* frc = a.force + 2 * gi.fl_off;
  uint8_t* msk = a.mask + gi.fl_off;
  int* own = a.owner + gi.fl_off;
  const float* fin = a.f_in + 9 * gi.fl_off;
  const int tid = threadIdx.x, bs = blockDim.x;
  int* cs = a.cell_seg + gi.fl_off;
  size_t off = 0;
  auto take = [&](size_t b) {
    unsigned char* p = base + off;
    off += al16(b);
    return p;
  };
  int* arr = (int*)take(4 * (size_t)ne);  // grouped entries: ids, then node ranks
  int4* nd = (int4*)take(16 * (size_t)nq);  // per node: sx | sy << 16, seg0, won (low, high)
  float2* corr = (float2*)take(8 * (size_t)nq);
  int* sstart = (int*)take(4 * ((size_t)ns + 1));
  int* scur = (int*)take(4 * (size_t)ns);
  int* scell = (int*)take(4 * (size_t)ns);
  float* sinv = (float*)take(4 * (size_t)ns);
  float* srho = (float*)take(4 * (size_t)ns);
  float2* su0 = (float2*)take(8 * (size_t)ns);
  float2* suc = (float2*)take(8 * (size_t)ns);
  float* wt = (float*)take(4 * (size_t)ne);   // tap-major: entry (r, t) at t nq + r
  int* eseg = (int*)take(4 * (size_t)ne);
  auto tix = [&](int e) { return (e % K2) * nq + e / K2; };

  const bool mine = tid < nq;
  const int r = tid;
  for (int s = tid; s <= ns; s += bs) sstart[s] = 0;
  if (mine) {
    nd[r] = make_int4(d.sx | (d.sy << 16), seg0, (int)(unsigned)d.won, (int)(unsigned)(d.won >> 32));
#pragma unroll
    for (int t = 0; t < K2; ++t) wt[t * nq + r] = d.w[t];
  }
  __syncthreads();
  auto ecell = [&](const int4& q, int t) {
    const int sx = q.x & 0xffff, sy = (int)((unsigned)q.x >> 16);
    return wrap1(sy + t % K, H) * W + wrap1(sx + t / K, W);
  };
  // won taps (one entry per thread, tap-major) enter their cell in the cell ->
  // segment map: tap t of node r is segment seg0(r) + (won taps of r before t)
  for (int i = tid; i < ne; i += bs) {
    const int t = i / nq, rr = i - t * nq;
    const int4 q = nd[rr];
    const unsigned long long wm = ((unsigned long long)(unsigned)q.w << 32) | (unsigned)q.z;
    if ((wm >> t) & 1ull) {
      const int g = ecell(q, t);
      const int sg = q.y + __popcll(wm & ((1ull << t) - 1ull));
      cs[g] = sg;
      scell[sg] = g;
    }
  }
  // the map's stores complete (at L2) before any wave of the workgroup reads
  // them back, past its L1
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  // per segment (at most two per thread here, more in a loop): the cell's
  // populations, in flight while the taps look their segments up
  constexpr int SPT2 = 2;
  float fp[SPT2][9];
#pragma unroll
  for (int k = 0; k < SPT2; ++k) {
    const int s = tid + k * bs;
    const int cell = s < ns ? scell[s] : 0;
#pragma unroll
    for (int i = 0; i < 9; ++i) fp[k][i] = s < ns ? fin[i * HW + cell] : 1.0f;
  }
  // every entry's segment (its cell's map entry, written above by the one
  // tap that claimed the cell) and the segments' entry counts; eight loads
  // in flight per thread
  for (int i0 = tid; i0 < ne; i0 += 8 * bs) {
    int sg[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int i = i0 + k * bs;
      sg[k] = 0;
      if (i < ne) {
        const int t = i / nq, rr = i - t * nq;
        sg[k] = __hip_atomic_load(&cs[ecell(nd[rr], t)], __ATOMIC_RELAXED,
                                  __HIP_MEMORY_SCOPE_AGENT);
      }
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int i = i0 + k * bs;
      if (i < ne) {
        eseg[i] = sg[k];
        atomicAdd(&sstart[sg[k]], 1);
      }
    }
  }
  // the cells' rho and clamped u (fluid.py:268-276)
#pragma unroll
  for (int k = 0; k < SPT2; ++k) {
    const int s = tid + k * bs;
    if (s < ns) {
      float f[9];
#pragma unroll
      for (int i = 0; i < 9; ++i) f[i] = fmax_nan(fp[k][i], MC_EPSIL);
      float rho, ux, uy;
      lbm_macro(f, umax, rho, ux, uy);
      srho[s] = rho;
      su0[s] = make_float2(ux, uy);
      suc[s] = make_float2(ux, uy);
    }
  }
  for (int s = tid + SPT2 * bs; s < ns; s += bs) {
    float rho;
    float2 u;
    cell_state(a.f_in, gi, scell[s], rho, u);
    srho[s] = rho;
    su0[s] = u;
    suc[s] = u;
  }
  __syncthreads();
  {
    int carry = 0;
    for (int b0 = 0; b0 < ns; b0 += bs) {
      const int s = b0 + tid;
      const int v = s < ns ? sstart[s] : 0;
      int tot;
      const int inc = block_inclusive_scan(v, wsum, &tot);
      if (s < ns) {
        sstart[s] = carry + inc - v;
        scur[s] = carry + inc - v;
      }
      carry += tot;
    }
    if (tid == 0) sstart[ns] = carry;  // == ne
  }
  __syncthreads();
  // entries to their segments (arrival order), and each node's segments
  for (int i = tid; i < ne; i += bs) {
    const int t = i / nq, rr = i - t * nq;
    arr[atomicAdd(&scur[eseg[i]], 1)] = rr * K2 + t;
  }
  int seg[K2];
  if (mine) {
#pragma unroll
    for (int t = 0; t < K2; ++t) seg[t] = eseg[t * nq + r];
  }
  __syncthreads();
  // each segment's entries ascending by entry id, then the kernel-weight fold
  // in that order and 1/max(kw, 1) (fluid.py:123-133,289-290); the grouped
  // list then holds each entry's node rank, and gw (the dead segment slots)
  // its weight, for the spreads
  float* gw = reinterpret_cast<float*>(eseg);
  for (int s = tid; s < ns; s += bs) {
    const int lo = sstart[s], hi = sstart[s + 1], len = hi - lo;
    float kw = 0.0f;
    if (len <= 8) {
      int v[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) v[k] = k < len ? arr[lo + k] : 0x7fffffff;
      sort_regs<8>(v);
#pragma unroll
      for (int k = 0; k < 8; ++k)
        if (k < len) {
          const float w = wt[tix(v[k])];
          arr[lo + k] = v[k] / K2;
          gw[lo + k] = w;
          kw = kw + w;
        }
    } else if (len <= 16) {
      int v[16];
#pragma unroll
      for (int k = 0; k < 16; ++k) v[k] = k < len ? arr[lo + k] : 0x7fffffff;
      sort_regs<16>(v);
#pragma unroll
      for (int k = 0; k < 16; ++k)
        if (k < len) {
          const float w = wt[tix(v[k])];
          arr[lo + k] = v[k] / K2;
          gw[lo + k] = w;
          kw = kw + w;
        }
    } else {
      for (int j = lo + 1; j < hi; ++j) {
        const int v = arr[j];
        int k = j - 1;
        while (k >= lo && arr[k] > v) {
          arr[k + 1] = arr[k];
          --k;
        }
        arr[k + 1] = v;
      }
      for (int j = lo; j < hi; ++j) {
        const int e = arr[j];
        const float w = wt[tix(e)];
        arr[j] = e / K2;
        gw[j] = w;
        kw = kw + w;
      }
    }
    sinv[s] = 1.0f / fmax_nan(kw, 1.0f);
  }
  __syncthreads();
  // a thread's first segment, when it has at most 8 entries, keeps their node
  // ranks and weights in registers for the spreads, with its 1 / max(kw, 1)
  // and its corrected velocity (only this thread writes it)
  constexpr int CSEG = 1;
  int cr[CSEG][8];
  float cw[CSEG][8];
  int clen[CSEG];
  float civ[CSEG];
  float2 cu[CSEG];
#pragma unroll
  for (int k = 0; k < CSEG; ++k) {
    const int s = tid + k * bs;
    const int lo = s < ns ? sstart[s] : 0, len = s < ns ? sstart[s + 1] - lo : 0;
    clen[k] = len <= 8 ? len : -1;
    civ[k] = s < ns ? sinv[s] : 0.0f;
    cu[k] = s < ns ? suc[s] : make_float2(0.0f, 0.0f);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const bool in = q < len && len <= 8;
      cr[k][q] = in ? arr[lo + q] : 0;
      cw[k][q] = in ? gw[lo + q] : 0.0f;
    }
  }
  // the node's tap weights, the same every iteration
  float w[K2];
  if (mine) {
#pragma unroll
    for (int t = 0; t < K2; ++t) w[t] = wt[t * nq + r];
  }
  // multi-direct forcing iterations (fluid.py:292-316)
  for (int it = 0; it < c.iters; ++it) {
    if (mine) {
      float2 u[K2];
#pragma unroll
      for (int t = 0; t < K2; ++t) u[t] = suc[seg[t]];
      float ux = 0.0f, uy = 0.0f;
#pragma unroll
      for (int t = 0; t < K2; ++t) {
        ux = ux + u[t].x * w[t];
        uy = uy + u[t].y * w[t];
      }
      const float cx = (d.vb.x - ux) * c.relax;
      const float cy = (d.vb.y - uy) * c.relax;
      const float mag = sqrtf((cx * cx + cy * cy) + MC_EPSIL);
      const float cf = fmin_nan(1.0f, umax / (mag + MC_EPSIL));
      corr[r] = make_float2(cx * cf, cy * cf);
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < CSEG; ++k) {
      const int s = tid + k * bs;
      if (s >= ns || clen[k] < 0) continue;
      float2 cv[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) cv[q] = corr[cr[k][q]];
      float ax = 0.0f, ay = 0.0f;
#pragma unroll
      for (int q = 0; q < 8; ++q)
        if (q < clen[k]) {
          ax = ax + cv[q].x * cw[k][q];
          ay = ay + cv[q].y * cw[k][q];
        }
      cu[k] = make_float2(cu[k].x + ax * civ[k], cu[k].y + ay * civ[k]);
      suc[s] = cu[k];
    }
    for (int s = tid; s < ns; s += bs) {
      const int k = (s - tid) / bs;
      if (k < CSEG && clen[k] >= 0) continue;  // done above
      const int lo = sstart[s], hi = sstart[s + 1];
      float ax = 0.0f, ay = 0.0f;
      // eight entries' loads in flight at a time, summed in order
      for (int j0 = lo; j0 < hi; j0 += 8) {
        int e[8];
        float wv[8];
        float2 cv[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          e[k] = j0 + k < hi ? arr[j0 + k] : 0;
          wv[k] = j0 + k < hi ? gw[j0 + k] : 0.0f;
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) cv[k] = corr[e[k]];
#pragma unroll
        for (int k = 0; k < 8; ++k)
          if (j0 + k < hi) {
            ax = ax + cv[k].x * wv[k];
            ay = ay + cv[k].y * wv[k];
          }
      }
      const float iv = sinv[s];
      const float2 u = suc[s];
      suc[s] = make_float2(u.x + ax * iv, u.y + ay * iv);
    }
    __syncthreads();
  }
  // post-clamp and force (fluid.py:327-339); F replaces u_corrected; the force
  // and support bits for the LBM pass; the covered cells return to idle
  for (int s = tid; s < ns; s += bs) {
    const int cell = scell[s];
    const float rho = srho[s];
    const float2 u0 = su0[s];
    float ucx = suc[s].x, ucy = suc[s].y;
    const float mag = sqrtf((ucx * ucx + ucy * ucy) + MC_EPSIL);
    const float cl = fmin_nan(1.0f, umax / (mag + MC_EPSIL));
    ucx = ucx * cl;
    ucy = ucy * cl;
    float fx = (ucx - u0.x) * rho;
    float fy = (ucy - u0.y) * rho;
    if (c.force_clamp) {
      const float fm = sqrtf((fx * fx + fy * fy) + MC_EPSIL);
      const float fc = fmin_nan(1.0f, umax / (fm + MC_EPSIL));
      fx = fx * fc;
      fy = fy * fc;
    }
    suc[s] = make_float2(fx, fy);
    frc[cell] = fx;
    frc[HW + cell] = fy;
    msk[cell] = 1;
    own[cell] = -1;
  }
  __syncthreads();
  // node drag (fluid.py:364-382) and the final wrap (simulate.py:97)
  if (mine) {
    float2 F[K2];
#pragma unroll
    for (int t = 0; t < K2; ++t) F[t] = suc[seg[t]];
    float fx = 0.0f, fy = 0.0f;
#pragma unroll
    for (int t = 0; t < K2; ++t) {
      fx = fx + F[t].x * w[t];
      fy = fy + F[t].y * w[t];
    }
    const int n = d.n;
    if (a.syn_force) {
      *reinterpret_cast<float2*>(a.syn_force + 2 * n) = make_float2(fx, fy);
    } else {
      const float dvx = (fx / c.node_mass) / (gi.scale_x * c.dt);
      const float dvy = (fy / c.node_mass) / (gi.scale_y * c.dt);
      *reinterpret_cast<float2*>(a.vel_io + 2 * n) = make_float2(d.v0.x - dvx, d.v0.y - dvy);
      *reinterpret_cast<float2*>(a.pos_out + 2 * n) =
          make_float2(fmodf_floor(d.p0.x, gi.dom_wf), fmodf_floor(d.p0.y, gi.dom_hf));
    }
  }
}

// General solve: any cluster size, arrays in LDS when they fit, else in the
// cluster's global scratch.  `nodes` holds the node ids in ascending order.
template <int K>
__device__ __forceinline__ void cluster_general(const ClusterArgs& a, const InstGeo& gi,
                                                unsigned char* smem,
                                                size_t avail, const CLayout& G, const GExtra& X,
                                                const int* nodes, int nq, int ns, int* wsum,
                                                int* s_seg) {
  constexpr int K2 = K * K;
  const Consts& c = a.c;
  const int ne = nq * K2;
  const int W = gi.Wf, H = gi.Hf;
  const size_t HW = (size_t)H * W;
  const float umax = gi.umax;
  float* frc = a.force + 2 * gi.fl_off;
  uint8_t* msk = a.mask + gi.fl_off;
  int* cseg = a.cell_seg + gi.fl_off;
  int* own = a.owner + gi.fl_off;
  const int tid = threadIdx.x, bs = blockDim.x;
  CLayout L;
  const size_t lbytes = llayout(smem, nq, ne, ns, L);
  const bool in_lds = lbytes <= avail;
  if (!in_lds) L = G;
  for (int i = tid; i < nq; i += bs) L.nodes[i] = nodes[i];
  for (int s = tid; s <= ns; s += bs) L.sstart[s] = 0;
  __syncthreads();
  // won taps claim segments (arbitrary numbering: a node's won taps take
  // consecutive ids); their cells return to idle
  for (int r = tid; r < nq; r += bs) {
    const int n = L.nodes[r];
    const unsigned long long won = a.st_won[n];
    const int sx = a.st_sx[n], sy = a.st_sy[n];
    int seg = won ? atomicAdd(s_seg, __popcll(won)) : 0;
#pragma unroll
    for (int t = 0; t < K2; ++t) {
      if ((won >> t) & 1ull) {
        const int g = wrap1(sy + t % K, H) * W + wrap1(sx + t / K, W);
        cseg[g] = seg;
        X.scell[seg] = g;
        own[g] = -1;
        ++seg;
      }
    }
  }
  __syncthreads();
  // the cells' rho and clamped u (fluid.py:268-276): four segments per thread
  // with their population loads in flight together
  {
    const float* fin = a.f_in + 9 * gi.fl_off;
    for (int s0 = tid; s0 < ns; s0 += 4 * bs) {
      int cellv[4];
      float fp[4][9];
#pragma unroll
      for (int k = 0; k < 4; ++k) cellv[k] = s0 + k * bs < ns ? X.scell[s0 + k * bs] : 0;
#pragma unroll
      for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int i = 0; i < 9; ++i) fp[k][i] = s0 + k * bs < ns ? fin[i * HW + cellv[k]] : 1.0f;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int sg = s0 + k * bs;
        if (sg >= ns) continue;
        float f[9];
#pragma unroll
        for (int i = 0; i < 9; ++i) f[i] = fmax_nan(fp[k][i], MC_EPSIL);
        float rho, ux, uy;
        lbm_macro(f, umax, rho, ux, uy);
        X.srho[sg] = rho;
        X.su0[sg] = make_float2(ux, uy);
        L.suc[sg] = make_float2(ux, uy);
      }
    }
  }
  __syncthreads();
  // every tap's segment and its arrival index there (LDS or L2 atomics)
  for (int r = tid; r < nq; r += bs) {
    const int n = L.nodes[r];
    const int sx = a.st_sx[n], sy = a.st_sy[n];
    int seg[K2];
#pragma unroll
    for (int t = 0; t < K2; ++t) seg[t] = cseg[wrap1(sy + t % K, H) * W + wrap1(sx + t / K, W)];
#pragma unroll
    for (int t = 0; t < K2; ++t) {
      L.tapseg[r * K2 + t] = seg[t];
      L.eidx[r * K2 + t] = atomicAdd(&L.sstart[seg[t]], 1);  // parked arrival index
    }
  }
  __syncthreads();
  {
    int carry = 0;
    for (int b0 = 0; b0 < ns; b0 += bs) {
      const int s = b0 + tid;
      const int v = s < ns ? (in_lds ? L.sstart[s] : ld_l2(&L.sstart[s])) : 0;
      int tot;
      const int inc = block_inclusive_scan(v, wsum, &tot);
      if (s < ns) L.sstart[s] = carry + inc - v;
      carry += tot;
    }
    if (tid == 0) L.sstart[ns] = carry;  // == ne
  }
  __syncthreads();
  // parked arrival indices become entry positions (all read before any
  // entry lands), then entries go to their segments in arrival order
  for (int e = tid; e < ne; e += bs) L.wt[e] = __int_as_float(L.sstart[L.tapseg[e]] + L.eidx[e]);
  __syncthreads();
  for (int r = tid; r < nq; r += bs) {
    const int n = L.nodes[r];
    float w[K2];
#pragma unroll
    for (int t = 0; t < K2; ++t) w[t] = a.st_w[(size_t)n * K2 + t];
    const float2 vb = make_float2(a.vib[2 * n], a.vib[2 * n + 1]);
#pragma unroll
    for (int t = 0; t < K2; ++t) {
      L.eidx[__float_as_int(L.wt[r * K2 + t])] = r * K2 + t;
      L.wt[r * K2 + t] = w[t];
    }
    L.vib[r] = vb;
  }
  __syncthreads();
  // each segment's entries ascending by entry id (by node rank), the
  // kernel-weight fold in that order and 1/max(kw, 1) (fluid.py:123-133,
  // 289-290); then eidx holds each entry's node rank and gw its weight, in
  // segment order, for the spreads
  for (int s = tid; s < ns; s += bs) {
    const int lo = L.sstart[s], hi = L.sstart[s + 1], len = hi - lo;
    float kw = 0.0f;
    if (len <= 16) {
      int v[16];
      float w[16];
#pragma unroll
      for (int k = 0; k < 16; ++k) v[k] = k < len ? L.eidx[lo + k] : 0x7fffffff;
      if (len <= 8) {
        int u[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) u[k] = v[k];
        sort_regs<8>(u);
#pragma unroll
        for (int k = 0; k < 8; ++k) v[k] = u[k];
      } else {
        sort_regs<16>(v);
      }
#pragma unroll
      for (int k = 0; k < 16; ++k) w[k] = k < len ? L.wt[v[k]] : 0.0f;
#pragma unroll
      for (int k = 0; k < 16; ++k)
        if (k < len) {
          kw = kw + w[k];
          L.eidx[lo + k] = v[k] / K2;
          L.gw[lo + k] = w[k];
        }
    } else {
      for (int j = lo + 1; j < hi; ++j) {
        const int v = L.eidx[j];
        int k = j - 1;
        while (k >= lo && L.eidx[k] > v) {
          L.eidx[k + 1] = L.eidx[k];
          --k;
        }
        L.eidx[k + 1] = v;
      }
      for (int j = lo; j < hi; ++j) {
        const int e = L.eidx[j];
        const float w = L.wt[e];
        kw = kw + w;
        L.eidx[j] = e / K2;
        L.gw[j] = w;
      }
    }
    L.sinv[s] = 1.0f / fmax_nan(kw, 1.0f);
  }
  __syncthreads();
  for (int it = 0; it < c.iters; ++it) {
    for (int r = tid; r < nq; r += bs) {
      float ux = 0.0f, uy = 0.0f;
#pragma unroll
      for (int t = 0; t < K2; ++t) {
        const float w = L.wt[r * K2 + t];
        const float2 u = L.suc[L.tapseg[r * K2 + t]];
        ux = ux + u.x * w;
        uy = uy + u.y * w;
      }
      const float2 vb = L.vib[r];
      const float cx = (vb.x - ux) * c.relax;
      const float cy = (vb.y - uy) * c.relax;
      const float mag = sqrtf((cx * cx + cy * cy) + MC_EPSIL);
      const float cf = fmin_nan(1.0f, umax / (mag + MC_EPSIL));
      L.corr[r] = make_float2(cx * cf, cy * cf);
    }
    __syncthreads();
    for (int s = tid; s < ns; s += bs) {
      float ax = 0.0f, ay = 0.0f;
      const int lo = L.sstart[s], hi = L.sstart[s + 1];
      // eight entries' loads in flight at a time, summed in order
      for (int j0 = lo; j0 < hi; j0 += 8) {
        int rk[8];
        float wv[8];
        float2 cv[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          rk[k] = j0 + k < hi ? L.eidx[j0 + k] : 0;
          wv[k] = j0 + k < hi ? L.gw[j0 + k] : 0.0f;
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) cv[k] = L.corr[rk[k]];
#pragma unroll
        for (int k = 0; k < 8; ++k)
          if (j0 + k < hi) {
            ax = ax + cv[k].x * wv[k];
            ay = ay + cv[k].y * wv[k];
          }
      }
      const float iv = L.sinv[s];
      const float2 u = L.suc[s];
      L.suc[s] = make_float2(u.x + ax * iv, u.y + ay * iv);
    }
    __syncthreads();
  }
  for (int s = tid; s < ns; s += bs) {
    const int cell = X.scell[s];
    const float rho = X.srho[s];
    const float2 u0 = X.su0[s];
    float ucx = L.suc[s].x, ucy = L.suc[s].y;
    const float mag = sqrtf((ucx * ucx + ucy * ucy) + MC_EPSIL);
    const float cl = fmin_nan(1.0f, umax / (mag + MC_EPSIL));
    ucx = ucx * cl;
    ucy = ucy * cl;
    float fx = (ucx - u0.x) * rho;
    float fy = (ucy - u0.y) * rho;
    if (c.force_clamp) {
      const float fm = sqrtf((fx * fx + fy * fy) + MC_EPSIL);
      const float fc = fmin_nan(1.0f, umax / (fm + MC_EPSIL));
      fx = fx * fc;
      fy = fy * fc;
    }
    L.suc[s] = make_float2(fx, fy);
    frc[cell] = fx;
    frc[HW + cell] = fy;
    msk[cell] = 1;
  }
  __syncthreads();
  for (int r = tid; r < nq; r += bs) {
    const int n = L.nodes[r];
    float fx = 0.0f, fy = 0.0f;
#pragma unroll
    for (int t = 0; t < K2; ++t) {
      const float w = L.wt[r * K2 + t];
      const float2 F = L.suc[L.tapseg[r * K2 + t]];
      fx = fx + F.x * w;
      fy = fy + F.y * w;
    }
    if (a.syn_force) {
      *reinterpret_cast<float2*>(a.syn_force + 2 * n) = make_float2(fx, fy);
      continue;
    }
    const float dvx = (fx / c.node_mass) / (gi.scale_x * c.dt);
    const float dvy = (fy / c.node_mass) / (gi.scale_y * c.dt);
    a.vel_io[2 * n] = a.vel_io[2 * n] - dvx;
    a.vel_io[2 * n + 1] = a.vel_io[2 * n + 1] - dvy;
    a.pos_out[2 * n] = fmodf_floor(a.pos[2 * n], gi.dom_wf);
    a.pos_out[2 * n + 1] = fmodf_floor(a.pos[2 * n + 1], gi.dom_hf);
  }
}

// ---- wide solve ----------------------------------------------------------
// Clusters of more than IB_WIDE_MIN nodes are solved together, as one
// problem, by the wide_wg workgroups of the launch after the cluster solve
// (the cluster solve skips them).  Clusters
// share no cell, so merging them changes no fold: a cell's sum is still over
// its covering nodes in ascending node id and a node's gather over its taps
// in tap order.  Entries are identified by e = n * K2 + t with the global
// node id n, so ascending e is ascending node id within a cell.
//
// Phases (grid-strided loops, separated by a grid barrier):
//   1. per node of a wide cluster: list it; per tap, an arrival index from a
//      per-cell counter; won taps claim a segment (the cell's rho / u);
//   2. per node: every tap's segment; a grid scan of the segments' entry
//      counts gives their entry ranges;
//   3. per node: entries placed in arrival order; the first gather
//      (fluid.py:292-305) -> correction;
//   4. per entry: its rank among its segment's entries (a count of smaller
//      entry ids, all loads independent) -> the segment's entries ascending;
//   5. per segment: kernel-weight fold, 1/max(kw, 1) and the first spread;
//      then per iteration a node gather and a segment spread;
//   6. per segment: post-clamp and force; per node: drag and wrap.
//
// The grid barrier: after a workgroup barrier, one thread per workgroup
// issues an agent-scope release fence, adds to the barrier counter, waits
// for all wide_wg arrivals and issues an agent-scope acquire fence; a second
// workgroup barrier releases the other threads.  The workgroups are
// co-resident: at most 256 workgroups of 256 threads (77 VGPRs, no dynamic
// LDS) take part, several fit on one CU, and the device has 256 CUs.  The
// wait is bounded nevertheless: after WideCtl::poll_limit polls (2^22 by
// default) it records the timeout in WideCtl::err and in the host flag, and
// the workgroup leaves the solve (it reads nothing another workgroup was to
// write after the barrier), so every workgroup ends.

// Scratch of the wide solve for N nodes and ne = N * K2 entries.
struct WLayout {
  int *list, *arr, *tapseg, *eplace, *esort, *scell, *sstart;
  float *srho, *sinv;
  float2 *su0, *suc, *corr;
};

__device__ __forceinline__ void wlayout(unsigned char* base, int N, int ne, WLayout& L) {
  size_t off = 0;
  auto take = [&](size_t b) {
    unsigned char* p = base + off;
    off += al16(b);
    return p;
  };
  L.list = (int*)take(4 * (size_t)N);
  L.corr = (float2*)take(8 * (size_t)N);
  L.arr = (int*)take(4 * (size_t)ne);
  L.tapseg = (int*)take(4 * (size_t)ne);
  L.eplace = (int*)take(4 * (size_t)ne);
  L.esort = (int*)take(4 * (size_t)ne);
  L.scell = (int*)take(4 * (size_t)ne);
  L.sstart = (int*)take(4 * ((size_t)ne + 1));
  L.srho = (float*)take(4 * (size_t)ne);
  L.sinv = (float*)take(4 * (size_t)ne);
  L.su0 = (float2*)take(8 * (size_t)ne);
  L.suc = (float2*)take(8 * (size_t)ne);
}

// Returns false (to every thread of the workgroup) when the wait gave up.
__device__ __forceinline__ bool grid_sync(WideCtl* ctl, unsigned int target, int* s_ok) {
  __syncthreads();
  if (threadIdx.x == 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __hip_atomic_fetch_add(&ctl->bar, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const int limit = ctl->poll_limit;
    int polls = 0, ok = 1;
    while (__hip_atomic_load(&ctl->bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
      if (polls >= limit) {
        ok = 0;
        __hip_atomic_store(&ctl->err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (ctl->host_err)
          __hip_atomic_store(ctl->host_err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
      ++polls;
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    *s_ok = ok;
  }
  __syncthreads();
  return *s_ok != 0;
}

__device__ __forceinline__ int ld_agent(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Exclusive scan of val(i), i in [0, m), across the wide workgroups (each
// scans a contiguous chunk, then adds the partial sums of the chunks before
// it); out may alias the storage val reads (each index is read by the thread
// that writes it).  Returns the total to every thread.  Ends with a barrier.
template <class F>
__device__ __forceinline__ int grid_scan(int m, F val, int* out, WideCtl* ctl, int h, int H,
                                         unsigned int& target, int* wsum, int* s_off, bool& ok) {
  const int tid = threadIdx.x, bs = blockDim.x;
  const int lo = (int)(((long long)m * h) / H), hi = (int)(((long long)m * (h + 1)) / H);
  int carry = 0;
  for (int b0 = lo; b0 < hi; b0 += bs) {
    const int i = b0 + tid;
    const int v = i < hi ? val(i) : 0;
    int tot;
    const int inc = block_inclusive_scan(v, wsum, &tot);
    if (i < hi) out[i] = carry + inc - v;
    carry += tot;
  }
  if (tid == 0) __hip_atomic_store(&ctl->bsum[h], carry, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  target += H;
  ok = grid_sync(ctl, target, &s_off[2]);
  if (!ok) return 0;
  if (tid == 0) {
    int off = 0, total = 0;
    for (int j = 0; j < H; ++j) {
      const int v = ld_agent(&ctl->bsum[j]);
      if (j < h) off += v;
      total += v;
    }
    s_off[0] = off;
    s_off[1] = total;
  }
  __syncthreads();
  const int off = s_off[0], total = s_off[1];
  if (off != 0)
    for (int i = lo + tid; i < hi; i += bs) out[i] += off;
  target += H;
  ok = grid_sync(ctl, target, &s_off[2]);
  return total;
}

template <int K>
__device__ __forceinline__ long long wide_cell(const InstGeo& gi, int sx, int sy, int t) {
  return gi.fl_off + (long long)wrap1(sy + t % K, gi.Hf) * gi.Wf + wrap1(sx + t / K, gi.Wf);
}

// Spread of one segment: the fold over its entries (ascending entry id) of
// corr[node] * weight (fluid.py:109-120), then u += sum * inv.
template <int K2>
__device__ __forceinline__ void wide_spread(const int* __restrict__ es, int lo, int hi,
                                            const float* __restrict__ w,
                                            const float2* __restrict__ corr, float iv,
                                            float2& u) {
  float ax = 0.0f, ay = 0.0f;
  int j = lo;
  for (; j + 4 <= hi; j += 4) {
    int e[4];
    float wv[4];
    float2 cv[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) e[k] = es[j + k];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      wv[k] = w[e[k]];
      cv[k] = corr[e[k] / K2];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      ax = ax + cv[k].x * wv[k];
      ay = ay + cv[k].y * wv[k];
    }
  }
  for (; j < hi; ++j) {
    const int e = es[j];
    const float wv = w[e];
    const float2 cc = corr[e / K2];
    ax = ax + cc.x * wv;
    ay = ay + cc.y * wv;
  }
  u = make_float2(u.x + ax * iv, u.y + ay * iv);
}

// Gather and correction of one node (fluid.py:90-97, 295-305).
template <int K2>
__device__ __forceinline__ float2 wide_corr(const Consts& c, const float* __restrict__ w,
                                            const int* __restrict__ seg,
                                            const float2* __restrict__ suc, float2 vb,
                                            float umax) {
  float2 u[K2];
#pragma unroll
  for (int t = 0; t < K2; ++t) u[t] = suc[seg[t]];
  float ux = 0.0f, uy = 0.0f;
#pragma unroll
  for (int t = 0; t < K2; ++t) {
    ux = ux + u[t].x * w[t];
    uy = uy + u[t].y * w[t];
  }
  const float cx = (vb.x - ux) * c.relax;
  const float cy = (vb.y - uy) * c.relax;
  const float mag = sqrtf((cx * cx + cy * cy) + MC_EPSIL);
  const float cf = fmin_nan(1.0f, umax / (mag + MC_EPSIL));
  return make_float2(cx * cf, cy * cf);
}

template <int K>
__device__ __forceinline__ bool cluster_wide(const ClusterArgs& a, int h, int H, int* wsum,
                                             int* s_off, unsigned int& target) {
  constexpr int K2 = K * K;
  const Consts& c = a.c;
  WideCtl* ctl = a.wctl;
  const int tid = threadIdx.x, bs = blockDim.x;
  const int gt = h * bs + tid, gs = H * bs;
  auto gsync = [&]() {
    target += H;
    return grid_sync(ctl, target, &s_off[2]);
  };
  WLayout L;
  wlayout(a.wscratch, a.N, a.N * K2, L);
  const float* wts = a.st_w;  // entry e = n * K2 + t has weight st_w[e]
  // 1. list the wide clusters' nodes; arrival index of every tap at its
  // cell; won taps claim segments (arbitrary numbering) and return their
  // cells to idle
  // (the wavefront walks the nodes together: list slots and segment ids come
  // from wavefront-aggregated counters)
  for (int nb = h * bs + (tid & ~63); nb < a.N; nb += gs) {
    const int n = nb + (tid & 63);
    const bool wide = n < a.N && a.cl_cnt[a.label[a.comp_of_node[n]]] > a.wide_min;
    const unsigned long long won = wide ? a.st_won[n] : 0ull;
    const int slot = wave_fetch_add(&ctl->nq, wide ? 1 : 0);
    int seg = wave_fetch_add(&ctl->nseg, __popcll(won));
    if (!wide) continue;
    L.list[slot] = n;
    const InstGeo& gi = a.bd.geo[a.bd.node_inst[n]];
    const int sx = a.st_sx[n], sy = a.st_sy[n];
#pragma unroll
    for (int t = 0; t < K2; ++t) {
      const long long cell = wide_cell<K>(gi, sx, sy, t);
      L.arr[n * K2 + t] = __hip_atomic_fetch_add(&a.wcount[cell], 1, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT);
      if ((won >> t) & 1ull) {
        a.cell_seg[cell] = seg;
        L.scell[seg] = (int)cell;  // global cell index
        a.owner[cell] = -1;
        float rho;
        float2 u;
        cell_state(a.f_in, gi, (int)(cell - gi.fl_off), rho, u);
        L.srho[seg] = rho;
        L.su0[seg] = u;
        L.suc[seg] = u;
        ++seg;
      }
    }
  }
  if (!gsync()) return false;
  const int nq = ld_agent(&ctl->nq);
  const int ns = ld_agent(&ctl->nseg);
  // 2. every tap's segment; entry ranges of the segments
  for (int r = gt; r < nq; r += gs) {
    const int n = L.list[r];
    const InstGeo& gi = a.bd.geo[a.bd.node_inst[n]];
    const int sx = a.st_sx[n], sy = a.st_sy[n];
    int seg[K2];
#pragma unroll
    for (int t = 0; t < K2; ++t) seg[t] = a.cell_seg[wide_cell<K>(gi, sx, sy, t)];
#pragma unroll
    for (int t = 0; t < K2; ++t) L.tapseg[n * K2 + t] = seg[t];
  }
  const int* wc = a.wcount;
  const int* sc = L.scell;
  bool ok = true;
  const int tot = grid_scan(ns, [&](int s) { return ld_agent(&wc[sc[s]]); }, L.sstart, ctl, h, H,
                            target, wsum, s_off, ok);
  if (!ok) return false;
  if (gt == 0) L.sstart[ns] = tot;
  // 3. entries in arrival order; the cell counters return to zero; first
  // gather and correction
  for (int r = gt; r < nq; r += gs) {
    const int n = L.list[r];
    const InstGeo& gi = a.bd.geo[a.bd.node_inst[n]];
    const unsigned long long won = a.st_won[n];
    const int sx = a.st_sx[n], sy = a.st_sy[n];
    int seg[K2];
    float w[K2];
#pragma unroll
    for (int t = 0; t < K2; ++t) {
      seg[t] = L.tapseg[n * K2 + t];
      w[t] = wts[n * K2 + t];
    }
#pragma unroll
    for (int t = 0; t < K2; ++t) {
      L.eplace[L.sstart[seg[t]] + L.arr[n * K2 + t]] = n * K2 + t;
      if ((won >> t) & 1ull) a.wcount[wide_cell<K>(gi, sx, sy, t)] = 0;
    }
    if (c.iters > 0)
      L.corr[n] = wide_corr<K2>(c, w, seg, L.suc, make_float2(a.vib[2 * n], a.vib[2 * n + 1]), gi.umax);
  }
  if (!gsync()) return false;
  // 4. each entry's rank among its segment's entries -> ascending entry ids
  for (int q = gt; q < nq * K2; q += gs) {
    const int r = q / K2, t = q - r * K2;
    const int e = L.list[r] * K2 + t;
    const int sg = L.tapseg[e];
    const int lo = L.sstart[sg], hi = L.sstart[sg + 1];
    int rank = 0;
    for (int j = lo; j < hi; ++j) rank += L.eplace[j] < e ? 1 : 0;
    L.esort[lo + rank] = e;
  }
  if (!gsync()) return false;
  // 5. kernel-weight fold and 1/max(kw, 1) (fluid.py:123-133,289-290), the
  // first spread; then the remaining iterations (fluid.py:292-316)
  for (int s = gt; s < ns; s += gs) {
    const int lo = L.sstart[s], hi = L.sstart[s + 1];
    float kw = 0.0f;
    for (int j = lo; j < hi; ++j) kw = kw + wts[L.esort[j]];
    const float iv = 1.0f / fmax_nan(kw, 1.0f);
    L.sinv[s] = iv;
    if (c.iters > 0) {
      float2 u = L.suc[s];
      wide_spread<K2>(L.esort, lo, hi, wts, L.corr, iv, u);
      L.suc[s] = u;
    }
  }
  if (!gsync()) return false;
  for (int it = 1; it < c.iters; ++it) {
    for (int r = gt; r < nq; r += gs) {
      const int n = L.list[r];
      int seg[K2];
      float w[K2];
#pragma unroll
      for (int t = 0; t < K2; ++t) {
        seg[t] = L.tapseg[n * K2 + t];
        w[t] = wts[n * K2 + t];
      }
      L.corr[n] = wide_corr<K2>(c, w, seg, L.suc, make_float2(a.vib[2 * n], a.vib[2 * n + 1]),
                                a.bd.geo[a.bd.node_inst[n]].umax);
    }
    if (!gsync()) return false;
    for (int s = gt; s < ns; s += gs) {
      float2 u = L.suc[s];
      wide_spread<K2>(L.esort, L.sstart[s], L.sstart[s + 1], wts, L.corr, L.sinv[s], u);
      L.suc[s] = u;
    }
    if (!gsync()) return false;
  }
  // 6. post-clamp and force (fluid.py:327-339); F replaces u_corrected
  for (int s = gt; s < ns; s += gs) {
    // the segment's instance: that of any covering node (its first entry)
    const InstGeo& gi = a.bd.geo[a.bd.node_inst[L.esort[L.sstart[s]] / K2]];
    const float umax = gi.umax;
    const size_t HW = (size_t)gi.Hf * gi.Wf;
    const int cell = (int)(L.scell[s] - gi.fl_off);
    const float rho = L.srho[s];
    const float2 u0 = L.su0[s];
    float ucx = L.suc[s].x, ucy = L.suc[s].y;
    const float mag = sqrtf((ucx * ucx + ucy * ucy) + MC_EPSIL);
    const float cl = fmin_nan(1.0f, umax / (mag + MC_EPSIL));
    ucx = ucx * cl;
    ucy = ucy * cl;
    float fx = (ucx - u0.x) * rho;
    float fy = (ucy - u0.y) * rho;
    if (c.force_clamp) {
      const float fm = sqrtf((fx * fx + fy * fy) + MC_EPSIL);
      const float fc = fmin_nan(1.0f, umax / (fm + MC_EPSIL));
      fx = fx * fc;
      fy = fy * fc;
    }
    L.suc[s] = make_float2(fx, fy);
    float* frc = a.force + 2 * gi.fl_off;
    frc[cell] = fx;
    frc[HW + cell] = fy;
    a.mask[gi.fl_off + cell] = 1;
  }
  if (!gsync()) return false;
  // node drag (fluid.py:364-382) and the final wrap (simulate.py:97)
  for (int r = gt; r < nq; r += gs) {
    const int n = L.list[r];
    const InstGeo& gi = a.bd.geo[a.bd.node_inst[n]];
    float2 F[K2];
    float w[K2];
#pragma unroll
    for (int t = 0; t < K2; ++t) {
      F[t] = L.suc[L.tapseg[n * K2 + t]];
      w[t] = wts[n * K2 + t];
    }
    float fx = 0.0f, fy = 0.0f;
#pragma unroll
    for (int t = 0; t < K2; ++t) {
      fx = fx + F[t].x * w[t];
      fy = fy + F[t].y * w[t];
    }
    if (a.syn_force) {
      *reinterpret_cast<float2*>(a.syn_force + 2 * n) = make_float2(fx, fy);
      continue;
    }
    const float dvx = (fx / c.node_mass) / (gi.scale_x * c.dt);
    const float dvy = (fy / c.node_mass) / (gi.scale_y * c.dt);
    a.vel_io[2 * n] = a.vel_io[2 * n] - dvx;
    a.vel_io[2 * n + 1] = a.vel_io[2 * n + 1] - dvy;
    a.pos_out[2 * n] = fmodf_floor(a.pos[2 * n], gi.dom_wf);
    a.pos_out[2 * n + 1] = fmodf_floor(a.pos[2 * n + 1], gi.dom_hf);
  }
  return true;
}

// A solve workgroup has read its instance's link bucket: the last workgroup
// of the launch returns every bucket's count to zero for the next step.
__device__ __forceinline__ void links_done(const ClusterArgs& a) {
  if (!a.bucketed || threadIdx.x != 0) return;
  const int d = __hip_atomic_fetch_add(&a.ictl->done, 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
  if (d == a.C - 1) {
    for (int i = 0; i < a.bd.n_inst; ++i)
      __hip_atomic_store(&a.nlinks[i], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&a.ictl->done, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Bucketed membership, at the end of the cluster's solve: every member's
// cluster root and the root's node count, which the wide solve reads.
__device__ __forceinline__ void write_labels(const ClusterArgs& a, const InstGeo& gi, int root,
                                             unsigned mbits, int nq) {
  for (int k = 0; k < 32 && (mbits >> k); ++k)
    if ((mbits >> k) & 1u) a.label[gi.comp0 + (int)threadIdx.x + k * (int)blockDim.x] = root;
  if (threadIdx.x == 0) a.cl_cnt[root] = nq;
}

// One workgroup per filament component; the workgroup of a cluster's smallest
// component solves the cluster (the others return).  Membership comes from a
// union-find over the instance's link bucket in LDS (bucketed), else from the
// cluster build's tables.  Every thread of a cluster is in this workgroup: the
// workgroup barrier (with its workgroup-scope fence) orders both the LDS and
// the global-memory phases.
template <int K>
__global__ void __launch_bounds__(cluster_threads<K>()) k_ib_solve(ClusterArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ int wsum[32];
  __shared__ int s_ns, s_seg, s_base, s_nm, s_nq;
  constexpr int K2 = K * K;
  const int root = a.comp_order ? a.comp_order[blockIdx.x] : (int)blockIdx.x;
  const int tid = threadIdx.x, bs = blockDim.x;
  const int inst = a.bd.comp_inst[root];  // a cluster never spans instances
  const InstGeo& gi = a.bd.geo[inst];
  if (tid == 0) s_seg = 0;
  // the rod launch of this step, which read the actuation counter, has completed
  if (root == 0 && tid == 0 && a.tick) *a.tick = *a.tick + 1;
  int nq = 0;
  unsigned mbits = 0;  // bucketed: member components tid + k bs, bit k
  // The cluster's node ids, ascending (the rank order of every fold), at the
  // top of the LDS when they fit there (twice: member order and ranked).
  // Bucketed membership keeps its tables at the bottom until the list is built.
  size_t in_bytes = 0;
  bool stage_lds = false;
  int* nodes_in = nullptr;
  int* ranked = nullptr;
  CLayout G;
  GExtra X;
  if (a.bucketed) {
    const int c0 = gi.comp0, nc = gi.ncomp;
    const int ncp = (nc + 3) & ~3;
    int* mem = reinterpret_cast<int*>(smem);  // members, ascending
    int* par = mem + ncp;                     // union-find forest
    int* cnn = par + ncp;                     // per component: node count
    int* cnf = cnn + ncp;                     // first id (contiguous) or ~(list offset)
    // the instance's component records and its links, all loads in flight
    const int nl = ld_l2(&a.nlinks[inst]);
    const int2* lk = a.links + a.link_off[inst];
    const int2 l0 = tid < nl ? lk[tid] : make_int2(c0, c0);
    for (int i = tid; i < nc; i += bs) {
      const CompMeta m = a.meta[c0 + i];
      par[i] = i;
      cnn[i] = m.nn;
      cnf[i] = m.nf >= 0 ? m.nf : ~m.n0;
    }
    __syncthreads();
    uf_union(par, l0.x - c0, l0.y - c0);
    for (int i = tid + bs; i < nl; i += bs) {
      const int2 l = lk[i];
      uf_union(par, l.x - c0, l.y - c0);
    }
    __syncthreads();
    const int me = root - c0;
    if (uf_find(par, me) != me) {
      links_done(a);
      return;
    }
    // members (ascending component ids: ascending smallest node ids), each
    // with its first position in the node list, and the node count: listed by
    // wavefront 0 alone (ballots and shuffles, no workgroup barrier per chunk)
    // per member: first position in the node list, written over the node
    // counts (member k's slot k is at most its component index, whose count
    // wavefront 0 has read by then)
    int* moff = cnn;
    if (tid < 64) {
      int carry = 0, nodes = 0;
      for (int b0 = 0; b0 < nc; b0 += 64) {
        const int i = b0 + tid;
        const bool in = i < nc && uf_find(par, i) == me;
        const unsigned long long bal = __ballot(in);
        const int cnt = in ? cnn[i] : 0;
        int inc = cnt;  // inclusive prefix of the node counts over the lanes
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
          const int u = __shfl_up(inc, d, 64);
          if (tid >= d) inc += u;
        }
        if (in) {
          const int pos = carry + __popcll(bal & ((1ull << tid) - 1ull));
          mem[pos] = i;
          moff[pos] = nodes + inc - cnt;
        }
        carry += __popcll(bal);
        nodes += __shfl(inc, 63, 64);
      }
      if (tid == 0) {
        s_nm = carry;
        s_nq = nodes;
      }
    }
    // this thread's member bits, for the labels every member gets at the end
    for (int b0 = 0, k = 0; b0 < nc; b0 += bs, ++k) {
      const int i = b0 + tid;
      if (i < nc && uf_find(par, i) == me) mbits |= 1u << k;
    }
    __syncthreads();
    const int nm = s_nm;
    nq = s_nq;
    if (a.wide_wg > 0 && nq > a.wide_min) {
      // solved by the wide workgroups
      write_labels(a, gi, root, mbits, nq);
      if (tid == 0) {
        atomicAdd(&a.wctl->wide, 1);
        atomicAdd(&a.wctl->wide_nodes, nq);
      }
      links_done(a);
      return;
    }
    in_bytes = al16(4 * (size_t)nq);
    stage_lds = 8 * in_bytes <= a.lds_bytes && 16 * (size_t)ncp + 2 * in_bytes <= a.lds_bytes;
    if (!stage_lds) {
      if (tid == 0) s_base = atomicAdd(&a.ictl->scratch_top, nq);
      __syncthreads();
      glayout(a.gscratch + (size_t)s_base * a.bpn, nq, nq * K2, G, X);
    }
    nodes_in = stage_lds ? reinterpret_cast<int*>(smem + a.lds_bytes - in_bytes) : G.tmp;
    ranked = stage_lds ? reinterpret_cast<int*>(smem + a.lds_bytes - 2 * in_bytes) : G.nodes;
    // members in order: their node ranges concatenated, one node per thread
    for (int i = tid; i < nq; i += bs) {
      int lo = 0, hi = nm - 1;  // last member whose first position is <= i
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (moff[mid] <= i) lo = mid; else hi = mid - 1;
      }
      const int f = cnf[mem[lo]], k = i - moff[lo];
      nodes_in[i] = f >= 0 ? f + k : a.comp_nodes[~f + k];
    }
    __syncthreads();
  } else {
    if (a.label[root] != root) return;
    nq = a.cl_cnt[root];
    if (a.wide_wg > 0 && nq > a.wide_min) return;  // solved by the wide workgroups
    in_bytes = al16(4 * (size_t)nq);
    stage_lds = 8 * in_bytes <= a.lds_bytes;
    glayout(a.gscratch + (size_t)a.cl_node_off[root] * a.bpn, nq, nq * K2, G, X);
    nodes_in = stage_lds ? reinterpret_cast<int*>(smem + a.lds_bytes - in_bytes) : G.tmp;
    ranked = stage_lds ? reinterpret_cast<int*>(smem + a.lds_bytes - 2 * in_bytes) : G.nodes;
    const size_t avail = stage_lds ? a.lds_bytes - 2 * in_bytes : a.lds_bytes;
    // members ordered by their smallest node id, so the gathered list is
    // already ascending unless components interleave; one record load per
    // member (node count, first id of a contiguous id range, list offset)
    const int m0 = a.cl_mem_off[root], nm = a.cl_mem_cnt[root];
    const bool m_lds = 24 * ((size_t)nm + 1) <= avail;
    int* mfirst = m_lds ? reinterpret_cast<int*>(smem) : G.eidx;
    int* msort = m_lds ? mfirst + (nm + 1) : G.tapseg;
    int* moff = m_lds ? msort + (nm + 1) : G.sstart;
    int* mnn = m_lds ? moff + (nm + 1) : reinterpret_cast<int*>(G.wt);
    int* mnf = m_lds ? mnn + (nm + 1) : reinterpret_cast<int*>(G.sinv);
    int* mn0 = m_lds ? mnf + (nm + 1) : reinterpret_cast<int*>(G.corr);
    for (int i = tid; i < nm; i += bs) {
      const CompMeta m = a.meta[a.cl_members[m0 + i]];
      // a component's node list is ascending: its first entry is the smallest id
      mfirst[i] = m.nf >= 0 ? m.nf : a.comp_nodes[m.n0];
      mnn[i] = m.nn;
      mnf[i] = m.nf;
      mn0[i] = m.n0;
    }
    __syncthreads();
    for (int i = tid; i < nm; i += bs) {
      const int v = mfirst[i];
      int rank = 0;
      for (int j = 0; j < nm; ++j) rank += mfirst[j] < v ? 1 : 0;
      msort[rank] = i;
    }
    __syncthreads();
    int carry = 0;
    for (int b0 = 0; b0 < nm; b0 += bs) {
      const int i = b0 + tid;
      const int cnt = i < nm ? mnn[msort[i]] : 0;
      int tot;
      const int inc = block_inclusive_scan(cnt, wsum, &tot);
      if (i < nm) moff[i] = carry + inc - cnt;
      carry += tot;
    }
    __syncthreads();
    for (int i = tid; i < nq; i += bs) {
      int lo = 0, hi = nm - 1;  // last member whose offset is <= i
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (moff[mid] <= i) lo = mid; else hi = mid - 1;
      }
      const int k = msort[lo], local = i - moff[lo];
      nodes_in[i] = mnf[k] >= 0 ? mnf[k] + local : a.comp_nodes[mn0[k] + local];
    }
    __syncthreads();
  }
  const int ne = nq * K2;
  const size_t avail = stage_lds ? a.lds_bytes - 2 * in_bytes : a.lds_bytes;
  int unsorted = 0;
  for (int i = tid; i + 1 < nq; i += bs) unsorted |= nodes_in[i] > nodes_in[i + 1] ? 1 : 0;
  const int* nodes = nodes_in;
  if (__syncthreads_or(unsorted)) {
    if (stage_lds) {
      // ranks by counting (ids are unique), four ids per LDS read
      for (int i = tid; i < nq; i += bs) {
        const int v = nodes_in[i];
        int rank = 0;
        int j = 0;
        for (; j + 4 <= nq; j += 4) {
          const int4 q = *reinterpret_cast<const int4*>(nodes_in + j);
          rank += (q.x < v) + (q.y < v) + (q.z < v) + (q.w < v);
        }
        for (; j < nq; ++j) rank += nodes_in[j] < v ? 1 : 0;
        ranked[rank] = v;
      }
    } else {
      for (int i = tid; i < nq; i += bs) ranked[i] = nodes_in[i];
      __syncthreads();
      sort_unique<int>(ranked, nq, reinterpret_cast<int*>(smem), 0, G.eidx, 0x7fffffff);
    }
    __syncthreads();
    nodes = ranked;
  }
  if (nq <= bs && stage_lds) {
    // the node of this thread's rank: all its data in one round trip; the
    // segments are numbered by the exclusive prefix of the won counts
    NodeIn<K2> d;
    d.won = 0;
    if (tid < nq) load_node<K2>(a, nodes[tid], d);
    const int cnt = tid < nq ? __popcll(d.won) : 0;
    int ns;
    const int inc = block_inclusive_scan(cnt, wsum, &ns);
    if (small_bytes(nq, ne, ns) <= avail) {
      solve_small<K>(a, gi, smem, d, nq, ns, inc - cnt, wsum);
      if (a.bucketed) {
        write_labels(a, gi, root, mbits, nq);
        links_done(a);
      }
      return;
    }
    if (tid == 0) s_ns = ns;
  } else {
    int won = 0;
    for (int i = tid; i < nq; i += bs) won += __popcll(a.st_won[nodes[i]]);
    int tot;
    block_inclusive_scan(won, wsum, &tot);
    if (tid == 0) s_ns = tot;  // segments = taps that won their cell
  }
  // beyond the small solve: the general solve in global scratch (a bucketed
  // cluster whose list was staged in LDS takes its scratch now)
  if (tid == 0 && a.bucketed && stage_lds) s_base = atomicAdd(&a.ictl->scratch_top, nq);
  if (tid == 0 && !a.bucketed) s_base = a.cl_node_off[root];
  __syncthreads();
  glayout(a.gscratch + (size_t)s_base * a.bpn, nq, ne, G, X);
  cluster_general<K>(a, gi, smem, avail, G, X, nodes, nq, s_ns, wsum, &s_seg);
  if (a.bucketed) {
    write_labels(a, gi, root, mbits, nq);
    links_done(a);
  }
}

// The wide solve: wide_wg workgroups launched after the cluster solve (a
// kernel of its own, so that its register demand does not reach the cluster
// kernel); they return at once when no cluster is wide.
constexpr int WIDE_THREADS = 256;  // small workgroups: a launch with nothing to do ends quickly

// The number of participating workgroups follows the problem: one per
// IB_WIDE_NODES_PER_WG wide nodes (the others return at once), so small
// problems pay for few barrier arrivals and large ones get the whole launch.
//
// With `predict`, the launch then makes the next step's inertial prediction
// (one wavefront per filament, predict_component): the filaments of wide
// clusters by the participating workgroups after a last grid barrier (their
// nodes' final state is complete), every other filament by any workgroup
// (its final state came from the cluster solve launch).  The launch then
// runs even when no instance can form a wide cluster (wide_wg 0).
template <int K>
__global__ void __launch_bounds__(WIDE_THREADS) k_ib_wide(ClusterArgs a, PredictArgs pr,
                                                          int predict) {
  __shared__ int wsum[16];
  __shared__ int s_off[3];
  const int nodes = a.wide_wg > 0 ? __hip_atomic_load(&a.wctl->wide_nodes, __ATOMIC_RELAXED,
                                                      __HIP_MEMORY_SCOPE_AGENT)
                                  : 0;
  const int H =
      nodes == 0 ? 0 : min(a.wide_wg, max(1, (nodes + IB_WIDE_NODES_PER_WG - 1) / IB_WIDE_NODES_PER_WG));
  bool wide_ready = false;  // this workgroup predicts the wide clusters' filaments
  if ((int)blockIdx.x < H) {
    if (__hip_atomic_load(&a.wctl->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
      // an earlier solve gave up at a barrier: its cell counters are not back
      // at zero until launch_ib_reset, so only the final wrap of the wide
      // clusters' nodes is done (simulate.py:97; synthetic points: by the
      // node pass); the host reports the error
      const int nw = a.syn_force ? 0 : a.N;
      for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < nw; n += H * blockDim.x) {
        if (a.cl_cnt[a.label[a.comp_of_node[n]]] <= a.wide_min) continue;
        const InstGeo& gi = a.bd.geo[a.bd.node_inst[n]];
        a.pos_out[2 * n] = fmodf_floor(a.pos[2 * n], gi.dom_wf);
        a.pos_out[2 * n + 1] = fmodf_floor(a.pos[2 * n + 1], gi.dom_hf);
      }
    } else {
      unsigned int target = 0;
      if (cluster_wide<K>(a, blockIdx.x, H, wsum, s_off, target) && predict) {
        target += H;
        wide_ready = grid_sync(a.wctl, target, &s_off[2]);
      }
    }
  }
  if (!predict) return;
  if (blockIdx.x == 0 && threadIdx.x == 0) pr.arena_top[0] = pr.arena_top[1] = 0;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int nwv = blockDim.x >> 6;
  auto is_wide = [&](int f) { return H > 0 && a.cl_cnt[a.label[f]] > a.wide_min; };
  for (int f = blockIdx.x * nwv + wave; f < a.C; f += gridDim.x * nwv)
    if (!is_wide(f)) predict_component(pr, f, lane);
  if (wide_ready)
    for (int f = blockIdx.x * nwv + wave; f < a.C; f += H * nwv)
      if (is_wide(f)) predict_component(pr, f, lane);
}

// Synthetic points: node n's drag is the mean of the forces at its points n
// and nodes + n (fluid.py:364-382), then the final wrap (simulate.py:97).
// bd is the node-level view.
__global__ void __launch_bounds__(256) k_ib_syn_drag(Consts c, BatchDev bd, int nodes,
                                                     const float2* __restrict__ F,
                                                     const float* __restrict__ pos,
                                                     float* __restrict__ vel_io,
                                                     float* __restrict__ pos_out) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nodes) return;
  const InstGeo& gi = bd.geo[bd.node_inst[n]];
  const float2 fa = F[n], fb = F[nodes + n];
  const float fx = (fa.x + fb.x) * 0.5f, fy = (fa.y + fb.y) * 0.5f;
  const float dvx = (fx / c.node_mass) / (gi.scale_x * c.dt);
  const float dvy = (fy / c.node_mass) / (gi.scale_y * c.dt);
  const float2 p0 = *reinterpret_cast<const float2*>(pos + 2 * n);
  const float2 v0 = *reinterpret_cast<const float2*>(vel_io + 2 * n);
  *reinterpret_cast<float2*>(vel_io + 2 * n) = make_float2(v0.x - dvx, v0.y - dvy);
  *reinterpret_cast<float2*>(pos_out + 2 * n) =
      make_float2(fmodf_floor(p0.x, gi.dom_wf), fmodf_floor(p0.y, gi.dom_hf));
}

__global__ void k_ib_reset(long long HW, int C, int* __restrict__ owner, int* __restrict__ parent,
                           int* __restrict__ nlinks, int nbuckets, int* __restrict__ wcount,
                           WideCtl* __restrict__ wctl, IbCtl* __restrict__ ictl) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nbuckets) nlinks[i] = 0;
  if (i == 0) {
    ictl->scratch_top = 0;
    ictl->done = 0;
    if (wctl) __hip_atomic_store(&wctl->err, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  if (i < HW) {
    owner[i] = -1;
    if (wcount) wcount[i] = 0;
  }
  if (i < C) parent[i] = (int)i;
}

}  // namespace

size_t ib_scratch_bytes_per_node(int kmax) {
  // glayout(nq, nq * K2): nq * (24 + 48 K2) + 4 + 14 arrays x 15 B of padding,
  // covered per node for nq >= 1
  const size_t K2 = (size_t)kmax * kmax;
  return 24 + 48 * K2 + 4 + 16 * 16;
}

size_t ib_wide_scratch_bytes(int N, int kmax) {
  // wlayout(N, N * K2): 12 B per node, 48 B per entry (+1 int), 12 arrays of padding
  const size_t K2 = (size_t)kmax * kmax;
  return 12 * (size_t)N + 48 * K2 * (size_t)N + 4 + 12 * 16;
}

size_t ib_lds_bytes(int lds_max) {
  // dynamic LDS of the cluster kernels: the device maximum less a reserve for
  // their static LDS (scan partials, the workgroup reduction's 256 B), a
  // multiple of 16
  return ((size_t)lds_max - 1024) & ~(size_t)15;
}

hipError_t ib_prepare(int lds_max) {
  const int lds = (int)ib_lds_bytes(lds_max);
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_ib_cluster_build),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, lds);
#define SET_LDS(KK)                                                                     \
  if (e == hipSuccess)                                                                  \
    e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_ib_solve<KK>),             \
                            hipFuncAttributeMaxDynamicSharedMemorySize, lds);
  SET_LDS(1) SET_LDS(2) SET_LDS(3) SET_LDS(4) SET_LDS(5) SET_LDS(6)
#undef SET_LDS
  return e;
}

#define MC_KSWITCH(LAUNCH)     \
  switch (c.K) {               \
    case 1: LAUNCH(1); break;  \
    case 2: LAUNCH(2); break;  \
    case 3: LAUNCH(3); break;  \
    case 4: LAUNCH(4); break;  \
    case 5: LAUNCH(5); break;  \
    default: LAUNCH(6); break; \
  }

hipError_t launch_ib_reset(const IBWork& w, hipStream_t s) {
  const long long n = std::max(std::max(w.cells, (long long)w.C), (long long)w.nbuckets);
  hipLaunchKernelGGL(k_ib_reset, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, w.cells, w.C,
                     w.owner, w.uf_parent, w.nlinks, w.nbuckets,
                     w.wide_wg > 0 ? w.wide_count : nullptr, w.wide_wg > 0 ? w.wide_ctl : nullptr,
                     w.ictl);
  return hipGetLastError();
}

hipError_t launch_ib(const Consts& c, const BatchDev& b, IBWork& w, const float* f_in,
                     const float* pos, float* vel_io, float* pos_out, long long* tick,
                     hipStream_t s, const PredictArgs* pred) {
  const int N = w.N, C = w.C;
  if (N == 0) {
    // no nodes: no support; the LBM pass takes the clamp-only force everywhere
    if (tick) return launch_tick(tick, s);
    return hipSuccess;
  }
  if (w.dirty) {
    // an earlier call stopped between its launches: restore the idle state
    hipError_t e = launch_ib_reset(w, s);
    if (e != hipSuccess) return e;
  }
  w.dirty = true;
  WideCtl* wctl = w.wide_wg > 0 ? w.wide_ctl : nullptr;
  // synthetic points: the claim and the solves see the per-point instances
  const bool syn = w.syn_force != nullptr;
  if (syn && pred) return hipErrorInvalidValue;
  BatchDev bp = b;
  if (syn) bp.node_inst = w.point_inst;
#define L_CLAIM(KK)                                                                               \
  hipLaunchKernelGGL(k_ib_claim<KK>,                                                              \
                     dim3((unsigned)(((long long)N * stencil_lanes<KK>() + 255) / 256)), dim3(256), \
                     0, s, c, bp, N, pos, vel_io, w.comp_of_node, w.st_sx, w.st_sy, w.st_w,        \
                     w.st_won, w.vib, w.owner, w.links, w.nlinks, w.link_off, w.bucketed ? 1 : 0,  \
                     wctl, w.ictl, w.nodes, syn ? w.syn_inc_off : nullptr, w.syn_inc)
  MC_KSWITCH(L_CLAIM);
#undef L_CLAIM
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (!w.bucketed) {
    hipLaunchKernelGGL(k_ib_cluster_build, dim3(1), dim3(1024), w.lds_bytes, s, C,
                       (int)(w.lds_bytes / sizeof(int)), w.comp_node_off, w.links, w.nlinks,
                       w.uf_parent, w.label, w.cl_cnt, w.cl_node_off, w.cl_mem_cnt, w.cl_mem_off,
                       w.cl_members, nullptr, wctl, w.wide_min);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  ClusterArgs a;
  a.c = c;
  a.comp_node_off = w.comp_node_off;
  a.comp_nodes = w.comp_nodes;
  a.meta = w.meta;
  a.label = w.label;
  a.cl_cnt = w.cl_cnt;
  a.cl_node_off = w.cl_node_off;
  a.cl_mem_cnt = w.cl_mem_cnt;
  a.cl_mem_off = w.cl_mem_off;
  a.cl_members = w.cl_members;
  a.st_sx = w.st_sx;
  a.st_sy = w.st_sy;
  a.st_won = w.st_won;
  a.st_w = w.st_w;
  a.vib = w.vib;
  a.f_in = f_in;
  a.pos = pos;
  a.vel_io = vel_io;
  a.pos_out = pos_out;
  a.force = w.force;
  a.syn_force = w.syn_force;
  a.cell_seg = w.cell_seg;
  a.owner = w.owner;
  a.mask = w.mask;
  a.comp_order = w.comp_order;
  a.bd = bp;
  a.gscratch = w.cl_scratch;
  a.bpn = w.cl_bpn;
  a.lds_bytes = w.lds_bytes;
  a.bucketed = w.bucketed ? 1 : 0;
  a.links = w.links;
  a.nlinks = w.nlinks;
  a.link_off = w.link_off;
  a.ictl = w.ictl;
  a.N = N;
  a.C = C;
  a.wide_wg = w.wide_wg;
  a.wide_min = w.wide_min;
  a.comp_of_node = w.comp_of_node;
  a.wscratch = w.wide_scratch;
  a.wcount = w.wide_count;
  a.wctl = w.wide_ctl;
  a.tick = tick;
#define L_CLUSTER(KK) \
  hipLaunchKernelGGL(k_ib_solve<KK>, dim3(C), dim3(cluster_threads<KK>()), w.lds_bytes, s, a)
  MC_KSWITCH(L_CLUSTER);
#undef L_CLUSTER
  e = hipGetLastError();
  if (e == hipSuccess && (w.wide_wg > 0 || pred)) {
    // the wide solve's workgroups, and with `pred` a wavefront per filament
    constexpr int waves = WIDE_THREADS / 64;
    const int G = std::max(std::max(w.wide_wg, 1), pred ? (C + waves - 1) / waves : 0);
    const PredictArgs pa = pred ? *pred : PredictArgs{};
    const int on = pred ? 1 : 0;
#define L_WIDE(KK) \
  hipLaunchKernelGGL(k_ib_wide<KK>, dim3(G), dim3(WIDE_THREADS), 0, s, a, pa, on)
    MC_KSWITCH(L_WIDE);
#undef L_WIDE
    e = hipGetLastError();
  }
  if (e == hipSuccess && syn) {
    hipLaunchKernelGGL(k_ib_syn_drag, dim3((unsigned)((w.nodes + 255) / 256)), dim3(256), 0, s, c,
                       b, w.nodes, reinterpret_cast<const float2*>(w.syn_force), pos, vel_io,
                       pos_out);
    e = hipGetLastError();
  }
  if (e == hipSuccess) w.dirty = false;
  return e;
}

}  // namespace mc
