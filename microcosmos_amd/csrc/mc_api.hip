// C ABI of the coupled fluid-filament step: plan construction (topology
// layout, FFT tables, workspace) and the stream-ordered step / rollout /
// sub-operator entry points declared in include/microcosmos_hip.h.
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/microcosmos_hip.h"
#include "mc_common.h"
#include "mc_internal.h"

#define MC_VERSION "microcosmos-hip 0.1.0 (gfx950)"

using namespace mc;

namespace {

thread_local std::string g_err;

mc_status fail(mc_status st, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
mc_status fail(mc_status st, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return st;
}

#define HIPCHK(expr)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      return fail(MC_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_),   \
                  __FILE__, __LINE__);                                                 \
  } while (0)

// Radix schedule: 4s first, then 2, 3, 5, 7, 11, 13.
bool factor(int n, std::vector<int>& out) {
  out.clear();
  if (n < 1) return false;
  while (n % 4 == 0) { out.push_back(4); n /= 4; }
  while (n % 2 == 0) { out.push_back(2); n /= 2; }
  for (int p = 3; p <= 13; p += 2) {
    while (n % p == 0) { out.push_back(p); n /= p; }
  }
  return n == 1;
}

}  // namespace

struct InstHost {
  int H, W, Hf, Wf, node0, nodes;
};

namespace {

// Plan memory comes from one stream-ordered pool per device (release
// threshold at maximum, so freed blocks stay cached for the next plan) and
// is allocated, zero-filled and released on a library-owned non-blocking
// stream.  Releasing a plan therefore never synchronises the device the way
// hipFree does, which matters for a server that retires and builds plans
// while other requests keep running.  When the pool cannot be created,
// plans fall back to hipMalloc / hipFree.
struct DevicePool {
  bool tried = false;
  hipMemPool_t pool = nullptr;
  hipStream_t stream = nullptr;
};
std::mutex g_pool_mu;
DevicePool g_pools[64];

DevicePool* device_pool(int dev) {
  if (dev < 0 || dev >= 64) return nullptr;
  std::lock_guard<std::mutex> lk(g_pool_mu);
  DevicePool& d = g_pools[dev];
  if (!d.tried) {
    d.tried = true;
    hipMemPoolProps props{};
    props.allocType = hipMemAllocationTypePinned;
    props.handleTypes = hipMemHandleTypeNone;
    props.location.type = hipMemLocationTypeDevice;
    props.location.id = dev;
    hipMemPool_t pool = nullptr;
    hipStream_t s = nullptr;
    if (hipMemPoolCreate(&pool, &props) == hipSuccess) {
      uint64_t keep = UINT64_MAX;
      if (hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep) == hipSuccess &&
          hipStreamCreateWithFlags(&s, hipStreamNonBlocking) == hipSuccess) {
        d.pool = pool;
        d.stream = s;
      } else {
        (void)hipMemPoolDestroy(pool);
      }
    }
    (void)hipGetLastError();
  }
  return d.pool ? &d : nullptr;
}

}  // namespace

// Key of the cached rollout graph: everything the captured kernels take.
struct GraphKey {
  mc_params q;
  double dt;
  mc_state sa, sb;
  mc_edge_params ed;
  Actuation act;
  int has_act;
  int has_rec;
  RecordDev rec;
  int64_t steps;
  TrackWork track;
};

struct mc_plan {
  int dev = 0;
  int N = 0, E = 0, B = 0, kmax = 0;
  int lds_max = 65536;
  // instances
  int n_inst = 0;
  std::vector<InstHost> inst;
  std::vector<InstGeo> geo_h;        // host copy of the geometry table
  std::vector<int> node_inst_h;      // instance of every node
  InstGeo* geo_d = nullptr;
  double geo_mfv = 0.0;              // max_fluid_velocity the device table holds
  bool geo_umax_set = false;         // u_max of the device table is valid
  long long grid_cells = 0, lat_cells = 0;
  int rows = 0, cols = 0, lbm_blocks = 0, max_w = 0, max_h = 0, min_hf = 0, min_wf = 0;
  BatchDev bd{};
  std::vector<void*> allocs;
  bool pooled = false;          // allocs come from the device pool
  size_t bytes = 0;
  // rod
  RodLayoutDev L{};
  int* comp_of_node = nullptr;  // component of every node
  size_t rod_lds_bytes = 0;
  int rod_use_lds = 1;
  bool rod_fuse = false;        // rollouts may run an LBM pass inside the rod launch
  bool rod_fuse_ok = false;     // the rod launch has room for it (rod_fuse may be tuned off)
  // rollouts may instead run a step's LBM pass on a second stream beside the
  // next step's steric and rod launches (it joins before that step's IB)
  bool lbm_side = false;
  hipStream_t side_stream = nullptr;
  hipEvent_t side_fork[ROLLOUT_GRAPH_STEPS] = {}, side_join[ROLLOUT_GRAPH_STEPS] = {};
  bool tuned = false;           // mc_plan_autotune chose the knobs
  bool all_chains = false;      // every component is a chain (rod: one-wavefront path only)
  float* rod_gscratch = nullptr;
  // steric
  float* pred_w = nullptr;
  int* cic_cell = nullptr;
  float* cic_w = nullptr;
  int4* comp_rows = nullptr;    // per component: steric row band, node count, ids
  int2* comp_cols = nullptr;    // per component: steric column band
  int *dep_arena = nullptr, *dep_top = nullptr, *row_sort = nullptr;
  unsigned long long* row_keys = nullptr;
  float2* spec = nullptr;
  std::vector<int2> shapes;     // distinct (H, W) steric grids, each with a kf table
  std::vector<long long> shape_kf_off;
  std::vector<bool> shape_fft;  // both sides of the shape have FFT plans
  bool fft_ok = false;
  std::string fft_why;
  // immersed boundary + LBM
  IBWork ib{};
  // the IB over synthetic points (synthetic_node_width > 0), set up on first
  // use from host tables built with the layout: per node its incident
  // (src, tgt) pairs in scatter order, and the points' component layout
  IBWork ib_syn{};
  bool ib_syn_ready = false;
  std::vector<int> syn_inc_off_h, syn_inc_h, syn_off_h, syn_pts_h, syn_pinst_h, syn_pcomp_h;
  std::vector<CompMeta> syn_meta_h;
  float* f_tmp = nullptr;
  float* node_tmp = nullptr;  // (N, 2) scratch
  long long* tick = nullptr;
  float* kf = nullptr;          // Gaussian transfer functions, one table per distinct shape
  double kf_sigma = -1.0;       // sigma the table holds (-1: none)
  bool filter_direct = false;   // some shape's table holds direct-filter taps
  bool filter_fft = false;      // some shape's table holds a transfer function
  bool filter_fused = false;    // every direct shape's kernels fit the one-launch 2-D filter
  // the CIC deposit inside the 2-D filter launch (every direct instance wider
  // and taller than a tile's window and at most one node per 32 cells)
  bool filter_dep_fused = false;
  DepEntry* tile_arena = nullptr;  // 2 x 16N entries, allocated on first use
  hipStream_t cap_stream = nullptr;
  // the instantiated S-step rollout graph of the last mc_rollout call, reused
  // while its key matches: slot 0 the ROLLOUT_GRAPH_STEPS-step graph
  // replayed for the bulk of a rollout, slot 1 the graph of its remainder;
  // `graph_done` follows the last launch of either
  GraphKey graph_key[2];
  hipGraphExec_t graph_exec[2] = {nullptr, nullptr};
  hipEvent_t graph_done = nullptr;
  // per-instance actuation time offsets (mc_batch_advance), host mirror
  long long* t_inst = nullptr;
  std::vector<long long> t_inst_h;
  // rod adjoint: largest component, checkpoints of the last recorded
  // rollout, the segment tape and the adjoint / working state
  int comp_max_nn = 0, comp_max_ne = 0, comp_max_nb = 0;
  float* adj_ckpt = nullptr;
  size_t adj_ckpt_cap = 0;
  float* adj_tape = nullptr;
  size_t adj_tape_cap = 0;
  float* adj_utape = nullptr;   // Cosserat node updates for the rest-length adjoint
  size_t adj_utape_cap = 0;
  float* adj_work = nullptr;
  float* adj_steric = nullptr;  // steric term: force (2N), pb1 (2N), V / Vbar / Dbar grids
  // fluid term: the lattice before every step of the recorded rollout, the
  // rod's recomputed outputs (x', v', debug, theta), the lattice cotangents
  // (two buffers) and the fluid adjoint's work arrays
  float* adj_ftape = nullptr;
  size_t adj_ftape_cap = 0;
  float* adj_fx = nullptr;
  float* adj_fb = nullptr;
  float* adj_fw = nullptr;
  int* adj_iw = nullptr;
  float* adj_stape = nullptr;   // steric term, per step: field, force, wrapped prediction
  size_t adj_stape_cap = 0;
  bool adj_stored = false;      // the last forward kept the steric tape
  bool adj_ok = false;
  int64_t adj_steps = 0;
  int adj_every = 0;
  double adj_dt = 0.0;
  mc_params adj_q;
  mc_edge_params adj_ed;
  // trajectory recording (mc_rollout_record): frame counter and the linear
  // resize taps of the fluid velocity, lattice -> grid
  long long* rec_tick = nullptr;
  int* rec_idx = nullptr;       // ry_off (H+1), ry_idx, rx_off (W+1), rx_idx
  float* rec_w = nullptr;       // ry_w, rx_w
  int rec_ny = 0, rec_nx = 0;   // tap counts
  int* host_err = nullptr;      // pinned: set by the wide IB solve when a barrier wait gave up
  TrackWork track{};            // displacement tracking (off: n_groups == 0)
  TrackWork inst_track{};       // per-instance tracker of mc_batch_rollout

  // While a plan is being built (`deferring`), pooled allocations, their
  // zero fills and the uploads are all queued on the pool's stream and the
  // host waits once, at the end (settle); the uploads' host data is staged
  // in `staged` until then.  Later allocations complete before they return.
  bool deferring = false;
  bool pool_pending = false;
  std::vector<std::vector<unsigned char>> staged;

  mc_status settle() {
    if (pool_pending) {
      DevicePool* dp = device_pool(dev);
      const hipError_t e = hipStreamSynchronize(dp->stream);
      if (e != hipSuccess) return fail(MC_ERR_HIP, "plan set-up: %s", hipGetErrorString(e));
      pool_pending = false;
    }
    staged.clear();
    return MC_OK;
  }

  // Host bytes to device memory allocated by `alloc`: queued behind the
  // zero fills while deferring, else a synchronous copy.
  mc_status put(void* dst, const void* src, size_t nb, const char* what) {
    if (nb == 0) return MC_OK;
    if (deferring && pooled) {
      staged.emplace_back(static_cast<const unsigned char*>(src),
                          static_cast<const unsigned char*>(src) + nb);
      const hipError_t e = hipMemcpyAsync(dst, staged.back().data(), nb, hipMemcpyHostToDevice,
                                          device_pool(dev)->stream);
      if (e != hipSuccess) return fail(MC_ERR_HIP, "upload (%s): %s", what, hipGetErrorString(e));
      pool_pending = true;
      return MC_OK;
    }
    const hipError_t e = hipMemcpy(dst, src, nb, hipMemcpyHostToDevice);
    if (e != hipSuccess) return fail(MC_ERR_HIP, "upload (%s): %s", what, hipGetErrorString(e));
    return MC_OK;
  }

  // A byte fill of memory allocated by `alloc`, ordered like `put`.
  mc_status fill(void* dst, int value, size_t nb, const char* what) {
    const hipError_t e = (deferring && pooled)
                             ? hipMemsetAsync(dst, value, nb, device_pool(dev)->stream)
                             : hipMemset(dst, value, nb);
    if (e != hipSuccess) return fail(MC_ERR_HIP, "fill (%s): %s", what, hipGetErrorString(e));
    if (deferring && pooled) pool_pending = true;
    return MC_OK;
  }

  template <class T>
  mc_status alloc(T** p, size_t count, const char* what) {
    *p = nullptr;
    if (count == 0) count = 1;
    void* q = nullptr;
    const size_t nb = count * sizeof(T);
    DevicePool* dp = device_pool(dev);
    if (allocs.empty()) pooled = dp != nullptr;
    if (pooled) {
      // zero-filled on the pool's stream; complete before return (or, while
      // the plan is being built, before it is returned): later synchronous
      // uploads on the null stream are ordered after the fill
      hipError_t e = hipMallocFromPoolAsync(&q, nb, dp->pool, dp->stream);
      if (e != hipSuccess) return fail(MC_ERR_OOM, "pool allocation (%s, %zu B): %s", what, nb, hipGetErrorString(e));
      allocs.push_back(q);
      e = hipMemsetAsync(q, 0, nb, dp->stream);
      if (e == hipSuccess && !deferring) e = hipStreamSynchronize(dp->stream);
      if (e != hipSuccess) return fail(MC_ERR_HIP, "zero fill (%s): %s", what, hipGetErrorString(e));
      if (deferring) pool_pending = true;
    } else {
      hipError_t e = hipMalloc(&q, nb);
      if (e != hipSuccess) return fail(MC_ERR_OOM, "hipMalloc(%s, %zu B): %s", what, nb, hipGetErrorString(e));
      allocs.push_back(q);
      e = hipMemset(q, 0, nb);
      if (e != hipSuccess) return fail(MC_ERR_HIP, "hipMemset(%s): %s", what, hipGetErrorString(e));
    }
    bytes += count * sizeof(T);
    *p = static_cast<T*>(q);
    return MC_OK;
  }
};

namespace {

#define MCCHK(expr)                 \
  do {                              \
    mc_status s_ = (expr);          \
    if (s_ != MC_OK) return s_;     \
  } while (0)

// Frees everything the plan holds; no work that uses it may be pending.
void release_plan(mc_plan* p) {
  if (p->host_err) (void)hipHostFree(p->host_err);
  for (hipGraphExec_t g : p->graph_exec)
    if (g) (void)hipGraphExecDestroy(g);
  if (p->graph_done) (void)hipEventDestroy(p->graph_done);
  DevicePool* dp = p->pooled ? device_pool(p->dev) : nullptr;
  for (void* q : p->allocs) {
    if (dp) (void)hipFreeAsync(q, dp->stream);
    else (void)hipFree(q);
  }
  if (dp) (void)hipStreamSynchronize(dp->stream);
  if (p->cap_stream) (void)hipStreamDestroy(p->cap_stream);
  if (p->side_stream) (void)hipStreamDestroy(p->side_stream);
  for (int k = 0; k < ROLLOUT_GRAPH_STEPS; ++k) {
    if (p->side_fork[k]) (void)hipEventDestroy(p->side_fork[k]);
    if (p->side_join[k]) (void)hipEventDestroy(p->side_join[k]);
  }
  delete p;
}

template <class T>
mc_status upload(mc_plan* p, T** dst, const std::vector<T>& v, const char* what) {
  MCCHK(p->alloc(dst, v.size(), what));
  return p->put(*dst, v.data(), v.size() * sizeof(T), what);
}

// Connected components of the edge graph (union by smaller root id), then
// the component-grouped layout with ordered incidence lists.
mc_status build_layout(mc_plan* p, const int32_t* pairs, const int32_t* bp) {
  const int N = p->N, E = p->E, B = p->B;
  for (int e = 0; e < E; ++e)
    for (int k = 0; k < 2; ++k)
      if (pairs[2 * e + k] < 0 || pairs[2 * e + k] >= N)
        return fail(MC_ERR_BAD_SHAPE, "edge %d endpoint %d out of range [0,%d)", e, pairs[2 * e + k], N);
  for (int b = 0; b < B; ++b)
    for (int k = 0; k < 2; ++k)
      if (bp[2 * b + k] < 0 || bp[2 * b + k] >= E)
        return fail(MC_ERR_BAD_SHAPE, "bending pair %d edge %d out of range [0,%d)", b, bp[2 * b + k], E);
  std::vector<int> parent(N);
  for (int i = 0; i < N; ++i) parent[i] = i;
  auto find = [&](int a) {
    int r = a;
    while (parent[r] != r) r = parent[r];
    while (parent[a] != r) { int nx = parent[a]; parent[a] = r; a = nx; }
    return r;
  };
  for (int e = 0; e < E; ++e)
    if (p->node_inst_h[pairs[2 * e]] != p->node_inst_h[pairs[2 * e + 1]])
      return fail(MC_ERR_BAD_SHAPE, "edge %d joins nodes of different instances", e);
  for (int e = 0; e < E; ++e) {
    int a = find(pairs[2 * e]), b = find(pairs[2 * e + 1]);
    if (a != b) { if (a < b) parent[b] = a; else parent[a] = b; }
  }
  std::vector<int> root(N), comp_of_node(N);
  for (int i = 0; i < N; ++i) root[i] = find(i);
  // component index = rank of the root (the smallest node id of the component)
  std::vector<int> comp_id(N, -1);
  int C = 0;
  for (int i = 0; i < N; ++i)
    if (root[i] == i) comp_id[i] = C++;
  for (int i = 0; i < N; ++i) comp_of_node[i] = comp_id[root[i]];
  std::vector<int> comp_of_edge(E), comp_of_hinge(B);
  for (int e = 0; e < E; ++e) comp_of_edge[e] = comp_of_node[pairs[2 * e]];
  for (int b = 0; b < B; ++b) {
    comp_of_hinge[b] = comp_of_edge[bp[2 * b]];
    if (comp_of_edge[bp[2 * b + 1]] != comp_of_hinge[b])
      return fail(MC_ERR_BAD_SHAPE, "bending pair %d joins edges of different components", b);
  }
  auto group = [](int n, int C, const std::vector<int>& comp, std::vector<int>& off,
                  std::vector<int>& order, std::vector<int>& local) {
    off.assign(C + 1, 0);
    for (int i = 0; i < n; ++i) off[comp[i] + 1]++;
    for (int c = 0; c < C; ++c) off[c + 1] += off[c];
    order.assign(n, 0);
    local.assign(n, 0);
    std::vector<int> fill(off.begin(), off.end() - 1);
    for (int i = 0; i < n; ++i) {  // stable: ascending ids within a component
      const int pos = fill[comp[i]]++;
      order[pos] = i;
      local[i] = pos - off[comp[i]];
    }
  };
  std::vector<int> noff, norder, nlocal, eoff, eorder, elocal, hoff, horder, hlocal;
  group(N, C, comp_of_node, noff, norder, nlocal);
  group(E, C, comp_of_edge, eoff, eorder, elocal);
  group(B, C, comp_of_hinge, hoff, horder, hlocal);
  std::vector<int> esrc(E), etgt(E), hin(B), hout(B);
  for (int q = 0; q < E; ++q) {
    const int e = eorder[q];
    esrc[q] = nlocal[pairs[2 * e]];
    etgt[q] = nlocal[pairs[2 * e + 1]];
  }
  for (int q = 0; q < B; ++q) {
    const int b = horder[q];
    hin[q] = elocal[bp[2 * b]];
    hout[q] = elocal[bp[2 * b + 1]];
  }
  // node incidences: src adds in edge order, then tgt adds in edge order
  std::vector<std::vector<int>> ninc(N);
  for (int e = 0; e < E; ++e) ninc[pairs[2 * e]].push_back((elocal[e] << 1) | 0);
  for (int e = 0; e < E; ++e) ninc[pairs[2 * e + 1]].push_back((elocal[e] << 1) | 1);
  std::vector<int> ninc_off(1, 0), ninc_flat;
  for (int q = 0; q < N; ++q) {
    for (int v : ninc[norder[q]]) ninc_flat.push_back(v);
    ninc_off.push_back((int)ninc_flat.size());
  }
  // edge incidences: e_in adds in hinge order, then e_out adds in hinge order
  std::vector<std::vector<int>> einc(E);
  for (int b = 0; b < B; ++b) einc[bp[2 * b]].push_back((hlocal[b] << 1) | 0);
  for (int b = 0; b < B; ++b) einc[bp[2 * b + 1]].push_back((hlocal[b] << 1) | 1);
  std::vector<int> einc_off(1, 0), einc_flat;
  for (int q = 0; q < E; ++q) {
    for (int v : einc[eorder[q]]) einc_flat.push_back(v);
    einc_off.push_back((int)einc_flat.size());
  }
  // LDS budget: the largest component
  size_t maxb = 0;
  for (int c = 0; c < C; ++c) {
    p->comp_max_nn = std::max(p->comp_max_nn, noff[c + 1] - noff[c]);
    p->comp_max_ne = std::max(p->comp_max_ne, eoff[c + 1] - eoff[c]);
    p->comp_max_nb = std::max(p->comp_max_nb, hoff[c + 1] - hoff[c]);
    const int nni = ninc_off[noff[c + 1]] - ninc_off[noff[c]];
    const int nei = einc_off[eoff[c + 1]] - einc_off[eoff[c]];
    const size_t b = rod_lds_bytes(noff[c + 1] - noff[c], eoff[c + 1] - eoff[c],
                                   hoff[c + 1] - hoff[c], nni, nei);
    maxb = std::max(maxb, b);
  }
  p->rod_lds_bytes = maxb;
  p->rod_use_lds = p->rod_lds_bytes + 1024 <= (size_t)p->lds_max ? 1 : 0;
  if (!p->rod_use_lds) MCCHK(p->alloc(&p->rod_gscratch, rod_scratch_floats(N, E, B), "rod scratch"));
  // the fused launch gives every workgroup the rod's LDS: keep it small enough
  // for the LBM workgroups to reach about 20 wavefronts per CU
  p->rod_fuse = !p->rod_use_lds || p->rod_lds_bytes <= 8192;
  p->rod_fuse_ok = p->rod_fuse;
  // per-component topology blocks and contiguous id ranges
  std::vector<int> blob, blob_off(1, 0), nfirst(C), efirst(C), hfirst(C);
  auto first_of = [](const std::vector<int>& off, const std::vector<int>& order, int c) {
    for (int k = off[c]; k < off[c + 1]; ++k)
      if (order[k] != order[off[c]] + (k - off[c])) return -1;
    return off[c + 1] > off[c] ? order[off[c]] : 0;
  };
  for (int c = 0; c < C; ++c) {
    for (int q = hoff[c]; q < hoff[c + 1]; ++q) blob.push_back(hin[q]);
    for (int q = hoff[c]; q < hoff[c + 1]; ++q) blob.push_back(hout[q]);
    for (int q = eoff[c]; q < eoff[c + 1]; ++q) blob.push_back(esrc[q]);
    for (int q = eoff[c]; q < eoff[c + 1]; ++q) blob.push_back(etgt[q]);
    for (int q = eoff[c]; q <= eoff[c + 1]; ++q) blob.push_back(einc_off[q] - einc_off[eoff[c]]);
    for (int q = einc_off[eoff[c]]; q < einc_off[eoff[c + 1]]; ++q) blob.push_back(einc_flat[q]);
    for (int q = noff[c]; q <= noff[c + 1]; ++q) blob.push_back(ninc_off[q] - ninc_off[noff[c]]);
    for (int q = ninc_off[noff[c]]; q < ninc_off[noff[c + 1]]; ++q) blob.push_back(ninc_flat[q]);
    while (blob.size() % 4) blob.push_back(0);  // blocks start on 16 bytes
    blob_off.push_back((int)blob.size());
    nfirst[c] = first_of(noff, norder, c);
    efirst[c] = first_of(eoff, eorder, c);
    hfirst[c] = first_of(hoff, horder, c);
  }
  std::vector<CompMeta> meta(C);
  for (int c = 0; c < C; ++c) {
    CompMeta& m = meta[c];
    m.n0 = noff[c];
    m.nn = noff[c + 1] - noff[c];
    m.e0 = eoff[c];
    m.ne = eoff[c + 1] - eoff[c];
    m.b0 = hoff[c];
    m.nb = hoff[c + 1] - hoff[c];
    m.nni = ninc_off[noff[c + 1]] - ninc_off[noff[c]];
    m.nei = einc_off[eoff[c + 1]] - einc_off[eoff[c]];
    m.nf = nfirst[c];
    m.ef = efirst[c];
    m.hf = hfirst[c];
    m.blob_off = blob_off[c];
    m.inst = p->node_inst_h[norder[noff[c]]];
    m.chain = m.pad1 = m.pad2 = 0;
    // a chain: at most 64 nodes with contiguous ids, edge e = (e, e + 1) and
    // hinge b = (b, b + 1) in component-local ids (the rod's one-wavefront path)
    bool chain = m.nn >= 1 && m.nn <= 64 && m.nf >= 0 && m.ef >= 0 && m.hf >= 0 &&
                 m.ne == m.nn - 1 && m.nb == std::max(m.ne - 1, 0);
    for (int e = 0; chain && e < m.ne; ++e)
      chain = esrc[m.e0 + e] == e && etgt[m.e0 + e] == e + 1;
    for (int b = 0; chain && b < m.nb; ++b)
      chain = hin[m.b0 + b] == b && hout[m.b0 + b] == b + 1;
    m.chain = chain ? 1 : 0;
  }
  p->all_chains = C > 0;
  for (int c = 0; c < C; ++c) p->all_chains = p->all_chains && meta[c].chain;
  // synthetic IB points (fluid.py:246-263): node n's incident edges as
  // (src, tgt) pairs, every edge at its source in edge order, then at its
  // target in edge order; point n < N is node n's + point, N + n its - point;
  // a component's points are its nodes, then its nodes + N (ascending)
  {
    std::vector<int> cnt(N + 1, 0);
    for (int e = 0; e < E; ++e) {
      cnt[pairs[2 * e] + 1]++;
      cnt[pairs[2 * e + 1] + 1]++;
    }
    for (int n = 0; n < N; ++n) cnt[n + 1] += cnt[n];
    p->syn_inc_off_h = cnt;
    p->syn_inc_h.assign(4 * (size_t)E, 0);
    std::vector<int> fill(cnt.begin(), cnt.end() - 1);
    for (int k = 0; k < 2; ++k)
      for (int e = 0; e < E; ++e) {
        const int at = fill[pairs[2 * e + k]]++;
        p->syn_inc_h[2 * at] = pairs[2 * e];
        p->syn_inc_h[2 * at + 1] = pairs[2 * e + 1];
      }
    p->syn_off_h.assign(C + 1, 0);
    p->syn_pts_h.clear();
    p->syn_pts_h.reserve(2 * (size_t)N);
    p->syn_meta_h = meta;
    for (int c = 0; c < C; ++c) {
      for (int q = noff[c]; q < noff[c + 1]; ++q) p->syn_pts_h.push_back(norder[q]);
      for (int q = noff[c]; q < noff[c + 1]; ++q) p->syn_pts_h.push_back(norder[q] + N);
      p->syn_off_h[c + 1] = 2 * noff[c + 1];
      CompMeta& m = p->syn_meta_h[c];
      m.n0 = 2 * noff[c];
      m.nn = 2 * (noff[c + 1] - noff[c]);
      m.nf = -1;
    }
    p->syn_pinst_h.assign(2 * (size_t)N, 0);
    p->syn_pcomp_h.assign(2 * (size_t)N, 0);
    for (int n = 0; n < N; ++n) {
      p->syn_pinst_h[n] = p->syn_pinst_h[N + n] = p->node_inst_h[n];
      p->syn_pcomp_h[n] = p->syn_pcomp_h[N + n] = comp_of_node[n];
    }
  }
  CompMeta* md;
  MCCHK(upload(p, &md, meta, "component records"));
  int *bl, *blo, *nf, *ef, *hf;
  MCCHK(upload(p, &bl, blob, "topology blocks"));
  MCCHK(upload(p, &blo, blob_off, "topology block offsets"));
  MCCHK(upload(p, &nf, nfirst, "node ranges"));
  MCCHK(upload(p, &ef, efirst, "edge ranges"));
  MCCHK(upload(p, &hf, hfirst, "hinge ranges"));
  int *a0, *a1, *a2, *a3, *a4, *a5, *a6, *a7, *a8, *a9, *a10, *a11, *a12, *a13;
  MCCHK(upload(p, &a0, noff, "comp_node_off"));
  MCCHK(upload(p, &a1, norder, "comp_nodes"));
  MCCHK(upload(p, &a2, eoff, "comp_edge_off"));
  MCCHK(upload(p, &a3, eorder, "comp_edges"));
  MCCHK(upload(p, &a4, hoff, "comp_hinge_off"));
  MCCHK(upload(p, &a5, horder, "comp_hinges"));
  MCCHK(upload(p, &a6, esrc, "edge_src"));
  MCCHK(upload(p, &a7, etgt, "edge_tgt"));
  MCCHK(upload(p, &a8, hin, "hinge_in"));
  MCCHK(upload(p, &a9, hout, "hinge_out"));
  MCCHK(upload(p, &a10, ninc_off, "node_inc_off"));
  MCCHK(upload(p, &a11, ninc_flat, "node_inc"));
  MCCHK(upload(p, &a12, einc_off, "edge_inc_off"));
  MCCHK(upload(p, &a13, einc_flat, "edge_inc"));
  MCCHK(upload(p, &p->comp_of_node, comp_of_node, "comp_of_node"));
  std::vector<int> comp_inst(C);
  for (int c = 0; c < C; ++c) comp_inst[c] = p->node_inst_h[norder[noff[c]]];
  // components are numbered by their smallest node and instances own
  // contiguous node ranges, so each instance's components are contiguous
  for (int i = 0, c = 0; i < p->n_inst; ++i) {
    p->geo_h[i].comp0 = c;
    while (c < C && comp_inst[c] == i) ++c;
    p->geo_h[i].ncomp = c - p->geo_h[i].comp0;
    if (i + 1 == p->n_inst && c != C)
      return fail(MC_ERR_BAD_SHAPE, "filaments are not grouped by instance");
  }
  int* ci_d;
  MCCHK(upload(p, &ci_d, comp_inst, "comp_inst"));
  p->bd.comp_inst = ci_d;
  p->L = RodLayoutDev{C, a0, a1, a2, a3, a4, a5, a6, a7, a8, a9, a10, a11, a12, a13,
                      bl, blo, nf, ef, hf, md};
  return MC_OK;
}

// Instance tables: node / row / column / LBM-workgroup ownership and the
// concatenation offsets of every per-instance array.
mc_status build_instances(mc_plan* p) {
  const int n = p->n_inst;
  p->geo_h.assign(n, InstGeo{});
  p->node_inst_h.assign(p->N, 0);
  std::vector<int> row_inst, col_inst, lbm_blk(n + 1, 0), ctile(n + 1, 0);
  long long go = 0, fo = 0;
  int r0 = 0, c0 = 0, b0 = 0;
  p->min_hf = 1 << 30;
  p->min_wf = 1 << 30;
  for (int i = 0; i < n; ++i) {
    const InstHost& h = p->inst[i];
    InstGeo& g = p->geo_h[i];
    g.H = h.H;
    g.W = h.W;
    g.Hf = h.Hf;
    g.Wf = h.Wf;
    g.dom_hf = (float)h.H;
    g.dom_wf = (float)h.W;
    // fluid.py:225-234 (scale_x uses grid_shape[0], as the reference does)
    g.scale_x = (float)((double)h.Wf / (double)h.H);
    g.scale_y = (float)((double)h.Hf / (double)h.W);
    g.mean_scale = (float)(0.5 * ((double)h.Wf / (double)h.H + (double)h.Hf / (double)h.W));
    g.grid_off = go;
    g.fl_off = fo;
    g.row0 = r0;
    g.col0 = c0;
    g.lbm_blk0 = b0;
    g.node0 = h.node0;
    g.nodes = h.nodes;
    const long long hwf = (long long)h.Hf * h.Wf;
    // floor(c / Wf) == (c * ceil(2^40 / Wf)) >> 40 for every c < hwf when hwf * Wf < 2^40
    g.wdiv = (unsigned long long)hwf * (unsigned long long)h.Wf < (1ull << 40)
                 ? ((1ull << 40) + (unsigned long long)h.Wf - 1) / (unsigned long long)h.Wf
                 : 0ull;
    go += (long long)h.H * h.W;
    fo += hwf;
    for (int y = 0; y < h.H; ++y) row_inst.push_back(i);
    for (int x = 0; x < h.W; ++x) col_inst.push_back(i);
    r0 += h.H;
    c0 += h.W;
    b0 += (int)((hwf + LBM_THREADS - 1) / LBM_THREADS);
    lbm_blk[i + 1] = b0;
    ctile[i + 1] = ctile[i] + ((h.H + FILTER_TILE_H - 1) / FILTER_TILE_H) *
                                  ((h.W + FILTER_TILE_W - 1) / FILTER_TILE_W);
    for (int k = 0; k < h.nodes; ++k) p->node_inst_h[h.node0 + k] = i;
    p->max_w = std::max(p->max_w, h.W);
    p->max_h = std::max(p->max_h, h.H);
    p->min_hf = std::min(p->min_hf, h.Hf);
    p->min_wf = std::min(p->min_wf, h.Wf);
  }
  p->grid_cells = go;
  p->lat_cells = fo;
  p->rows = r0;
  p->cols = c0;
  p->lbm_blocks = b0;
  int *ni, *ri, *cc, *lb, *ct;
  MCCHK(upload(p, &ni, p->node_inst_h, "node_inst"));
  MCCHK(upload(p, &ri, row_inst, "row_inst"));
  MCCHK(upload(p, &cc, col_inst, "col_inst"));
  MCCHK(upload(p, &lb, lbm_blk, "lbm blocks"));
  MCCHK(upload(p, &ct, ctile, "filter tiles"));
  MCCHK(p->alloc(&p->geo_d, (size_t)n, "instance geometry"));
  BatchDev& b = p->bd;
  b.n_inst = n;
  b.geo = p->geo_d;
  b.node_inst = ni;
  b.row_inst = ri;
  b.col_inst = cc;
  b.lbm_blk = lb;
  b.ctile = ct;
  b.ctiles = ctile[n];
  b.rows = p->rows;
  b.cols = p->cols;
  b.lbm_blocks = p->lbm_blocks;
  b.max_w = p->max_w;
  b.max_h = p->max_h;
  b.max_nodes = 0;
  for (const InstHost& h : p->inst) b.max_nodes = std::max(b.max_nodes, h.nodes);
  return MC_OK;
}

// Stockham plans for every distinct side length, twiddles, and one transfer-
// function table per distinct (H, W) shape.  The geometry table is uploaded
// by ensure_geo.
mc_status build_fft(mc_plan* p) {
  std::vector<int> lens;
  for (const InstHost& h : p->inst) {
    lens.push_back(h.W);
    lens.push_back(h.H);
  }
  std::sort(lens.begin(), lens.end());
  lens.erase(std::unique(lens.begin(), lens.end()), lens.end());
  std::vector<FftDev> plans;
  std::vector<float2> tw;
  std::vector<int> no_fft;  // sides without an FFT plan: the direct filter only
  for (int n : lens) {
    std::vector<int> r;
    if (!factor(n, r) || (int)r.size() > FFT_MAX_RADICES) {
      FftDev f{};
      f.n = n;
      plans.push_back(f);
      no_fft.push_back(n);
      continue;
    }
    FftDev f{};
    f.n = n;
    f.tw_off = (int)tw.size();
    f.nrad = (int)r.size();
    for (size_t k = 0; k < r.size(); ++k) f.radix[k] = r[k];
    for (int k = 0; k < n; ++k) {
      const double a = -2.0 * M_PI * (double)k / (double)n;
      tw.push_back(make_float2((float)cos(a), (float)sin(a)));
    }
    // the same values per stage, contiguous in the butterfly index k: for
    // radix R after Ns points, tw[r k n/(Ns R)] at (r-1) Ns + k, then for
    // R > 4 the R-th roots tw[m n/R]
    f.stw_off = (int)tw.size();
    int Ns = 1;
    for (int R : r) {
      const int tstride = n / (Ns * R), nb = n / R;
      for (int q = 1; q < R; ++q)
        for (int k = 0; k < Ns; ++k) {
          const float2 v = tw[f.tw_off + q * k * tstride];
          tw.push_back(v);
        }
      if (R > 4)
        for (int m = 0; m < R; ++m) {
          const float2 v = tw[f.tw_off + m * nb];
          tw.push_back(v);
        }
      Ns *= R;
    }
    f.stw_len = (int)tw.size() - f.stw_off;
    plans.push_back(f);
  }
  // row pass: W complex + max(2 W complex, deposit scratch); column pass: 3 H complex
  if (fft_rows_lds_bytes(p->max_w) + 256 > (size_t)p->lds_max ||
      fft_cols_lds_bytes(p->max_h) + 256 > (size_t)p->lds_max) {
    p->fft_why = "steric FFT row/column does not fit in LDS";
    return MC_OK;
  }
  auto plan_of = [&](int n) {
    return (int)(std::lower_bound(lens.begin(), lens.end(), n) - lens.begin());
  };
  long long kf_cells = 0;
  for (size_t i = 0; i < p->inst.size(); ++i) {
    InstGeo& g = p->geo_h[i];
    g.fft_w = plan_of(g.W);
    g.fft_h = plan_of(g.H);
    size_t sh = 0;
    while (sh < p->shapes.size() && !(p->shapes[sh].x == g.H && p->shapes[sh].y == g.W)) ++sh;
    if (sh == p->shapes.size()) {
      p->shapes.push_back(make_int2(g.H, g.W));
      p->shape_kf_off.push_back(kf_cells);
      p->shape_fft.push_back(
          !std::binary_search(no_fft.begin(), no_fft.end(), g.W) &&
          !std::binary_search(no_fft.begin(), no_fft.end(), g.H));
      kf_cells += (long long)g.H * g.W;
    }
    g.kf_off = p->shape_kf_off[sh];
  }
  FftDev* fd;
  float2* twd;
  MCCHK(upload(p, &fd, plans, "fft plans"));
  MCCHK(upload(p, &twd, tw, "twiddles"));
  p->bd.fft = fd;
  p->bd.tw = twd;
  MCCHK(p->alloc(&p->spec, (size_t)p->grid_cells, "spectrum"));
  MCCHK(p->alloc(&p->kf, (size_t)kf_cells, "transfer functions"));
  p->fft_ok = true;
  return MC_OK;
}

mc_status check_params(const mc_plan* p, const mc_params* q, bool need_fluid, bool need_steric) {
  if (!q) return fail(MC_ERR_INVALID, "params is NULL");
  if (q->solver_type != MC_SOLVER_PBD && q->solver_type != MC_SOLVER_COSSERAT)
    return fail(MC_ERR_UNSUPPORTED, "solver_type %d", q->solver_type);
  if (q->bending_constraint < MC_CONSTRAINT_AUTO || q->bending_constraint > MC_CONSTRAINT_COSSERAT)
    return fail(MC_ERR_UNSUPPORTED, "bending_constraint %d", q->bending_constraint);
  if (q->cycles_per_step < 0) return fail(MC_ERR_BAD_SHAPE, "cycles_per_step < 0");
  if (need_fluid && q->enable_fluid) {
    if (!(q->synthetic_node_width >= 0.0))
      return fail(MC_ERR_BAD_SHAPE, "synthetic_node_width must be >= 0");
    if (q->ibm_kernel_size < 1 || q->ibm_kernel_size > IB_KMAX || q->ibm_kernel_size > p->kmax)
      return fail(MC_ERR_UNSUPPORTED, "ibm_kernel_size %d (plan supports 1..%d)", q->ibm_kernel_size, p->kmax);
    if (q->ibm_kernel_size > p->min_hf || q->ibm_kernel_size > p->min_wf)
      return fail(MC_ERR_BAD_SHAPE, "ibm_kernel_size %d exceeds a lattice side (smallest %dx%d)",
                  q->ibm_kernel_size, p->min_hf, p->min_wf);
    if (q->ibm_iterations < 0) return fail(MC_ERR_BAD_SHAPE, "ibm_iterations < 0");
  }
  if (need_steric && q->enable_steric) {
    if (!p->fft_ok) return fail(MC_ERR_UNSUPPORTED, "%s", p->fft_why.c_str());
    if (q->steric_neighbor_skip < 0) return fail(MC_ERR_BAD_SHAPE, "steric_neighbor_skip < 0");
  }
  return MC_OK;
}

Consts derive(const mc_plan* p, const mc_params* q, double dt) {
  Consts c;
  memset(&c, 0, sizeof(c));
  c.dt = (float)dt;
  c.damping = (float)q->damping;
  c.enable_fluid = q->enable_fluid != 0;
  c.enable_steric = q->enable_steric != 0;
  const double s = q->steric_sigma;
  c.steric_strength = (float)q->steric_strength;
  c.steric_sv = (float)q->steric_scatter_value;
  c.steric_2s2 = (float)(2.0 * s * s);                       // forces.py:76
  c.steric_norm = (float)(2.0 * M_PI * s * s);               // forces.py:76
  c.steric_svs = (float)(q->steric_scatter_value / (s * s)); // forces.py:77
  c.fft_coef = (float)(-2.0 * (M_PI * M_PI) * (s * s));      // utils.py:47
  c.steric_nskip = q->steric_neighbor_skip;
  // the kernel follows bending_constraint (simulate.py:74-77); solver_type
  // only decides whether the inertial target is set (simulate.py:71-72)
  c.solver = q->bending_constraint == MC_CONSTRAINT_AUTO
                 ? q->solver_type
                 : (q->bending_constraint == MC_CONSTRAINT_COSSERAT ? 1 : 0);
  c.cos_target = q->solver_type == MC_SOLVER_COSSERAT;
  c.cycles = q->cycles_per_step;
  c.k_st = (float)q->stiffness_stretch;
  c.k_sh = (float)q->stiffness_shear;
  c.cos_w = (float)(1.0 / (q->config_dt * q->config_dt));   // stable_cosserat.py:186
  c.K = q->ibm_kernel_size;
  c.half = (c.K - 1) / 2;                                    // fluid.py:48
  c.half_width = (float)(c.K / 2.0);                         // fluid.py:49
  c.iters = q->ibm_iterations;
  c.force_clamp = q->enable_force_clamp != 0;
  c.relax = (float)q->ibm_relaxation;
  c.node_mass = (float)q->node_mass;
  c.syn_d = (float)q->synthetic_node_width;
  c.synthetic = q->enable_fluid && q->synthetic_node_width > 0.0;
  // fluid.py:155-161,342-344, float32 arithmetic after the double 1/(3 nu + 0.5)
  const float eps = 1.1920928955078125e-07f;
  float op = (float)(1.0 / (3.0 * q->viscosity + 0.5));
  op = op < 0.0f ? 0.0f : (op > 1.9f ? 1.9f : op);
  const float tpo = 1.0f / std::max(op, eps) - 0.5f;
  const float om = 1.0f / ((float)q->lambda_trt / std::max(tpo, eps) + 0.5f);
  c.omega_p = op;
  c.omega_m = om;
  c.cp = 1.0f - 0.5f * op;
  c.cm = 1.0f - 0.5f * om;
  (void)p;
  return c;
}

// Per-instance u_max = max_fluid_velocity * mean scale (fluid.py:225-234, in
// double, then float32) into the device geometry table.  Called before any
// stream capture; uploads only when max_fluid_velocity changes.
mc_status ensure_geo(mc_plan* p, const mc_params* q) {
  if (p->geo_umax_set && p->geo_mfv == q->max_fluid_velocity) return MC_OK;
  for (size_t i = 0; i < p->inst.size(); ++i) {
    const InstHost& h = p->inst[i];
    const double sx = (double)h.Wf / (double)h.H;
    const double sy = (double)h.Hf / (double)h.W;
    p->geo_h[i].umax = (float)(q->max_fluid_velocity * (0.5 * (sx + sy)));
  }
  HIPCHK(hipMemcpy(p->geo_d, p->geo_h.data(), p->geo_h.size() * sizeof(InstGeo),
                   hipMemcpyHostToDevice));
  p->geo_mfv = q->max_fluid_velocity;
  p->geo_umax_set = true;
  return MC_OK;
}

// The periodic 1-D kernel of the Gaussian filter along an axis of length n:
// h[m] = (1/n) sum_k exp(-2 pi^2 s^2 f_k^2) cos(2 pi k m / n) (the inverse DFT
// of one factor of the transfer function, utils.py:41-47), k summed in order,
// in double.  Keeps offsets lo .. lo + count - 1: the smallest radius whose
// tail holds at most 2^-27 of the kernel's |mass|, or all n offsets once when
// the kernel wraps.  False when more than FILTER_TAPS_MAX taps are needed.
static bool gauss_taps(int n, double sigma, int* lo, std::vector<float>* taps) {
  std::vector<double> T(n), C(n), h(n, 0.0);
  for (int k = 0; k < n; ++k) {
    const double f = (double)(k < (n + 1) / 2 ? k : k - n) / (double)n;
    T[k] = exp((-2.0 * M_PI * M_PI * sigma * sigma) * (f * f));
    C[k] = cos(2.0 * M_PI * (double)k / (double)n);
  }
  for (int k = 0; k < n; ++k)
    for (int m = 0; m < n; ++m) h[m] = h[m] + T[k] * C[(int)(((long long)k * m) % n)];
  double total = 0.0;
  for (int m = 0; m < n; ++m) {
    h[m] = h[m] / (double)n;
    total += fabs(h[m]);
  }
  int R = 0;
  for (;; ++R) {
    if (R >= n / 2) break;
    double tail = 0.0;
    for (int m = 0; m < n; ++m)
      if (std::min(m, n - m) > R) tail += fabs(h[m]);
    if (tail <= std::ldexp(1.0, -27) * total) break;
  }
  int cnt;
  if (2 * R + 1 >= n) {
    *lo = -((n - 1) / 2);
    cnt = n;
  } else {
    *lo = -R;
    cnt = 2 * R + 1;
  }
  if (cnt > FILTER_TAPS_MAX) return false;
  taps->resize(cnt);
  for (int t = 0; t < cnt; ++t) (*taps)[t] = (float)h[(((*lo + t) % n) + n) % n];
  return true;
}

// Tabulates the Gaussian filter for `sigma` when the plan does not hold it
// yet, per steric grid shape: the direct filter's taps when both of the
// shape's kernels are short, else the transfer function of the FFT path
// (utils.py:41-47).  The kernels tell the two apart by the table's first word:
// lo_x <= 0 as an int, or exp(0) = 1.0f.  Called before any stream capture.
mc_status ensure_kf(mc_plan* p, double sigma, hipStream_t s) {
  if (!p->fft_ok || p->kf_sigma == sigma) return MC_OK;
  const float coef = (float)(-2.0 * (M_PI * M_PI) * (sigma * sigma));  // utils.py:47
  bool any_direct = false, any_fft = false;
  size_t max_taps = 0;
  std::vector<float> tab;
  for (size_t i = 0; i < p->shapes.size(); ++i) {
    const int H = p->shapes[i].x, W = p->shapes[i].y;
    int lx, ly;
    std::vector<float> tx, ty;
    const bool direct = gauss_taps(W, sigma, &lx, &tx) && gauss_taps(H, sigma, &ly, &ty) &&
                        4 + tx.size() + ty.size() <= (size_t)H * W;
    if (!direct) {
      if (!p->shape_fft[i]) {
        p->kf_sigma = -1.0;  // the tables are no longer those of any sigma
        return fail(MC_ERR_UNSUPPORTED,
                    "steric filter: sigma %g on a %dx%d grid needs more than %d taps and an FFT, "
                    "whose sides must have prime factors <= 13", sigma, H, W, FILTER_TAPS_MAX);
      }
      any_fft = true;
      HIPCHK(launch_gauss_table(H, W, coef, p->kf + p->shape_kf_off[i], s));
      continue;
    }
    any_direct = true;
    max_taps = std::max(max_taps, std::max(tx.size(), ty.size()));
    const int hd[4] = {lx, (int)tx.size(), ly, (int)ty.size()};
    tab.assign(4, 0.0f);
    memcpy(tab.data(), hd, sizeof(hd));
    tab.insert(tab.end(), tx.begin(), tx.end());
    tab.insert(tab.end(), ty.begin(), ty.end());
    // ordered after every earlier use of the table on `s`; the wait keeps the
    // host staging alive until the copy is done
    HIPCHK(hipMemcpyAsync(p->kf + p->shape_kf_off[i], tab.data(), tab.size() * sizeof(float),
                          hipMemcpyHostToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
  }
  p->filter_direct = any_direct;
  p->filter_fft = any_fft;
  p->filter_fused = any_direct && max_taps <= 33 && !getenv("MC_FILTER_TWO_PASS");
  // the deposit joins the 2-D filter launch when each window holds every
  // grid cell at most once and the nodes are sparse (a window's entries then
  // fit the LDS); MC_DEPOSIT_ROWS keeps the row deposit launch
  bool dep_fused = p->filter_fused && !getenv("MC_DEPOSIT_ROWS") && p->N > 0;
  for (size_t i = 0; i < p->shapes.size() && dep_fused; ++i) {
    const int H = p->shapes[i].x, W = p->shapes[i].y;
    int lx, ly;
    std::vector<float> tx, ty;
    if (!(gauss_taps(W, sigma, &lx, &tx) && gauss_taps(H, sigma, &ly, &ty))) continue;
    dep_fused = H >= FILTER_TILE_H + (int)ty.size() - 1 && W >= FILTER_TILE_W + (int)tx.size() - 1;
  }
  for (const InstHost& h : p->inst)
    dep_fused = dep_fused && (long long)h.nodes * 32 <= (long long)h.H * h.W;
  if (dep_fused && !p->tile_arena)
    MCCHK(p->alloc(&p->tile_arena, 32 * (size_t)p->N, "fused deposit arena"));
  p->filter_dep_fused = dep_fused;
  p->kf_sigma = sigma;
  return MC_OK;
}

// The Gaussian filter of the plan's current tables over every instance: of
// real_in, or of the CIC deposit described by `dep` (also written to
// density_out when not NULL).  Each path's kernels skip the other's shapes;
// the direct path's row-pass scratch is the spectrum buffer, free by then.
hipError_t run_filter(mc_plan* p, const float* real_in, const StericDep* dep, float* density_out,
                      float* out, hipStream_t s) {
  if (p->filter_direct) {
    const hipError_t e = launch_steric_conv(p->bd, real_in, dep, reinterpret_cast<float*>(p->spec),
                                            p->kf, density_out, out, p->filter_fused, s);
    if (e != hipSuccess) return e;
  }
  if (p->filter_fft)
    return launch_steric_fft_T(p->bd, real_in, dep, p->spec, p->kf, density_out, out, s);
  return hipSuccess;
}

// Steric pipeline on `pos` (+ optional velocity prediction): fills p->pred_w,
// writes the filtered field to steric_out.
mc_status run_steric(mc_plan* p, const Consts& c, const float* pos, const float* vel,
                     float* steric_out, float* density_out, hipStream_t s,
                     float* pred_out = nullptr, bool pred_ready = false) {
  if (!pred_ready)
    HIPCHK(launch_predict(c, p->bd, p->L, pos, vel, pred_out ? pred_out : p->pred_w, p->cic_cell,
                          p->cic_w, p->comp_rows, p->comp_cols, p->dep_top, s));
  if (!c.enable_steric) return MC_OK;
  StericDep d;
  d.L = p->L;
  d.comp_rows = p->comp_rows;
  d.comp_cols = p->comp_cols;
  d.cic_cell = p->cic_cell;
  d.cic_w = p->cic_w;
  d.arena = p->dep_arena;
  d.arena_top = p->dep_top;
  d.sort_tmp = p->row_sort;
  d.key_tmp = p->row_keys;
  d.tile_arena = p->filter_dep_fused ? p->tile_arena : nullptr;
  d.tile_top = p->dep_top + 1;
  d.key_half = (size_t)4 * p->N;
  HIPCHK(run_filter(p, nullptr, &d, density_out, steric_out, s));
  return MC_OK;
}

// Optional timing events around the LBM launch of one step.
struct StepSync {
  hipEvent_t t0 = nullptr, t1 = nullptr;
  bool on_rod = false;  // the events bracket the rod launch (carrying an LBM pass)
};

// The previous step's LBM pass on a second stream, forked from `s` at the
// start of a step and joined before its IB (see enqueue_step).
struct SideLbm {
  const LbmArgs* lbm = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
};

mc_status ensure_side(mc_plan* p) {
  if (p->side_stream) return MC_OK;
  HIPCHK(hipStreamCreateWithFlags(&p->side_stream, hipStreamNonBlocking));
  for (int k = 0; k < ROLLOUT_GRAPH_STEPS; ++k) {
    HIPCHK(hipEventCreateWithFlags(&p->side_fork[k], hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&p->side_join[k], hipEventDisableTiming));
  }
  return MC_OK;
}

// Whether the steps of a sequence on one stream can hand the next step's
// inertial prediction to the IB launch that ends the step before (the fluid
// runs; MC_PREDICT_LAUNCH keeps the prediction launch of every step).
bool chain_predict(const mc_plan* p, const Consts& c) {
  static const bool off = getenv("MC_PREDICT_LAUNCH") != nullptr;
  // synthetic points end the IB with the node pass, not the wide solve launch
  return c.enable_fluid && p->N > 0 && !off && !c.synthetic;
}

// The IB work of the call: over the nodes, or over the synthetic points
// (set up by ensure_synth).
IBWork& ib_of(mc_plan* p, const Consts& c) {
  if (!c.synthetic) return p->ib;
  p->ib_syn.wide_min = p->ib.wide_min;  // tuned on the plan's own IB
  return p->ib_syn;
}

// Enqueues one coupled step on `s`.  The LBM pass of the step may be deferred
// (defer_lbm != NULL: its arguments are returned instead of launched); the
// next step then runs it inside its rod launch (fuse_lbm), where it fills the
// CUs the filaments leave idle.  That is valid because the rod and steric
// work of the next step neither read the lattice nor the IB force, and its IB
// (which reads the new lattice) is enqueued after the rod launch.
//
// With `side`, the previous step's deferred LBM pass runs on a second stream
// instead, beside this step's steric and rod launches (which read neither the
// lattice nor the IB force and mask), and joins `s` before this step's IB,
// which reads the lattice it writes and rewrites the force and mask it reads.
//
// pred_ready: the previous step on this stream (whose output is `in`) already
// made this step's inertial prediction; tail_pred: this step makes the next
// step's, at the end of its IB (both only where chain_predict holds).
mc_status enqueue_step(mc_plan* p, const mc_params* q, const Consts& c, const mc_state* in,
                       const mc_state* out, const mc_edge_params* ed, const Actuation* act,
                       hipStream_t s, const StepSync& sy = StepSync(),
                       const LbmArgs* fuse_lbm = nullptr, LbmArgs* defer_lbm = nullptr,
                       const SideLbm& side = SideLbm(), bool pred_ready = false,
                       bool tail_pred = false) {
  if (side.lbm) {
    HIPCHK(hipEventRecord(side.fork, s));
    HIPCHK(hipStreamWaitEvent(side.stream, side.fork, 0));
    const LbmArgs& l = *side.lbm;
    HIPCHK(launch_lbm(c, p->bd, l.fin, l.fout, l.ustar, l.mask, l.force, side.stream, nullptr,
                      nullptr));
    HIPCHK(hipEventRecord(side.join, side.stream));
  }
  MCCHK(run_steric(p, c, in->position, in->velocity, out->steric, nullptr, s, nullptr,
                   pred_ready));
  if (!c.enable_steric && out->steric != in->steric)
    HIPCHK(hipMemcpyAsync(out->steric, in->steric, (size_t)p->grid_cells * sizeof(float),
                          hipMemcpyDeviceToDevice, s));
  HIPCHK(launch_rod(c, p->bd, p->L, p->N, p->rod_lds_bytes, p->rod_use_lds, p->rod_gscratch, 0,
                    in->position, in->velocity, nullptr, in->theta, ed->rest_lengths,
                    ed->bending_rest_angles, ed->bending_stiffness, act, out->steric, p->pred_w,
                    out->position, out->velocity, out->debug_vector, out->theta, fuse_lbm, s,
                    p->all_chains, sy.on_rod ? sy.t0 : nullptr, sy.on_rod ? sy.t1 : nullptr));
  const size_t hwf = (size_t)p->lat_cells;
  long long* tick = act ? p->tick : nullptr;
  if (side.lbm) HIPCHK(hipStreamWaitEvent(s, side.join, 0));
  if (c.enable_fluid) {
    PredictArgs pa{c, p->bd, p->L, out->position, out->velocity, p->pred_w, p->cic_cell,
                   p->cic_w, p->comp_rows, p->comp_cols, p->dep_top};
    if (c.synthetic && !p->ib_syn_ready)
      return fail(MC_ERR_INVALID, "synthetic IB points are not set up on this plan");
    HIPCHK(launch_ib(c, p->bd, ib_of(p, c), in->f_grid, out->position, out->velocity, out->position,
                     tick, s, tail_pred ? &pa : nullptr));
    if (defer_lbm) {
      *defer_lbm = make_lbm_args(c, p->bd, in->f_grid, out->f_grid, out->fluid_velocity, p->ib.mask,
                                 p->ib.force);
    } else {
      HIPCHK(launch_lbm(c, p->bd, in->f_grid, out->f_grid, out->fluid_velocity, p->ib.mask,
                        p->ib.force, s, sy.on_rod ? nullptr : sy.t0, sy.on_rod ? nullptr : sy.t1));
    }
  } else {
    if (tick) HIPCHK(launch_tick(tick, s));
    if (out->f_grid != in->f_grid)
      HIPCHK(hipMemcpyAsync(out->f_grid, in->f_grid, 9 * hwf * sizeof(float), hipMemcpyDeviceToDevice, s));
    if (out->fluid_velocity != in->fluid_velocity)
      HIPCHK(hipMemcpyAsync(out->fluid_velocity, in->fluid_velocity, 2 * hwf * sizeof(float),
                            hipMemcpyDeviceToDevice, s));
  }
  if (p->track.n_groups > 0 && p->N > 0)
    HIPCHK(launch_track(p->track, out->position, 1, s));
  (void)q;
  return MC_OK;
}

mc_status check_state(const mc_plan* p, const mc_state* st, const char* which) {
  if (!st) return fail(MC_ERR_INVALID, "%s state is NULL", which);
  if (p->N > 0 && (!st->position || !st->velocity || !st->debug_vector))
    return fail(MC_ERR_INVALID, "%s node arrays are NULL", which);
  if (p->E > 0 && !st->theta) return fail(MC_ERR_INVALID, "%s theta is NULL", which);
  if (!st->steric || !st->fluid_velocity || !st->f_grid)
    return fail(MC_ERR_INVALID, "%s field arrays are NULL", which);
  return MC_OK;
}

mc_status check_edges(const mc_plan* p, const mc_edge_params* ed) {
  if (!ed) return fail(MC_ERR_INVALID, "edge params are NULL");
  if (p->E > 0 && !ed->rest_lengths) return fail(MC_ERR_INVALID, "rest_lengths is NULL");
  if (p->B > 0 && (!ed->bending_rest_angles || !ed->bending_stiffness))
    return fail(MC_ERR_INVALID, "bending arrays are NULL");
  return MC_OK;
}

}  // namespace

extern "C" {

const char* mc_version(void) { return MC_VERSION; }
const char* mc_last_error(void) { return g_err.c_str(); }

}  // extern "C"

namespace {

// The wide solve's cell counters and control block (shared by the plan's IB
// over nodes and over synthetic points), with the barrier-timeout record: a
// pinned host flag the wide solve sets, and the poll bound of its barrier
// waits (MC_IB_WIDE_POLL_LIMIT overrides the default 2^22, e.g. to exercise
// the error path).
mc_status alloc_wide_ctl(mc_plan* p) {
  IBWork& w = p->ib;
  if (w.wide_ctl) return MC_OK;
  MCCHK(p->alloc(&w.wide_count, (size_t)p->lat_cells, "wide cell counters"));
  MCCHK(p->alloc(&w.wide_ctl, 1, "wide control"));
  if (hipHostMalloc(reinterpret_cast<void**>(&p->host_err), sizeof(int)) != hipSuccess)
    return fail(MC_ERR_OOM, "pinned barrier-timeout flag");
  *p->host_err = 0;
  int limit = 1 << 22;
  if (const char* env = getenv("MC_IB_WIDE_POLL_LIMIT")) {
    const long v = atol(env);
    if (v >= 0 && v <= (1L << 30)) limit = (int)v;
  }
  int* dev_flag = nullptr;
  if (hipHostGetDevicePointer(reinterpret_cast<void**>(&dev_flag), p->host_err, 0) != hipSuccess)
    return fail(MC_ERR_HIP, "device address of the barrier-timeout flag");
  WideCtl init{};
  init.poll_limit = limit;
  init.host_err = dev_flag;
  return p->put(w.wide_ctl, &init, sizeof(init), "wide control");
}

// MC_IB_WIDE_WG (1..IB_WIDE_WG) caps the wide solve's workgroup count, for tuning.
int wide_wg_cap() {
  const char* env = getenv("MC_IB_WIDE_WG");
  const int v = env ? atoi(env) : 0;
  return (v >= 1 && v <= IB_WIDE_WG) ? v : IB_WIDE_WG;
}

// On the first call with synthetic points, sets up their IB work: the per-point arrays for 2N
// points, the point layout and the node -> edge table; the lattice-cell
// arrays, cluster tables and counters are the plan's own IB's (one IB runs
// at a time on a plan).  The wide solve launches whenever an instance holds
// more than 64 points (the smallest wide_min), as the tuning may lower it.
mc_status ensure_synth(mc_plan* p, const mc_params* q) {
  if (!(q->enable_fluid && q->synthetic_node_width > 0.0) || p->ib_syn_ready) return MC_OK;
  const IBWork& b = p->ib;
  IBWork w = b;
  const int N = p->N;
  const size_t P = 2 * (size_t)N;
  const size_t taps = (size_t)p->kmax * p->kmax;
  w.N = (int)P;
  w.nodes = N;
  int *pcomp, *poff, *pts, *ioff, *pinst;
  MCCHK(upload(p, &pcomp, p->syn_pcomp_h, "point components"));
  MCCHK(upload(p, &poff, p->syn_off_h, "component point offsets"));
  MCCHK(upload(p, &pts, p->syn_pts_h, "component points"));
  w.comp_of_node = pcomp;
  w.comp_node_off = poff;
  w.comp_nodes = pts;
  CompMeta* md;
  MCCHK(upload(p, &md, p->syn_meta_h, "component point records"));
  w.meta = md;
  MCCHK(upload(p, &ioff, p->syn_inc_off_h, "node edge offsets"));
  w.syn_inc_off = ioff;
  int* inc;
  MCCHK(upload(p, &inc, p->syn_inc_h, "node edges"));
  w.syn_inc = reinterpret_cast<const int2*>(inc);
  MCCHK(upload(p, &pinst, p->syn_pinst_h, "point instances"));
  w.point_inst = pinst;
  MCCHK(p->alloc(&w.syn_force, 2 * P, "point drag"));
  MCCHK(p->fill(w.syn_force, 0, 2 * P * sizeof(float), "point drag"));
  MCCHK(p->alloc(&w.st_sx, P, "point st sx"));
  MCCHK(p->alloc(&w.st_sy, P, "point st sy"));
  MCCHK(p->alloc(&w.st_w, taps * P, "point st w"));
  MCCHK(p->alloc(&w.st_won, P, "point st won"));
  MCCHK(p->alloc(&w.vib, 2 * P, "point vib"));
  MCCHK(p->alloc(&w.links, taps * P, "point cluster links"));
  {
    std::vector<int> loff(p->n_inst);
    long long acc = 0;
    for (int i = 0; i < p->n_inst; ++i) {
      loff[i] = (int)acc;
      acc += 2LL * p->inst[i].nodes * (long long)taps;
    }
    MCCHK(upload(p, &w.link_off, loff, "point link bucket offsets"));
  }
  MCCHK(p->alloc(&w.cl_scratch, w.cl_bpn * P, "point cluster scratch"));
  w.wide_wg = 0;
  for (const InstHost& h : p->inst)
    if (2LL * h.nodes > 64) w.wide_wg = b.wide_wg > 0 ? b.wide_wg : wide_wg_cap();
  w.wide_scratch = nullptr;
  if (w.wide_wg > 0) {
    MCCHK(alloc_wide_ctl(p));
    w.wide_count = p->ib.wide_count;
    w.wide_ctl = p->ib.wide_ctl;
    MCCHK(p->alloc(&w.wide_scratch, ib_wide_scratch_bytes((int)P, p->kmax), "point wide scratch"));
  }
  w.dirty = false;
  p->ib_syn = w;
  p->ib_syn_ready = true;
  return MC_OK;
}

mc_status create_plan(mc_plan** out, int32_t n_inst, const int32_t* shape, const int32_t* node_off,
                      const int32_t* pairs_host, int32_t num_edges,
                      const int32_t* bending_pairs_host, int32_t num_hinges,
                      int32_t max_kernel_size) {
  if (!out) return fail(MC_ERR_INVALID, "plan pointer is NULL");
  *out = nullptr;
  if (n_inst < 1) return fail(MC_ERR_BAD_SHAPE, "num_instances %d < 1", n_inst);
  if (!shape || !node_off) return fail(MC_ERR_INVALID, "instance arrays are NULL");
  for (int i = 0; i < n_inst; ++i) {
    const int* sh = shape + 4 * i;
    if (sh[0] < 2 || sh[1] < 2 || sh[2] < 2 || sh[3] < 2)
      return fail(MC_ERR_BAD_SHAPE, "instance %d: grid %dx%d / lattice %dx%d must be at least 2x2",
                  i, sh[0], sh[1], sh[2], sh[3]);
    if (node_off[i + 1] < node_off[i])
      return fail(MC_ERR_BAD_SHAPE, "node offsets descend at instance %d", i);
  }
  if (node_off[0] != 0) return fail(MC_ERR_BAD_SHAPE, "node offsets must start at 0");
  if (num_edges < 0 || num_hinges < 0) return fail(MC_ERR_BAD_SHAPE, "negative count");
  if ((num_edges > 0 && !pairs_host) || (num_hinges > 0 && !bending_pairs_host))
    return fail(MC_ERR_INVALID, "topology arrays are NULL");
  if (max_kernel_size < 1 || max_kernel_size > IB_KMAX)
    return fail(MC_ERR_UNSUPPORTED, "max_kernel_size %d (1..%d)", max_kernel_size, IB_KMAX);
  mc_plan* p = new mc_plan();
  HIPCHK(hipGetDevice(&p->dev));
  int lds = 0;
  if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, p->dev) == hipSuccess && lds > 0)
    p->lds_max = lds;
  p->n_inst = n_inst;
  for (int i = 0; i < n_inst; ++i)
    p->inst.push_back(InstHost{shape[4 * i], shape[4 * i + 1], shape[4 * i + 2], shape[4 * i + 3],
                               node_off[i], node_off[i + 1] - node_off[i]});
  p->N = node_off[n_inst];
  p->E = num_edges;
  p->B = num_hinges;
  p->kmax = max_kernel_size;
  auto bail = [&](mc_status st) {
    std::string keep = g_err;
    release_plan(p);
    g_err = keep;
    return st;
  };
  mc_status st;
  // the set-up's device work is queued and waited for once, at the end
  p->deferring = true;
  if ((st = build_instances(p)) != MC_OK) return bail(st);
  if ((st = build_layout(p, pairs_host, bending_pairs_host)) != MC_OK) return bail(st);
  if ((st = build_fft(p)) != MC_OK) return bail(st);
  // geometry without u_max (set by ensure_geo from the parameters of a call)
  if ((st = p->put(p->geo_d, p->geo_h.data(), p->geo_h.size() * sizeof(InstGeo),
                   "instance geometry")) != MC_OK)
    return bail(st);
  const int N = p->N;
  const size_t hwf = (size_t)p->lat_cells;
  if ((st = p->alloc(&p->pred_w, 2 * (size_t)N, "pred")) != MC_OK) return bail(st);
  if ((st = p->alloc(&p->node_tmp, 2 * (size_t)N, "node tmp")) != MC_OK) return bail(st);
  if ((st = p->alloc(&p->cic_cell, 4 * (size_t)N, "cic cell")) != MC_OK) return bail(st);
  if ((st = p->alloc(&p->cic_w, 4 * (size_t)N, "cic w")) != MC_OK) return bail(st);
  if ((st = p->alloc(&p->comp_rows, (size_t)p->L.num_comp, "comp rows")) != MC_OK) return bail(st);
  if ((st = p->alloc(&p->comp_cols, (size_t)p->L.num_comp, "comp cols")) != MC_OK) return bail(st);
  if ((st = p->alloc(&p->dep_arena, 2 * (size_t)N, "deposit arena")) != MC_OK) return bail(st);
  if ((st = p->alloc(&p->dep_top, 2, "deposit arena tops")) != MC_OK) return bail(st);
  if ((st = p->alloc(&p->row_sort, 2 * (size_t)N, "row sort")) != MC_OK) return bail(st);
  if ((st = p->alloc(&p->row_keys, 8 * (size_t)N, "row keys")) != MC_OK) return bail(st);
  // immersed boundary
  IBWork& w = p->ib;
  w.N = N;
  w.C = p->L.num_comp;
  w.comp_of_node = p->comp_of_node;
  w.comp_node_off = p->L.comp_node_off;
  w.comp_nodes = p->L.comp_nodes;
  w.meta = p->L.meta;
  w.cells = p->lat_cells;
  w.cl_bpn = ib_scratch_bytes_per_node(max_kernel_size);
  w.lds_bytes = ib_lds_bytes(p->lds_max);
  w.comp_order = nullptr;
  if (p->n_inst > 1 && p->L.num_comp > 0) {
    // the solve workgroups of the densest instances (stencil taps per lattice
    // cell) are dispatched first: their large clusters then start at once
    // instead of in the last wave of workgroups
    std::vector<int> io(p->n_inst);
    for (int i = 0; i < p->n_inst; ++i) io[i] = i;
    auto dens = [&](int i) {
      return (double)p->inst[i].nodes / ((double)p->inst[i].Hf * (double)p->inst[i].Wf);
    };
    std::stable_sort(io.begin(), io.end(), [&](int x, int y) { return dens(x) > dens(y); });
    std::vector<int> order;
    order.reserve(p->L.num_comp);
    for (int i : io)
      for (int c = 0; c < p->geo_h[i].ncomp; ++c) order.push_back(p->geo_h[i].comp0 + c);
    int* od;
    if ((st = upload(p, &od, order, "solve order")) != MC_OK) return bail(st);
    w.comp_order = od;
  }
  w.dirty = false;
  const size_t taps = (size_t)max_kernel_size * max_kernel_size;
  const size_t C = (size_t)w.C;
  if ((st = p->alloc(&w.st_sx, (size_t)N, "st sx")) != MC_OK) return bail(st);
  if ((st = p->alloc(&w.st_sy, (size_t)N, "st sy")) != MC_OK) return bail(st);
  if ((st = p->alloc(&w.st_w, taps * N, "st w")) != MC_OK) return bail(st);
  if ((st = p->alloc(&w.st_won, (size_t)N, "st won")) != MC_OK) return bail(st);
  if ((st = p->alloc(&w.vib, 2 * (size_t)N, "vib")) != MC_OK) return bail(st);
  if ((st = p->alloc(&w.owner, hwf, "cell owner")) != MC_OK) return bail(st);
  if ((st = p->alloc(&w.cell_seg, hwf, "cell seg")) != MC_OK) return bail(st);
  if ((st = p->alloc(&w.links, taps * N, "cluster links")) != MC_OK) return bail(st);
  for (const InstHost& h : p->inst)
    if (h.Hf > 32767 || h.Wf > 32767)
      return bail(fail(MC_ERR_BAD_SHAPE, "lattice side above 32767 (%d x %d)", h.Hf, h.Wf));
  // link buckets per instance when every instance's components fit the
  // per-workgroup union-find
  w.bucketed = true;
  for (const InstGeo& gh : p->geo_h) w.bucketed = w.bucketed && gh.ncomp <= IB_UF_MAX;
  w.nbuckets = w.bucketed ? p->n_inst : 1;
  if ((st = p->alloc(&w.nlinks, (size_t)w.nbuckets, "cluster link counts")) != MC_OK) return bail(st);
  {
    std::vector<int> loff(p->n_inst);
    long long acc = 0;
    for (int i = 0; i < p->n_inst; ++i) {
      loff[i] = (int)acc;
      acc += (long long)p->inst[i].nodes * (long long)taps;
    }
    if ((st = upload(p, &w.link_off, loff, "link bucket offsets")) != MC_OK) return bail(st);
  }
  if ((st = p->alloc(&w.ictl, 1, "solve counters")) != MC_OK) return bail(st);
  if ((st = p->alloc(&w.label, C, "cluster label")) != MC_OK) return bail(st);
  if ((st = p->alloc(&w.cl_cnt, C, "cluster nodes")) != MC_OK) return bail(st);
  if ((st = p->alloc(&w.cl_node_off, C, "cluster node off")) != MC_OK) return bail(st);
  if ((st = p->alloc(&w.cl_mem_cnt, C, "cluster members")) != MC_OK) return bail(st);
  if ((st = p->alloc(&w.cl_mem_off, C, "cluster member off")) != MC_OK) return bail(st);
  if ((st = p->alloc(&w.cl_members, C, "cluster member list")) != MC_OK) return bail(st);
  if ((st = p->alloc(&w.cl_scratch, w.cl_bpn * N, "cluster scratch")) != MC_OK) return bail(st);
  {
    std::vector<int> ident(C);
    for (size_t i = 0; i < C; ++i) ident[i] = (int)i;
    if ((st = upload(p, &w.uf_parent, ident, "union-find")) != MC_OK) return bail(st);
    if ((st = p->fill(w.owner, 0xff, hwf * sizeof(int), "cell owner")) != MC_OK) return bail(st);
  }
  if (ib_prepare(p->lds_max) != hipSuccess)
    return bail(fail(MC_ERR_HIP, "cluster solve: %zu B of dynamic LDS refused", w.lds_bytes));
  if (rod_prepare(p->lds_max) != hipSuccess)
    return bail(fail(MC_ERR_HIP, "rod kernels: %d B of dynamic LDS refused", p->lds_max - 1024));
  if (steric_prepare(p->lds_max) != hipSuccess)
    return bail(fail(MC_ERR_HIP, "steric kernels: %d B of dynamic LDS refused", p->lds_max - 1024));
  // the wide solve exists when an instance can hold a cluster of more than
  // IB_WIDE_MIN nodes
  w.wide_wg = 0;
  w.wide_min = IB_WIDE_MIN;
  {
    // MC_IB_WIDE_MIN overrides the cluster size above which the wide solve
    // takes over (64..IB_CLUSTER_MAX), for tuning
    const char* env = getenv("MC_IB_WIDE_MIN");
    const int v = env ? atoi(env) : 0;
    if (v >= 64 && v <= IB_CLUSTER_MAX) w.wide_min = v;
  }
  for (const InstHost& h : p->inst)
    if (h.nodes > w.wide_min) w.wide_wg = IB_WIDE_WG;
  if (w.wide_wg > 0) {
    // MC_IB_WIDE_WG (1..IB_WIDE_WG) caps the workgroup count, for tuning
    const char* env = getenv("MC_IB_WIDE_WG");
    const int v = env ? atoi(env) : 0;
    if (v >= 1 && v <= IB_WIDE_WG) w.wide_wg = v;
  }
  if (w.wide_wg > 0) {
    unsigned char* ws = nullptr;
    if ((st = p->alloc(&ws, ib_wide_scratch_bytes(N, max_kernel_size), "wide scratch")) != MC_OK)
      return bail(st);
    w.wide_scratch = ws;
    if ((st = alloc_wide_ctl(p)) != MC_OK) return bail(st);
  }
  if ((st = p->alloc(&w.force, 2 * hwf, "force U")) != MC_OK) return bail(st);
  if ((st = p->alloc(&w.mask, (size_t)p->lat_cells, "U flags")) != MC_OK) return bail(st);
  if ((st = p->alloc(&p->f_tmp, 9 * hwf, "f ping-pong")) != MC_OK) return bail(st);
  if ((st = p->alloc(&p->tick, 1, "tick")) != MC_OK) return bail(st);
  p->deferring = false;
  if ((st = p->settle()) != MC_OK) return bail(st);
  *out = p;
  return MC_OK;
}

// A wide IB solve of work enqueued earlier gave up at a grid barrier (its
// workgroups were not all resident): report it once, and restore the idle IB
// state on `s` ahead of the caller's next work (the device flag stays set,
// and the wide solve does nothing but wrap positions, until then).
mc_status check_async(mc_plan* p, hipStream_t s) {
  if (!p->host_err || __atomic_load_n(p->host_err, __ATOMIC_ACQUIRE) == 0) return MC_OK;
  __atomic_store_n(p->host_err, 0, __ATOMIC_RELEASE);
  hipError_t e = launch_ib_reset(p->ib, s);
  if (e == hipSuccess && p->ib_syn_ready) e = launch_ib_reset(p->ib_syn, s);
  if (e != hipSuccess) return fail(MC_ERR_HIP, "IB reset after a barrier timeout: %s", hipGetErrorString(e));
  return fail(MC_ERR_HIP,
              "immersed-boundary wide solve: a grid barrier wait gave up (its workgroups were not "
              "all resident on the device); node velocities and IB forces of the steps enqueued "
              "since the last successful call on this plan are invalid");
}

}  // namespace

extern "C" {

mc_status mc_plan_status(mc_plan* p, void* stream) {
  g_err.clear();
  if (!p) return fail(MC_ERR_INVALID, "plan is NULL");
  HIPCHK(hipStreamSynchronize((hipStream_t)stream));
  return check_async(p, (hipStream_t)stream);
}

mc_status mc_plan_create(mc_plan** out, int32_t grid_h, int32_t grid_w, int32_t fluid_h,
                         int32_t fluid_w, int32_t num_nodes, const int32_t* pairs_host,
                         int32_t num_edges, const int32_t* bending_pairs_host, int32_t num_hinges,
                         int32_t max_kernel_size) {
  g_err.clear();
  if (!out) return fail(MC_ERR_INVALID, "plan pointer is NULL");
  if (num_nodes < 0) return fail(MC_ERR_BAD_SHAPE, "negative count");
  const int32_t shape[4] = {grid_h, grid_w, fluid_h, fluid_w};
  const int32_t off[2] = {0, num_nodes};
  return create_plan(out, 1, shape, off, pairs_host, num_edges, bending_pairs_host, num_hinges,
                     max_kernel_size);
}

mc_status mc_batch_plan_create(mc_plan** out, int32_t num_instances, const int32_t* instance_shapes,
                               const int32_t* node_offsets, const int32_t* pairs_host,
                               int32_t num_edges, const int32_t* bending_pairs_host,
                               int32_t num_hinges, int32_t max_kernel_size) {
  g_err.clear();
  return create_plan(out, num_instances, instance_shapes, node_offsets, pairs_host, num_edges,
                     bending_pairs_host, num_hinges, max_kernel_size);
}

mc_status mc_plan_destroy(mc_plan* p) {
  if (!p) return MC_OK;
  // work using the plan may still be queued on any stream
  (void)hipDeviceSynchronize();
  release_plan(p);
  return MC_OK;
}

mc_status mc_plan_destroy_after(mc_plan* p, void* stream) {
  g_err.clear();
  if (!p) return MC_OK;
  // everything that used the plan was enqueued on `stream`: wait for that
  // stream only, then hand the memory back to the pool
  hipEvent_t ev;
  HIPCHK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  hipError_t e = hipEventRecord(ev, (hipStream_t)stream);
  if (e == hipSuccess) e = hipEventSynchronize(ev);
  (void)hipEventDestroy(ev);
  if (e != hipSuccess) return fail(MC_ERR_HIP, "waiting for the plan's stream: %s", hipGetErrorString(e));
  const bool timed_out = p->host_err && __atomic_load_n(p->host_err, __ATOMIC_ACQUIRE) != 0;
  release_plan(p);
  if (timed_out)
    return fail(MC_ERR_HIP,
                "immersed-boundary wide solve: a grid barrier wait gave up (its workgroups were "
                "not all resident on the device); the plan's last results are invalid");
  return MC_OK;
}

}  // extern "C"

namespace {

mc_status drop_graph(mc_plan* p, int slot);
mc_status rollout_impl(mc_plan* p, const mc_params* q, const mc_state* st, const mc_edge_params* ed,
                       const mc_actuation* act, const long long* t_inst, double dt,
                       int64_t num_steps, float* rec_pos, float* rec_vel, int32_t use_graph,
                       void* stream, const mc_actions* rows = nullptr,
                       const mc_record* frames = nullptr);

// The tuned knobs of plans with the same instances and physics, shared by
// every plan of the process (a re-formed batch of a known shape reuses them).
struct TuneChoice {
  int wide_min;
  bool rod_fuse;
  bool lbm_side;
};
std::mutex g_tune_mu;
std::map<std::string, TuneChoice> g_tune;

std::string tune_key(const mc_plan* p, const mc_params* q, double dt) {
  std::string k;
  auto add = [&](long long v) {
    k += std::to_string(v);
    k += ',';
  };
  add(p->dev);
  add(p->n_inst);
  for (const InstHost& h : p->inst) {
    add(h.H);
    add(h.W);
    add(h.Hf);
    add(h.Wf);
    add(h.nodes);
  }
  add(p->E);
  add(p->B);
  add(q->ibm_kernel_size);
  add(q->ibm_iterations);
  add(q->cycles_per_step);
  add(q->enable_fluid);
  add(q->enable_steric);
  add(q->solver_type);
  add(q->bending_constraint);
  add((long long)(dt * 1e9));
  return k;
}

// `steps` coupled steps of the rollout graphs' shape on `s`, eagerly: the
// LBM pass of each step inside the next step's rod launch when fused.
mc_status enqueue_steps(mc_plan* p, const mc_params* q, const Consts& c, const mc_state& sa,
                        const mc_state& sb, const mc_edge_params* ed, int steps, hipStream_t s) {
  LbmArgs pend[2];
  const LbmArgs* fuse = nullptr;
  const bool side_mode = p->lbm_side && q->enable_fluid;
  if (side_mode) MCCHK(ensure_side(p));
  for (int k = 0; k < steps; ++k) {
    const int par = k & 1;
    LbmArgs* defer =
        (q->enable_fluid && (p->rod_fuse || side_mode) && k + 1 < steps) ? &pend[par] : nullptr;
    SideLbm sd;
    if (side_mode && fuse) {
      sd.lbm = fuse;
      sd.stream = p->side_stream;
      sd.fork = p->side_fork[k % ROLLOUT_GRAPH_STEPS];
      sd.join = p->side_join[k % ROLLOUT_GRAPH_STEPS];
    }
    const bool ch = chain_predict(p, c);
    MCCHK(enqueue_step(p, q, c, par ? &sb : &sa, par ? &sa : &sb, ed, nullptr, s, StepSync(),
                       side_mode ? nullptr : fuse, defer, sd, ch && k > 0, ch && k + 1 < steps));
    fuse = defer;
  }
  return MC_OK;
}

}  // namespace

extern "C" {

mc_status mc_plan_autotune(mc_plan* p, const mc_params* q, const mc_state* st,
                           const mc_edge_params* ed, double dt, int32_t steps, void* stream) {
  g_err.clear();
  if (!p) return fail(MC_ERR_INVALID, "plan is NULL");
  MCCHK(check_params(p, q, true, true));
  MCCHK(ensure_synth(p, q));
  MCCHK(check_state(p, st, "state"));
  MCCHK(check_edges(p, ed));
  if (steps < 1) return fail(MC_ERR_BAD_SHAPE, "steps < 1");
  if (!(dt > 0.0)) return fail(MC_ERR_BAD_SHAPE, "dt must be positive");
  hipStream_t s = (hipStream_t)stream;
  const std::string key = tune_key(p, q, dt);
  {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    auto it = g_tune.find(key);
    if (it != g_tune.end()) {
      const int wm = p->ib.wide_wg > 0 ? it->second.wide_min : p->ib.wide_min;
      const bool fz = it->second.rod_fuse && p->rod_fuse_ok;
      const bool sd = it->second.lbm_side;
      const bool changed = wm != p->ib.wide_min || fz != p->rod_fuse || sd != p->lbm_side;
      p->ib.wide_min = wm;
      p->rod_fuse = fz;
      p->lbm_side = sd;
      p->tuned = true;
      if (changed)  // the graphs captured with the old knobs are stale
        for (int slot = 0; slot < 2; ++slot) MCCHK(drop_graph(p, slot));
      return MC_OK;
    }
  }
  MCCHK(check_async(p, s));
  const Consts c = derive(p, q, dt);
  MCCHK(ensure_geo(p, q));
  if (q->enable_steric) MCCHK(ensure_kf(p, q->steric_sigma, s));
  // a scratch copy of the state: the candidates step it, not the caller's
  const size_t N2 = 2 * (size_t)p->N, E = (size_t)p->E, G = (size_t)p->grid_cells,
               L = (size_t)p->lat_cells;
  const size_t words = 3 * N2 + E + G + 2 * L + 18 * L + 8;
  float* scratch = nullptr;
  HIPCHK(hipMallocAsync(reinterpret_cast<void**>(&scratch), words * sizeof(float), s));
  mc_state sa;
  sa.position = scratch;
  sa.velocity = sa.position + N2;
  sa.debug_vector = sa.velocity + N2;
  sa.theta = sa.debug_vector + N2;
  sa.steric = sa.theta + E;
  sa.fluid_velocity = sa.steric + G;
  sa.f_grid = sa.fluid_velocity + 2 * L;
  mc_state sb = sa;
  sb.f_grid = sa.f_grid + 9 * L;
  auto copy = [&](float* dst, const float* src, size_t n) -> mc_status {
    if (n && src) HIPCHK(hipMemcpyAsync(dst, src, n * sizeof(float), hipMemcpyDeviceToDevice, s));
    return MC_OK;
  };
  mc_status r = MC_OK;
  if (r == MC_OK) r = copy(sa.position, st->position, N2);
  if (r == MC_OK) r = copy(sa.velocity, st->velocity, N2);
  if (r == MC_OK) r = copy(sa.debug_vector, st->debug_vector, N2);
  if (r == MC_OK) r = copy(sa.theta, st->theta, E);
  if (r == MC_OK) r = copy(sa.steric, st->steric, G);
  if (r == MC_OK) r = copy(sa.fluid_velocity, st->fluid_velocity, 2 * L);
  if (r == MC_OK) r = copy(sa.f_grid, st->f_grid, 9 * L);
  // the candidates: the cluster size above which the wide solve takes over
  // (when the plan has one) and the LBM pass inside the rod launch or not;
  // every choice computes the same result bit for bit
  const TrackWork keep_track = p->track;
  p->track.n_groups = 0;  // the scratch steps feed no tracker
  // (the current choice first: another must beat it by 5 % to replace it).
  // Sizes that no cluster of the scratch steps exceeds change nothing there
  // and are not timed: the largest cluster is read after two steps.
  std::vector<int> wides = {p->ib.wide_min};
  if (p->ib.wide_wg > 0 && q->enable_fluid && r == MC_OK) {
    r = enqueue_steps(p, q, c, sa, sb, ed, 2, s);
    std::vector<int> cnt(std::max(p->ib.C, 1));
    if (r == MC_OK && p->ib.C > 0 &&
        (hipMemcpyAsync(cnt.data(), p->ib.cl_cnt, (size_t)p->ib.C * sizeof(int),
                        hipMemcpyDeviceToHost, s) != hipSuccess ||
         hipStreamSynchronize(s) != hipSuccess))
      r = fail(MC_ERR_HIP, "cluster sizes");
    int largest = 0;
    for (int v : cnt) largest = std::max(largest, v);
    // (sizes below the default put clusters a long rollout keeps forming on
    // the slower multi-workgroup path: only the default and the largest the
    // per-cluster solve takes are tried)
    for (int v : {IB_WIDE_MIN, IB_CLUSTER_MAX})
      if (v >= 64 && v <= IB_CLUSTER_MAX && v != p->ib.wide_min &&
          (v < largest || p->ib.wide_min < largest))
        wides.push_back(v);
  }
  // the LBM pass: 1 inside the next rod launch, 2 on a second stream beside
  // the next step's steric and rod launches, 0 a launch after the step's IB
  const int cur_mode = p->lbm_side ? 2 : (p->rod_fuse ? 1 : 0);
  std::vector<int> modes = {cur_mode};
  if (q->enable_fluid)
    for (int m : {1, 2, 0})
      if (m != cur_mode && (m != 1 || p->rod_fuse_ok)) modes.push_back(m);
  const int keep_wide = p->ib.wide_min;
  const bool keep_fuse = p->rod_fuse, keep_side = p->lbm_side;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  float best = 0.0f;
  TuneChoice choice{keep_wide, keep_fuse, keep_side};
  if (r == MC_OK && hipEventCreate(&e0) != hipSuccess) r = fail(MC_ERR_HIP, "event");
  if (r == MC_OK && hipEventCreate(&e1) != hipSuccess) r = fail(MC_ERR_HIP, "event");
  for (int wm : wides) {
    for (int mode : modes) {
      if (r != MC_OK) break;
      const bool fz = mode == 1, sd = mode == 2;
      p->ib.wide_min = wm;
      p->rod_fuse = fz;
      p->lbm_side = sd;
      // graph-replayed rollouts of the scratch state, as mc_rollout runs
      // them: one to capture the graphs, then the timed ones twice; the
      // faster repeat
      for (int slot = 0; slot < 2 && r == MC_OK; ++slot) r = drop_graph(p, slot);
      const int64_t len = (int64_t)ROLLOUT_GRAPH_STEPS *
                          std::max<int64_t>(1, (steps + ROLLOUT_GRAPH_STEPS - 1) / ROLLOUT_GRAPH_STEPS);
      float ms = 0.0f;
      if (r == MC_OK)
        r = rollout_impl(p, q, &sa, ed, nullptr, nullptr, dt, ROLLOUT_GRAPH_STEPS, nullptr,
                         nullptr, 1, s);
      for (int rep = 0; rep < 2 && r == MC_OK; ++rep) {
        if (hipEventRecord(e0, s) != hipSuccess) r = fail(MC_ERR_HIP, "event record");
        if (r == MC_OK)
          r = rollout_impl(p, q, &sa, ed, nullptr, nullptr, dt, len, nullptr, nullptr, 1, s);
        if (r == MC_OK && hipEventRecord(e1, s) != hipSuccess) r = fail(MC_ERR_HIP, "event record");
        float t = 0.0f;
        if (r == MC_OK && (hipEventSynchronize(e1) != hipSuccess ||
                           hipEventElapsedTime(&t, e0, e1) != hipSuccess))
          r = fail(MC_ERR_HIP, "timing");
        if (r == MC_OK) ms = rep == 0 ? t : std::min(ms, t);
      }
      if (r == MC_OK) r = check_async(p, s);
      if (r == MC_OK && (best == 0.0f || ms < 0.95f * best)) {
        best = ms;
        choice = TuneChoice{wm, fz, sd};
      }
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  p->track = keep_track;
  // the graphs captured on the scratch state go with it
  for (int slot = 0; slot < 2; ++slot) {
    const mc_status d = drop_graph(p, slot);
    if (r == MC_OK) r = d;
  }
  (void)hipFreeAsync(scratch, s);
  if (r != MC_OK) {
    p->ib.wide_min = keep_wide;
    p->rod_fuse = keep_fuse;
    p->lbm_side = keep_side;
    return r;
  }
  p->ib.wide_min = choice.wide_min;
  p->rod_fuse = choice.rod_fuse;
  p->lbm_side = choice.lbm_side;
  p->tuned = true;
  // the graphs captured with the old knobs are stale
  for (int slot = 0; slot < 2; ++slot) MCCHK(drop_graph(p, slot));
  std::lock_guard<std::mutex> lock(g_tune_mu);
  g_tune[key] = choice;
  return MC_OK;
}

mc_status mc_plan_set_tuning(mc_plan* p, int32_t wide_min, int32_t lbm_mode) {
  g_err.clear();
  if (!p) return fail(MC_ERR_INVALID, "plan is NULL");
  if (lbm_mode < 0 || lbm_mode > 2) return fail(MC_ERR_INVALID, "lbm_mode %d (0, 1 or 2)", lbm_mode);
  if (lbm_mode == 1 && !p->rod_fuse_ok)
    return fail(MC_ERR_UNSUPPORTED, "the rod launch of this plan has no room for the LBM pass");
  if (wide_min < 64 || wide_min > IB_CLUSTER_MAX)
    return fail(MC_ERR_BAD_SHAPE, "wide_min %d (64 .. %d)", wide_min, IB_CLUSTER_MAX);
  // without the wide solve every cluster is solved per workgroup: the size
  // stays above every instance's node count
  const int wm = p->ib.wide_wg > 0 ? wide_min : p->ib.wide_min;
  const bool fz = lbm_mode == 1, sd = lbm_mode == 2;
  if (wm != p->ib.wide_min || fz != p->rod_fuse || sd != p->lbm_side)
    for (int slot = 0; slot < 2; ++slot) MCCHK(drop_graph(p, slot));
  p->ib.wide_min = wm;
  p->rod_fuse = fz;
  p->lbm_side = sd;
  return MC_OK;
}

mc_status mc_plan_tuning(const mc_plan* p, int32_t* wide_min, int32_t* lbm_mode) {
  g_err.clear();
  if (!p || !wide_min || !lbm_mode) return fail(MC_ERR_INVALID, "NULL argument");
  *wide_min = p->ib.wide_min;
  *lbm_mode = p->lbm_side ? 2 : (p->rod_fuse ? 1 : 0);
  return MC_OK;
}

mc_status mc_plan_workspace_bytes(const mc_plan* p, size_t* bytes) {
  if (!p || !bytes) return fail(MC_ERR_INVALID, "NULL argument");
  *bytes = p->bytes;
  return MC_OK;
}

mc_status mc_step(mc_plan* p, const mc_params* q, const mc_state* in, const mc_state* out,
                  const mc_edge_params* ed, double dt, void* stream) {
  g_err.clear();
  if (!p) return fail(MC_ERR_INVALID, "plan is NULL");
  MCCHK(check_params(p, q, true, true));
  MCCHK(ensure_synth(p, q));
  MCCHK(check_state(p, in, "input"));
  MCCHK(check_state(p, out, "output"));
  MCCHK(check_edges(p, ed));
  if (q->enable_fluid && in->f_grid == out->f_grid)
    return fail(MC_ERR_INVALID, "f_grid of out must differ from in (streaming is out of place)");
  if (!(dt > 0.0)) return fail(MC_ERR_BAD_SHAPE, "dt must be positive");
  hipStream_t s = (hipStream_t)stream;
  MCCHK(check_async(p, s));
  const Consts c = derive(p, q, dt);
  MCCHK(ensure_geo(p, q));
  if (q->enable_steric) MCCHK(ensure_kf(p, q->steric_sigma, s));
  return enqueue_step(p, q, c, in, out, ed, nullptr, s);
}

}  // extern "C"

namespace {

// Drops the cached rollout graph after its last launch has finished.
mc_status drop_graph(mc_plan* p, int slot) {
  if (!p->graph_exec[slot]) return MC_OK;
  if (p->graph_done) HIPCHK(hipEventSynchronize(p->graph_done));
  HIPCHK(hipGraphExecDestroy(p->graph_exec[slot]));
  p->graph_exec[slot] = nullptr;
  return MC_OK;
}

// mc_rollout with optional per-instance actuation offsets t_inst (device).
// The recording taps of a single-instance plan (built on first use).
mc_status ensure_record(mc_plan* p) {
  if (p->rec_tick) return MC_OK;
  const InstHost& h = p->inst[0];
  std::vector<int> yo, yi, xo, xi;
  std::vector<float> yw, xw;
  resize_taps(h.Hf, h.H, yo, yi, yw);
  resize_taps(h.Wf, h.W, xo, xi, xw);
  p->rec_ny = (int)yi.size();
  p->rec_nx = (int)xi.size();
  std::vector<int> idx;
  idx.insert(idx.end(), yo.begin(), yo.end());
  idx.insert(idx.end(), yi.begin(), yi.end());
  idx.insert(idx.end(), xo.begin(), xo.end());
  idx.insert(idx.end(), xi.begin(), xi.end());
  std::vector<float> w(yw);
  w.insert(w.end(), xw.begin(), xw.end());
  long long* t = nullptr;
  MCCHK(p->alloc(&p->rec_idx, idx.size(), "resize taps"));
  MCCHK(p->alloc(&p->rec_w, w.size(), "resize weights"));
  HIPCHK(hipMemcpy(p->rec_idx, idx.data(), idx.size() * sizeof(int), hipMemcpyHostToDevice));
  if (!w.empty())
    HIPCHK(hipMemcpy(p->rec_w, w.data(), w.size() * sizeof(float), hipMemcpyHostToDevice));
  MCCHK(p->alloc(&t, 1, "frame counter"));
  p->rec_tick = t;
  return MC_OK;
}

RecordDev record_dev(const mc_plan* p, const mc_record* r, const mc_state* st) {
  const InstHost& h = p->inst[0];
  RecordDev d;
  memset(&d, 0, sizeof(d));
  d.pos = r->position;
  d.vel = r->velocity;
  d.dbg = r->debug_vector;
  d.steric = r->steric;
  d.fv = r->fluid_velocity;
  d.s_pos = st->position;
  d.s_vel = st->velocity;
  d.s_dbg = st->debug_vector;
  d.s_steric = st->steric;
  d.s_fv = st->fluid_velocity;
  d.counter = p->rec_tick;
  d.ry_off = p->rec_idx;
  d.ry_idx = d.ry_off + (h.H + 1);
  d.rx_off = d.ry_idx + p->rec_ny;
  d.rx_idx = d.rx_off + (h.W + 1);
  d.ry_w = p->rec_w;
  d.rx_w = p->rec_w + p->rec_ny;
  d.stride = (long long)r->stride;
  d.N = p->N;
  d.H = h.H;
  d.W = h.W;
  d.Hf = h.Hf;
  d.Wf = h.Wf;
  d.resize = (h.H != h.Hf || h.W != h.Wf) ? 1 : 0;
  return d;
}

mc_status rollout_impl(mc_plan* p, const mc_params* q, const mc_state* st, const mc_edge_params* ed,
                       const mc_actuation* act, const long long* t_inst, double dt,
                       int64_t num_steps, float* rec_pos, float* rec_vel, int32_t use_graph,
                       void* stream, const mc_actions* rows, const mc_record* frames) {
  if (!p) return fail(MC_ERR_INVALID, "plan is NULL");
  MCCHK(check_params(p, q, true, true));
  MCCHK(ensure_synth(p, q));
  MCCHK(check_state(p, st, "state"));
  MCCHK(check_edges(p, ed));
  if (num_steps < 0) return fail(MC_ERR_BAD_SHAPE, "num_steps < 0");
  if (!(dt > 0.0)) return fail(MC_ERR_BAD_SHAPE, "dt must be positive");
  if (act && p->B > 0 && (!act->amp || !act->omega || !act->phase))
    return fail(MC_ERR_INVALID, "actuation arrays are NULL");
  if (act && rows) return fail(MC_ERR_INVALID, "actuation and action rows are exclusive");
  if (rows && rows->num_rows < num_steps)
    return fail(MC_ERR_BAD_SHAPE, "%lld action rows for %lld steps", (long long)rows->num_rows,
                (long long)num_steps);
  MCCHK(check_async(p, (hipStream_t)stream));
  if (num_steps == 0) return MC_OK;
  hipStream_t us = (hipStream_t)stream;
  const Consts c = derive(p, q, dt);
  const size_t hwf = (size_t)p->lat_cells;
  MCCHK(ensure_geo(p, q));
  if (q->enable_steric) MCCHK(ensure_kf(p, q->steric_sigma, us));
  Actuation A{};
  const Actuation* ap = nullptr;
  if (act) {
    A = Actuation{act->amp, act->omega, act->phase, p->tick, t_inst};
    ap = &A;
    HIPCHK(launch_set_tick(p->tick, (long long)act->t0, us));
  } else if (rows) {
    // step k of this call reads row k: the counter starts at 0
    A = Actuation{nullptr, nullptr, nullptr, p->tick, nullptr, rows->d_rest_length,
                  rows->d_bending_angle, (long long)p->E, (long long)p->B, 1, 0};
    ap = &A;
    HIPCHK(launch_set_tick(p->tick, 0LL, us));
  }
  RecordDev R;
  memset(&R, 0, sizeof(R));
  if (frames) {
    MCCHK(ensure_record(p));
    R = record_dev(p, frames, st);
    HIPCHK(launch_set_tick(p->rec_tick, 0LL, us));
  }
  auto frame = [&](hipStream_t s) -> mc_status {
    if (!frames) return MC_OK;
    HIPCHK(launch_record(R, s));
    HIPCHK(launch_tick(p->rec_tick, s));
    return MC_OK;
  };
  // ping-pong: even steps read state->f_grid, odd steps read f_tmp
  mc_state sa = *st, sb = *st;
  sb.f_grid = p->f_tmp;
  auto record = [&](int64_t k, hipStream_t s) -> mc_status {
    if (rec_pos && p->N > 0)
      HIPCHK(hipMemcpyAsync(rec_pos + (size_t)k * 2 * p->N, st->position, 2 * (size_t)p->N * sizeof(float),
                            hipMemcpyDeviceToDevice, s));
    if (rec_vel)  // both parities write u* into st->fluid_velocity
      HIPCHK(hipMemcpyAsync(rec_vel + (size_t)k * 2 * hwf, st->fluid_velocity, 2 * hwf * sizeof(float),
                            hipMemcpyDeviceToDevice, s));
    return MC_OK;
  };
  const bool recording = rec_pos || rec_vel;
  if (!use_graph || recording) {
    // step by step on the caller's stream (records between steps)
    const bool ch = chain_predict(p, c);
    for (int64_t k = 0; k < num_steps; ++k) {
      const int par = (int)(k & 1);
      MCCHK(enqueue_step(p, q, c, par ? &sb : &sa, par ? &sa : &sb, ed, ap, us, StepSync(),
                         nullptr, nullptr, SideLbm(), ch && k > 0, ch && k + 1 < num_steps));
      MCCHK(record(k, us));
      MCCHK(frame(us));
    }
  } else {
    // The bulk of the rollout replays one graph of ROLLOUT_GRAPH_STEPS steps
    // (even, so every replay starts on the same lattice buffer), the
    // remainder one graph of its own length; inside a graph the LBM pass of
    // step k runs inside step k+1's rod launch and the last step's LBM pass
    // runs on its own.  Both instantiated graphs are kept and replayed by
    // later calls with the same arguments; they are captured on a private
    // stream and launched on the caller's.
    const int64_t S = ROLLOUT_GRAPH_STEPS;
    const int64_t bulk = (num_steps / S) * S, tail = num_steps - bulk;
    auto graph_for = [&](int slot, int64_t len) -> mc_status {
      GraphKey key;
      memset(&key, 0, sizeof(key));
      key.q = *q;
      key.dt = dt;
      key.sa = sa;
      key.sb = sb;
      key.ed = *ed;
      if (ap) key.act = A;
      key.has_act = ap != nullptr;
      key.has_rec = frames != nullptr;
      key.rec = R;
      key.steps = len;
      key.track = p->track;
      if (p->graph_exec[slot] && memcmp(&key, &p->graph_key[slot], sizeof(key)) != 0)
        MCCHK(drop_graph(p, slot));
      if (p->graph_exec[slot]) return MC_OK;
      if (!p->cap_stream) HIPCHK(hipStreamCreateWithFlags(&p->cap_stream, hipStreamNonBlocking));
      if (!p->graph_done) HIPCHK(hipEventCreateWithFlags(&p->graph_done, hipEventDisableTiming));
      hipStream_t cs = p->cap_stream;
      hipGraph_t g;
      mc_status r = MC_OK;
      HIPCHK(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
      LbmArgs pend[2];
      const LbmArgs* fuse = nullptr;
      const bool side_mode = p->lbm_side && q->enable_fluid && !frames;
      if (side_mode) r = ensure_side(p);
      for (int64_t k = 0; k < len && r == MC_OK; ++k) {
        const int par = (int)(k & 1);
        // a recorded frame needs the step's lattice pass done: no deferral
        LbmArgs* defer = (q->enable_fluid && (p->rod_fuse || side_mode) && k + 1 < len && !frames)
                             ? &pend[par]
                             : nullptr;
        SideLbm sd;
        if (side_mode && fuse) {
          sd.lbm = fuse;
          sd.stream = p->side_stream;
          sd.fork = p->side_fork[k % ROLLOUT_GRAPH_STEPS];
          sd.join = p->side_join[k % ROLLOUT_GRAPH_STEPS];
        }
        const bool ch = chain_predict(p, c);
        r = enqueue_step(p, q, c, par ? &sb : &sa, par ? &sa : &sb, ed, ap, cs, StepSync(),
                         side_mode ? nullptr : fuse, defer, sd, ch && k > 0, ch && k + 1 < len);
        if (r == MC_OK) r = frame(cs);
        fuse = defer;
      }
      hipError_t ec = hipStreamEndCapture(cs, &g);
      if (r != MC_OK) return r;
      if (ec != hipSuccess) return fail(MC_ERR_HIP, "stream capture failed: %s", hipGetErrorString(ec));
      hipGraphExec_t ge = nullptr;
      ec = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (ec != hipSuccess) return fail(MC_ERR_HIP, "graph instantiation failed: %s", hipGetErrorString(ec));
      p->graph_exec[slot] = ge;
      p->graph_key[slot] = key;
      return MC_OK;
    };
    if (bulk > 0) {
      MCCHK(graph_for(0, S));
      for (int64_t k = 0; k < bulk; k += S) HIPCHK(hipGraphLaunch(p->graph_exec[0], us));
    }
    if (tail > 0) {
      // the remainder starts on an even step (bulk is even): same parity as step 0
      MCCHK(graph_for(1, tail));
      HIPCHK(hipGraphLaunch(p->graph_exec[1], us));
    }
    HIPCHK(hipEventRecord(p->graph_done, us));
  }
  if (num_steps & 1) {
    // final lattice is in f_tmp; bring it back into the caller's buffer
    if (q->enable_fluid)
      HIPCHK(hipMemcpyAsync(st->f_grid, p->f_tmp, 9 * hwf * sizeof(float), hipMemcpyDeviceToDevice, us));
  }
  return MC_OK;
}

}  // namespace

extern "C" {

mc_status mc_rollout(mc_plan* p, const mc_params* q, const mc_state* st, const mc_edge_params* ed,
                     const mc_actuation* act, double dt, int64_t num_steps, float* rec_pos,
                     float* rec_vel, int32_t use_graph, void* stream) {
  g_err.clear();
  return rollout_impl(p, q, st, ed, act, nullptr, dt, num_steps, rec_pos, rec_vel, use_graph,
                      stream);
}

mc_status mc_profile_step(mc_plan* p, const mc_params* q, const mc_state* st,
                          const mc_edge_params* ed, const mc_actuation* act, double dt,
                          int64_t num_steps, double* lbm_ms_avg, double* step_ms_avg, void* stream) {
  g_err.clear();
  if (!p || !lbm_ms_avg || !step_ms_avg) return fail(MC_ERR_INVALID, "NULL argument");
  MCCHK(check_params(p, q, true, true));
  MCCHK(ensure_synth(p, q));
  MCCHK(check_state(p, st, "state"));
  MCCHK(check_edges(p, ed));
  if (num_steps < 1) return fail(MC_ERR_BAD_SHAPE, "num_steps < 1");
  if (!(dt > 0.0)) return fail(MC_ERR_BAD_SHAPE, "dt must be positive");
  hipStream_t s = (hipStream_t)stream;
  const Consts c = derive(p, q, dt);
  MCCHK(ensure_geo(p, q));
  if (q->enable_steric) MCCHK(ensure_kf(p, q->steric_sigma, s));
  Actuation A{};
  const Actuation* ap = nullptr;
  if (act) {
    A = Actuation{act->amp, act->omega, act->phase, p->tick, nullptr};
    ap = &A;
    long long t0 = (long long)act->t0;
    HIPCHK(hipMemcpyAsync(p->tick, &t0, sizeof(t0), hipMemcpyHostToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
  }
  mc_state sa = *st, sb = *st;
  sb.f_grid = p->f_tmp;
  std::vector<hipEvent_t> ev(3 * (size_t)num_steps + 1);
  for (auto& e : ev) HIPCHK(hipEventCreate(&e));
  HIPCHK(hipEventRecord(ev[0], s));
  for (int64_t k = 0; k < num_steps; ++k) {
    const mc_state* in = (k & 1) ? &sb : &sa;
    const mc_state* out = (k & 1) ? &sa : &sb;
    StepSync sy;
    sy.t0 = ev[3 * k + 1];
    sy.t1 = ev[3 * k + 2];
    MCCHK(enqueue_step(p, q, c, in, out, ed, ap, s, sy));
    HIPCHK(hipEventRecord(ev[3 * k + 3], s));
  }
  HIPCHK(hipEventSynchronize(ev[3 * num_steps]));
  double lbm = 0.0;
  for (int64_t k = 0; k < num_steps; ++k) {
    float ms = 0.0f;
    if (q->enable_fluid) {
      HIPCHK(hipEventElapsedTime(&ms, ev[3 * k + 1], ev[3 * k + 2]));
      lbm += ms;
    }
  }
  float tot = 0.0f;
  HIPCHK(hipEventElapsedTime(&tot, ev[0], ev[3 * num_steps]));
  for (auto& e : ev) (void)hipEventDestroy(e);
  if ((num_steps & 1) && q->enable_fluid)
    HIPCHK(hipMemcpyAsync(st->f_grid, p->f_tmp, 9 * (size_t)p->lat_cells * sizeof(float),
                          hipMemcpyDeviceToDevice, s));
  *lbm_ms_avg = lbm / (double)num_steps;
  *step_ms_avg = (double)tot / (double)num_steps;
  return MC_OK;
}

mc_status mc_profile_fused_step(mc_plan* p, const mc_params* q, const mc_state* st,
                                const mc_edge_params* ed, const mc_actuation* act, double dt,
                                int64_t num_steps, double* kernel_ms_avg, double* step_ms_avg,
                                void* stream) {
  g_err.clear();
  if (!p || !kernel_ms_avg || !step_ms_avg) return fail(MC_ERR_INVALID, "NULL argument");
  MCCHK(check_params(p, q, true, true));
  MCCHK(ensure_synth(p, q));
  MCCHK(check_state(p, st, "state"));
  MCCHK(check_edges(p, ed));
  if (num_steps < 2 || (num_steps & 1))
    return fail(MC_ERR_BAD_SHAPE, "num_steps must be even and at least 2");
  if (!(dt > 0.0)) return fail(MC_ERR_BAD_SHAPE, "dt must be positive");
  if (!q->enable_fluid || !p->rod_fuse_ok)
    return fail(MC_ERR_UNSUPPORTED, "no LBM pass is carried by the rod launch on this plan");
  hipStream_t s = (hipStream_t)stream;
  const Consts c = derive(p, q, dt);
  MCCHK(ensure_geo(p, q));
  if (q->enable_steric) MCCHK(ensure_kf(p, q->steric_sigma, s));
  Actuation A{};
  const Actuation* ap = nullptr;
  if (act) {
    A = Actuation{act->amp, act->omega, act->phase, p->tick, nullptr};
    ap = &A;
    long long t0 = (long long)act->t0;
    HIPCHK(hipMemcpyAsync(p->tick, &t0, sizeof(t0), hipMemcpyHostToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
  }
  mc_state sa = *st, sb = *st;
  sb.f_grid = p->f_tmp;
  std::vector<hipEvent_t> ev(2 * (size_t)num_steps + 2);
  for (auto& e : ev) HIPCHK(hipEventCreate(&e));
  HIPCHK(hipEventRecord(ev[0], s));
  LbmArgs pend[2];
  const LbmArgs* fuse = nullptr;
  for (int64_t k = 0; k < num_steps; ++k) {
    const int par = (int)(k & 1);
    LbmArgs* defer = k + 1 < num_steps ? &pend[par] : nullptr;
    StepSync sy;
    sy.t0 = ev[2 * k + 1];
    sy.t1 = ev[2 * k + 2];
    sy.on_rod = true;
    const bool ch = chain_predict(p, c);
    MCCHK(enqueue_step(p, q, c, par ? &sb : &sa, par ? &sa : &sb, ed, ap, s, sy, fuse, defer,
                       SideLbm(), ch && k > 0, ch && k + 1 < num_steps));
    fuse = defer;
  }
  HIPCHK(hipEventRecord(ev[2 * num_steps + 1], s));
  HIPCHK(hipEventSynchronize(ev[2 * num_steps + 1]));
  // steps 1 .. n-1 carry the previous step's LBM pass in their rod launch
  double kern = 0.0;
  for (int64_t k = 1; k < num_steps; ++k) {
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, ev[2 * k + 1], ev[2 * k + 2]));
    kern += ms;
  }
  float tot = 0.0f;
  HIPCHK(hipEventElapsedTime(&tot, ev[0], ev[2 * num_steps + 1]));
  for (auto& e : ev) (void)hipEventDestroy(e);
  *kernel_ms_avg = kern / (double)(num_steps - 1);
  *step_ms_avg = (double)tot / (double)num_steps;
  return MC_OK;
}

mc_status mc_gaussian_filter(mc_plan* p, const float* data, float* out, double sigma, void* stream) {
  g_err.clear();
  if (!p || !data || !out) return fail(MC_ERR_INVALID, "NULL argument");
  if (!p->fft_ok) return fail(MC_ERR_UNSUPPORTED, "%s", p->fft_why.c_str());
  hipStream_t s = (hipStream_t)stream;
  MCCHK(ensure_kf(p, sigma, s));
  HIPCHK(run_filter(p, data, nullptr, nullptr, out, s));
  return MC_OK;
}

mc_status mc_steric_field(mc_plan* p, const mc_params* q, const float* position, float* steric_out,
                          float* force_out, float* density_out, void* stream) {
  g_err.clear();
  if (!p || !q || (!position && p->N > 0) || !steric_out) return fail(MC_ERR_INVALID, "NULL argument");
  mc_params qq = *q;
  qq.enable_steric = 1;
  MCCHK(check_params(p, &qq, false, true));
  hipStream_t s = (hipStream_t)stream;
  const Consts c = derive(p, &qq, 1.0);
  MCCHK(ensure_geo(p, &qq));
  MCCHK(ensure_kf(p, qq.steric_sigma, s));
  MCCHK(run_steric(p, c, position, nullptr, steric_out, density_out, s));
  if (force_out) HIPCHK(launch_steric_nodes(c, p->bd, p->N, steric_out, p->pred_w, force_out, s));
  return MC_OK;
}

mc_status mc_ib_lbm_step(mc_plan* p, const mc_params* q, const float* position, const float* velocity,
                         const float* f_in, float* velocity_out, float* fluid_velocity_out, float* f_out,
                         float* pos_out, double dt, void* stream) {
  g_err.clear();
  if (!p || !f_in || !f_out || !fluid_velocity_out) return fail(MC_ERR_INVALID, "NULL argument");
  if (p->N > 0 && (!position || !velocity || !velocity_out)) return fail(MC_ERR_INVALID, "NULL node argument");
  if (f_in == f_out) return fail(MC_ERR_INVALID, "f_out must differ from f_in");
  mc_params qq = *q;
  qq.enable_fluid = 1;
  MCCHK(check_params(p, &qq, true, false));
  MCCHK(ensure_synth(p, &qq));
  if (!(dt > 0.0)) return fail(MC_ERR_BAD_SHAPE, "dt must be positive");
  hipStream_t s = (hipStream_t)stream;
  MCCHK(check_async(p, s));
  const Consts c = derive(p, &qq, dt);
  MCCHK(ensure_geo(p, &qq));
  if (p->N > 0 && velocity_out != velocity)
    HIPCHK(hipMemcpyAsync(velocity_out, velocity, 2 * (size_t)p->N * sizeof(float), hipMemcpyDeviceToDevice, s));
  HIPCHK(launch_ib(c, p->bd, ib_of(p, c), f_in, position, velocity_out, pos_out ? pos_out : p->node_tmp,
                   nullptr, s));
  HIPCHK(launch_lbm(c, p->bd, f_in, f_out, fluid_velocity_out, p->ib.mask, p->ib.force, s));
  return MC_OK;
}

mc_status mc_rod_solve(mc_plan* p, const mc_params* q, float* position, float* theta, const float* y,
                       const mc_edge_params* ed, int32_t cycles, void* stream) {
  g_err.clear();
  if (!p || (!position && p->N > 0) || (!theta && p->E > 0)) return fail(MC_ERR_INVALID, "NULL argument");
  MCCHK(check_params(p, q, false, false));
  MCCHK(check_edges(p, ed));
  if (cycles < 0) return fail(MC_ERR_BAD_SHAPE, "cycles < 0");
  hipStream_t s = (hipStream_t)stream;
  Consts c = derive(p, q, 1.0);
  c.cycles = cycles;
  HIPCHK(launch_rod(c, p->bd, p->L, p->N, p->rod_lds_bytes, p->rod_use_lds, p->rod_gscratch, 1, position,
                    nullptr, y, theta, ed->rest_lengths, ed->bending_rest_angles, ed->bending_stiffness,
                    nullptr, nullptr, nullptr, position, nullptr, nullptr, theta, nullptr, s));
  return MC_OK;
}

mc_status mc_center_of_mass(const float* position, int32_t n_groups, const int32_t* group_offsets,
                            int32_t grid_h, int32_t grid_w, int32_t periodic, float* com_out,
                            void* stream) {
  g_err.clear();
  if (n_groups < 0) return fail(MC_ERR_BAD_SHAPE, "n_groups < 0");
  if (n_groups == 0) return MC_OK;
  if (!position || !group_offsets || !com_out) return fail(MC_ERR_INVALID, "NULL argument");
  if (grid_h < 1 || grid_w < 1) return fail(MC_ERR_BAD_SHAPE, "grid %dx%d", grid_h, grid_w);
  HIPCHK(launch_com(n_groups, group_offsets, position, grid_h, grid_w, periodic ? 1 : 0, com_out,
                    (hipStream_t)stream));
  return MC_OK;
}

mc_status mc_track_enable(mc_plan* p, int32_t n_groups, const int32_t* group_offsets_host,
                          int32_t start_step) {
  g_err.clear();
  if (!p) return fail(MC_ERR_INVALID, "plan is NULL");
  if (n_groups < 0) return fail(MC_ERR_BAD_SHAPE, "n_groups < 0");
  if (n_groups == 0) {
    p->track.n_groups = 0;
    return MC_OK;
  }
  if (!group_offsets_host) return fail(MC_ERR_INVALID, "group_offsets is NULL");
  if (group_offsets_host[0] < 0 || group_offsets_host[n_groups] > p->N)
    return fail(MC_ERR_BAD_SHAPE, "group offsets outside [0, %d]", p->N);
  for (int g = 0; g < n_groups; ++g)
    if (group_offsets_host[g + 1] <= group_offsets_host[g])
      return fail(MC_ERR_BAD_SHAPE, "group %d is empty or offsets descend", g);
  std::vector<int2> dims(n_groups);
  for (int g = 0; g < n_groups; ++g) {
    const int i0 = p->node_inst_h[group_offsets_host[g]];
    if (p->node_inst_h[group_offsets_host[g + 1] - 1] != i0)
      return fail(MC_ERR_BAD_SHAPE, "group %d spans two instances", g);
    dims[g] = make_int2(p->inst[i0].H, p->inst[i0].W);
  }
  TrackWork t{};
  t.n_groups = n_groups;
  t.start_step = start_step;
  std::vector<int> off(group_offsets_host, group_offsets_host + n_groups + 1);
  MCCHK(upload(p, &t.off, off, "track offsets"));
  MCCHK(upload(p, &t.dims, dims, "track dims"));
  MCCHK(p->alloc(&t.prev, 2 * (size_t)n_groups, "track prev"));
  MCCHK(p->alloc(&t.cum, 2 * (size_t)n_groups, "track cumulative"));
  MCCHK(p->alloc(&t.count, (size_t)n_groups, "track count"));
  p->track = t;
  return MC_OK;
}

mc_status mc_track_reset(mc_plan* p, const float* position, void* stream) {
  g_err.clear();
  if (!p) return fail(MC_ERR_INVALID, "plan is NULL");
  if (p->track.n_groups == 0) return fail(MC_ERR_INVALID, "tracking is not enabled");
  if (!position) return fail(MC_ERR_INVALID, "position is NULL");
  HIPCHK(launch_track(p->track, position, 0, (hipStream_t)stream));
  return MC_OK;
}

mc_status mc_track_read(mc_plan* p, float* cumulative_out, float* com_out, void* stream) {
  g_err.clear();
  if (!p) return fail(MC_ERR_INVALID, "plan is NULL");
  if (p->track.n_groups == 0) return fail(MC_ERR_INVALID, "tracking is not enabled");
  hipStream_t s = (hipStream_t)stream;
  MCCHK(check_async(p, s));
  const size_t b = 2 * (size_t)p->track.n_groups * sizeof(float);
  if (cumulative_out)
    HIPCHK(hipMemcpyAsync(cumulative_out, p->track.cum, b, hipMemcpyDeviceToDevice, s));
  if (com_out) HIPCHK(hipMemcpyAsync(com_out, p->track.prev, b, hipMemcpyDeviceToDevice, s));
  return MC_OK;
}

}  // extern "C"

namespace {

// One tracker group per instance (mc_batch_rollout / mc_batch_advance),
// allocated on first use.
mc_status ensure_inst_track(mc_plan* p) {
  if (p->inst_track.n_groups != 0) return MC_OK;
  for (int i = 0; i < p->n_inst; ++i)
    if (p->inst[i].nodes == 0) return fail(MC_ERR_BAD_SHAPE, "instance %d has no nodes to track", i);
  std::vector<int> off(1, 0);
  std::vector<int2> dims;
  for (const InstHost& h : p->inst) {
    off.push_back(h.node0 + h.nodes);
    dims.push_back(make_int2(h.H, h.W));
  }
  TrackWork t{};
  t.n_groups = p->n_inst;
  t.start_step = 0;
  MCCHK(upload(p, &t.off, off, "instance offsets"));
  MCCHK(upload(p, &t.dims, dims, "instance dims"));
  MCCHK(p->alloc(&t.prev, 2 * (size_t)p->n_inst, "instance prev"));
  MCCHK(p->alloc(&t.cum, 2 * (size_t)p->n_inst, "instance cumulative"));
  MCCHK(p->alloc(&t.count, (size_t)p->n_inst, "instance count"));
  p->inst_track = t;
  return MC_OK;
}

// Rollout of every instance with the per-instance tracker swapped in.
// mode 0 starts the displacement from zero, mode 2 from `disp_in` (device).
mc_status tracked_rollout(mc_plan* p, const mc_params* q, const mc_state* st,
                          const mc_edge_params* ed, const mc_actuation* act,
                          const long long* t_inst, double dt, int64_t num_steps,
                          const float* disp_in, float* disp_out, hipStream_t s,
                          const mc_actions* rows = nullptr) {
  if (disp_out) {
    if (!st || (p->N > 0 && !st->position)) return fail(MC_ERR_INVALID, "state is NULL");
    MCCHK(ensure_inst_track(p));
    if (disp_in)
      HIPCHK(hipMemcpyAsync(p->inst_track.cum, disp_in, 2 * (size_t)p->n_inst * sizeof(float),
                            hipMemcpyDeviceToDevice, s));
    HIPCHK(launch_track(p->inst_track, st->position, disp_in ? 2 : 0, s));
  }
  const TrackWork keep = p->track;
  if (disp_out) p->track = p->inst_track;
  const mc_status r =
      rollout_impl(p, q, st, ed, act, t_inst, dt, num_steps, nullptr, nullptr, 1, s, rows);
  p->track = keep;
  if (r != MC_OK) return r;
  if (disp_out)
    HIPCHK(hipMemcpyAsync(disp_out, p->inst_track.cum, 2 * (size_t)p->n_inst * sizeof(float),
                          hipMemcpyDeviceToDevice, s));
  return MC_OK;
}

}  // namespace

extern "C" {

mc_status mc_batch_rollout(mc_plan* p, const mc_params* q, const mc_state* st,
                           const mc_edge_params* ed, const mc_actuation* act, double dt,
                           int64_t num_steps, float* displacement_out, void* stream) {
  g_err.clear();
  if (!p) return fail(MC_ERR_INVALID, "plan is NULL");
  return tracked_rollout(p, q, st, ed, act, nullptr, dt, num_steps, nullptr, displacement_out,
                         (hipStream_t)stream);
}

mc_status mc_rollout_record(mc_plan* p, const mc_params* q, const mc_state* st,
                            const mc_edge_params* ed, const mc_actuation* act, double dt,
                            int64_t num_steps, const mc_record* rec, int32_t use_graph,
                            void* stream) {
  g_err.clear();
  if (!p) return fail(MC_ERR_INVALID, "plan is NULL");
  if (!rec) return fail(MC_ERR_INVALID, "record is NULL");
  if (p->n_inst != 1) return fail(MC_ERR_UNSUPPORTED, "recording needs a single-instance plan");
  if (rec->stride < 1) return fail(MC_ERR_BAD_SHAPE, "stride must be at least 1");
  if (!st) return fail(MC_ERR_INVALID, "state is NULL");
  if ((rec->position && !st->position) || (rec->velocity && !st->velocity) ||
      (rec->debug_vector && !st->debug_vector) || (rec->steric && !st->steric) ||
      (rec->fluid_velocity && !st->fluid_velocity))
    return fail(MC_ERR_INVALID, "a recorded state array is NULL");
  return rollout_impl(p, q, st, ed, act, nullptr, dt, num_steps, nullptr, nullptr, use_graph,
                      stream, nullptr, rec);
}

mc_status mc_rollout_actions(mc_plan* p, const mc_params* q, const mc_state* st,
                             const mc_edge_params* ed, const mc_actions* actions, double dt,
                             int64_t num_steps, float* displacement_out, void* stream) {
  g_err.clear();
  if (!p) return fail(MC_ERR_INVALID, "plan is NULL");
  if (!actions) return fail(MC_ERR_INVALID, "actions is NULL");
  return tracked_rollout(p, q, st, ed, nullptr, nullptr, dt, num_steps, nullptr, displacement_out,
                         (hipStream_t)stream, actions);
}

mc_status mc_batch_advance(mc_plan* p, const mc_params* q, const mc_state* st,
                           const mc_edge_params* ed, const mc_actuation* act,
                           const int64_t* instance_t0, double dt, int64_t num_steps,
                           float* displacement_io, void* stream) {
  g_err.clear();
  if (!p) return fail(MC_ERR_INVALID, "plan is NULL");
  hipStream_t s = (hipStream_t)stream;
  const long long* t_inst = nullptr;
  if (act && instance_t0) {
    if (!p->t_inst) MCCHK(p->alloc(&p->t_inst, (size_t)p->n_inst, "instance time offsets"));
    bool same = p->t_inst_h.size() == (size_t)p->n_inst;
    for (int i = 0; same && i < p->n_inst; ++i) same = p->t_inst_h[i] == (long long)instance_t0[i];
    if (!same) {
      // earlier launches on this stream may still read the old offsets
      HIPCHK(hipStreamSynchronize(s));
      p->t_inst_h.assign(instance_t0, instance_t0 + p->n_inst);
      HIPCHK(hipMemcpy(p->t_inst, p->t_inst_h.data(), (size_t)p->n_inst * sizeof(long long),
                       hipMemcpyHostToDevice));
    }
    t_inst = p->t_inst;
  }
  return tracked_rollout(p, q, st, ed, act, t_inst, dt, num_steps, displacement_io,
                         displacement_io, s);
}

}  // extern "C"

namespace {

mc_status check_rod_adjoint(const mc_params* q) {
  if (!q) return fail(MC_ERR_INVALID, "params is NULL");
  if (q->solver_type != MC_SOLVER_PBD && q->solver_type != MC_SOLVER_COSSERAT)
    return fail(MC_ERR_UNSUPPORTED, "solver_type %d", q->solver_type);
  const int kern = q->bending_constraint == MC_CONSTRAINT_AUTO
                       ? q->solver_type
                       : (q->bending_constraint == MC_CONSTRAINT_COSSERAT ? 1 : 0);
  if (kern != q->solver_type)
    return fail(MC_ERR_UNSUPPORTED,
                "the rod adjoint needs bending_constraint and solver_type to agree "
                "(a Cosserat constraint with solver_type pbd, or the reverse)");
  if (q->enable_fluid && q->synthetic_node_width > 0.0)
    return fail(MC_ERR_UNSUPPORTED, "the adjoint of the fluid-coupled step takes "
                "synthetic_node_width 0 (IB on the nodes)");
  return MC_OK;
}

// Largest steric tape the adjoint keeps (field, force and prediction of every
// step); longer or larger rollouts recompute them in the reverse pass.
constexpr size_t ADJ_STERIC_TAPE_MAX = (size_t)2 << 30;

// Steric-term scratch of the adjoint: per-node force and pb1, three grids.
// The fluid adjoint's buffers of a single-instance plan (allocated once).
mc_status ensure_adj_fluid(mc_plan* p, const mc_params* q) {
  if (p->n_inst != 1)
    return fail(MC_ERR_UNSUPPORTED, "the adjoint of the fluid-coupled step takes single-instance "
                "plans");
  if (p->adj_fw) return MC_OK;
  const long long HW = p->lat_cells;
  const int kmax = p->kmax;
  MCCHK(p->alloc(&p->adj_fx, 6 * (size_t)p->N + (size_t)p->E + 1, "fluid adjoint rod outputs"));
  MCCHK(p->alloc(&p->adj_fb, 18 * (size_t)HW, "lattice cotangents"));
  MCCHK(p->alloc(&p->adj_fw, fluid_adj_floats(p->N, kmax, std::max(q->ibm_iterations, 1), HW),
                 "fluid adjoint work"));
  MCCHK(p->alloc(&p->adj_iw, fluid_adj_ints(p->N, kmax, HW), "fluid adjoint lists"));
  return MC_OK;
}

mc_status ensure_adj_steric(mc_plan* p) {
  if (!p->adj_steric)
    MCCHK(p->alloc(&p->adj_steric, 4 * (size_t)p->N + 3 * (size_t)p->grid_cells,
                   "adjoint steric scratch"));
  return MC_OK;
}

}  // namespace

extern "C" {

mc_status mc_rod_adjoint_forward(mc_plan* p, const mc_params* q, const mc_state* st,
                                 const mc_edge_params* ed, double dt, int64_t num_steps,
                                 int32_t checkpoint_every, void* stream) {
  g_err.clear();
  if (!p) return fail(MC_ERR_INVALID, "plan is NULL");
  MCCHK(check_rod_adjoint(q));
  MCCHK(check_params(p, q, false, true));
  MCCHK(check_state(p, st, "state"));
  MCCHK(check_edges(p, ed));
  if (num_steps < 1) return fail(MC_ERR_BAD_SHAPE, "num_steps < 1");
  if (checkpoint_every < 1) return fail(MC_ERR_BAD_SHAPE, "checkpoint_every < 1");
  if (!(dt > 0.0)) return fail(MC_ERR_BAD_SHAPE, "dt must be positive");
  // the steric term and the fluid couple all filaments every step: every
  // step's state is kept
  const bool steric = q->enable_steric != 0;
  const bool fluid = q->enable_fluid != 0;
  if (steric || fluid) checkpoint_every = 1;
  if (fluid) MCCHK(ensure_adj_fluid(p, q));
  const size_t lds = rod_segment_lds_bytes(q->cycles_per_step, p->comp_max_nn, p->comp_max_ne,
                                           p->comp_max_nb, true);
  if (lds + 1024 > (size_t)p->lds_max)
    return fail(MC_ERR_UNSUPPORTED, "a filament of %d nodes with %d cycles needs %zu B of LDS for "
                "the adjoint (device: %d B)", p->comp_max_nn, q->cycles_per_step, lds, p->lds_max);
  hipStream_t s = (hipStream_t)stream;
  p->adj_ok = false;
  const size_t S = 4 * (size_t)p->N + (size_t)p->E;
  const int64_t nck = (num_steps + checkpoint_every - 1) / checkpoint_every;
  if (p->adj_ckpt_cap < (size_t)nck * S) {
    // earlier work on the plan may still read the old buffer: it stays
    // allocated until the plan is destroyed
    MCCHK(p->alloc(&p->adj_ckpt, (size_t)nck * S, "adjoint checkpoints"));
    p->adj_ckpt_cap = (size_t)nck * S;
  }
  RodAdjArgs a{};
  a.c = derive(p, q, dt);
  a.bd = p->bd;
  a.L = p->L;
  a.N = p->N;
  a.pos = st->position;
  a.vel = st->velocity;
  a.theta = st->theta;
  a.rest_len = ed->rest_lengths;
  a.rest_ang = ed->bending_rest_angles;
  a.kb = ed->bending_stiffness;
  const size_t lds_f = rod_segment_lds_bytes(q->cycles_per_step, p->comp_max_nn, p->comp_max_ne,
                                             p->comp_max_nb, false);
  // steric tape slot per step: field (grid), force (2N), wrapped prediction (2N)
  const size_t slot = (size_t)p->grid_cells + 4 * (size_t)p->N;
  size_t tape_max = ADJ_STERIC_TAPE_MAX;
  if (const char* env = getenv("MC_ADJ_STERIC_TAPE_MAX"))  // bytes; 0 always recomputes
    tape_max = (size_t)strtoull(env, nullptr, 10);
  const bool store = steric && !fluid && (size_t)num_steps * slot * sizeof(float) <= tape_max;
  if (steric) {
    MCCHK(ensure_geo(p, q));
    MCCHK(ensure_kf(p, q->steric_sigma, s));
    MCCHK(ensure_adj_steric(p));
    a.fst = p->adj_steric;
    if (store && p->adj_stape_cap < (size_t)num_steps * slot) {
      MCCHK(p->alloc(&p->adj_stape, (size_t)num_steps * slot, "adjoint steric tape"));
      p->adj_stape_cap = (size_t)num_steps * slot;
    }
  }
  if (fluid) {
    // coupled steps as mc_rollout takes them (unfused), each after a copy of
    // its input state to the checkpoint slot and of its lattice to the tape
    MCCHK(ensure_geo(p, q));
    MCCHK(check_async(p, s));
    const size_t hw9 = 9 * (size_t)p->lat_cells;
    if (p->adj_ftape_cap < (size_t)num_steps * hw9) {
      MCCHK(p->alloc(&p->adj_ftape, (size_t)num_steps * hw9, "adjoint lattice tape"));
      p->adj_ftape_cap = (size_t)num_steps * hw9;
    }
    mc_state sa = *st, sb = *st;
    sb.f_grid = p->f_tmp;
    const size_t n2 = 2 * (size_t)p->N * sizeof(float);
    for (int64_t k = 0; k < num_steps; ++k) {
      const mc_state* in = (k & 1) ? &sb : &sa;
      const mc_state* out = (k & 1) ? &sa : &sb;
      float* ck = p->adj_ckpt + (size_t)k * S;
      if (p->N) {
        HIPCHK(hipMemcpyAsync(ck, in->position, n2, hipMemcpyDeviceToDevice, s));
        HIPCHK(hipMemcpyAsync(ck + 2 * (size_t)p->N, in->velocity, n2, hipMemcpyDeviceToDevice, s));
      }
      if (p->E)
        HIPCHK(hipMemcpyAsync(ck + 4 * (size_t)p->N, in->theta, (size_t)p->E * sizeof(float),
                              hipMemcpyDeviceToDevice, s));
      HIPCHK(hipMemcpyAsync(p->adj_ftape + (size_t)k * hw9, in->f_grid, hw9 * sizeof(float),
                            hipMemcpyDeviceToDevice, s));
      MCCHK(enqueue_step(p, q, a.c, in, out, ed, nullptr, s));
    }
    if (num_steps & 1)
      HIPCHK(hipMemcpyAsync(st->f_grid, p->f_tmp, hw9 * sizeof(float), hipMemcpyDeviceToDevice, s));
  }
  for (int64_t seg = 0; seg < (fluid ? 0 : nck); ++seg) {
    const int len = (int)std::min<int64_t>(checkpoint_every, num_steps - seg * checkpoint_every);
    if (store) {
      // the step's field, force and prediction, kept for the reverse pass
      float* vt = p->adj_stape + (size_t)seg * slot;
      float* ft = vt + p->grid_cells;
      float* pt = ft + 2 * (size_t)p->N;
      MCCHK(run_steric(p, a.c, st->position, st->velocity, vt, nullptr, s, pt));
      HIPCHK(launch_steric_nodes(a.c, p->bd, p->N, vt, pt, ft, s));
      a.fst = ft;
    } else if (steric) {
      // the step's prediction, field and per-node force, as the coupled step
      MCCHK(run_steric(p, a.c, st->position, st->velocity, st->steric, nullptr, s));
      HIPCHK(launch_steric_nodes(a.c, p->bd, p->N, st->steric, p->pred_w, p->adj_steric, s));
    }
    HIPCHK(launch_rod_segment(a, false, len, p->adj_ckpt + (size_t)seg * S, st->debug_vector,
                              lds_f, s));
  }
  if (store)  // the state's field is the last step's, as after the coupled step
    HIPCHK(hipMemcpyAsync(st->steric, p->adj_stape + (size_t)(nck - 1) * slot,
                          (size_t)p->grid_cells * sizeof(float), hipMemcpyDeviceToDevice, s));
  p->adj_stored = store;
  p->adj_ok = true;
  p->adj_steps = num_steps;
  p->adj_every = checkpoint_every;
  p->adj_dt = dt;
  p->adj_q = *q;
  p->adj_ed = *ed;
  return MC_OK;
}

}  // extern "C"

namespace {

// The reverse pass with position cotangents at frames (step counts, each at
// a checkpoint boundary or the last step): a frame's cotangent joins the
// position adjoint when the reverse pass reaches the state after that step.
mc_status adjoint_backward(mc_plan* p, const mc_params* q, const mc_edge_params* ed, double dt,
                           int nframes, const int64_t* frame_steps, const float* frame_bars,
                           const float* velocity_bar, float* rest_angle_bar,
                           float* rest_length_bar, float* position0_bar, float* velocity0_bar,
                           float* theta0_bar, void* stream) {
  if (!p) return fail(MC_ERR_INVALID, "plan is NULL");
  MCCHK(check_rod_adjoint(q));
  if (!p->adj_ok) return fail(MC_ERR_INVALID, "no rollout recorded by mc_rod_adjoint_forward");
  if (!ed || memcmp(q, &p->adj_q, sizeof(mc_params)) != 0 || dt != p->adj_dt ||
      memcmp(ed, &p->adj_ed, sizeof(mc_edge_params)) != 0)
    return fail(MC_ERR_INVALID, "params, dt and edge arrays must be those of the recorded rollout");
  if (nframes < 1 || !frame_steps || !frame_bars || (p->B > 0 && !rest_angle_bar))
    return fail(MC_ERR_INVALID, "frames / position cotangents / rest_angle_bar missing");
  for (int f = 0; f < nframes; ++f) {
    const int64_t t = frame_steps[f];
    if (t < 1 || t > p->adj_steps || (t != p->adj_steps && t % p->adj_every != 0))
      return fail(MC_ERR_BAD_SHAPE, "frame step %lld: must be in 1..%lld and a multiple of the "
                  "checkpoint interval %d (or the last step)", (long long)t,
                  (long long)p->adj_steps, p->adj_every);
  }
  const size_t lds = rod_segment_lds_bytes(q->cycles_per_step, p->comp_max_nn, p->comp_max_ne,
                                           p->comp_max_nb, true);
  hipStream_t s = (hipStream_t)stream;
  const int N = p->N, E = p->E, B = p->B;
  const size_t S = 4 * (size_t)N + (size_t)E;
  const int every = p->adj_every;
  const int64_t T = p->adj_steps;
  const size_t tape = 4 * (size_t)every * (size_t)std::max(q->cycles_per_step, 1) * (size_t)E;
  if (p->adj_tape_cap < tape) {
    MCCHK(p->alloc(&p->adj_tape, tape, "adjoint tape"));
    p->adj_tape_cap = tape;
  }
  const size_t utape = 2 * (size_t)every * (size_t)std::max(q->cycles_per_step, 1) * (size_t)N;
  if (rest_length_bar && q->solver_type == MC_SOLVER_COSSERAT && p->adj_utape_cap < utape) {
    MCCHK(p->alloc(&p->adj_utape, utape, "adjoint node-update tape"));
    p->adj_utape_cap = utape;
  }
  if (!p->adj_work) MCCHK(p->alloc(&p->adj_work, S, "adjoint state"));
  float* xb = p->adj_work;
  float* vb = xb + 2 * (size_t)N;
  float* thb = vb + 2 * (size_t)N;
  const size_t n2 = 2 * (size_t)N * sizeof(float);
  HIPCHK(hipMemsetAsync(xb, 0, n2, s));
  if (velocity_bar) HIPCHK(hipMemcpyAsync(vb, velocity_bar, n2, hipMemcpyDeviceToDevice, s));
  else HIPCHK(hipMemsetAsync(vb, 0, n2, s));
  if (E) HIPCHK(hipMemsetAsync(thb, 0, (size_t)E * sizeof(float), s));
  if (B) HIPCHK(hipMemsetAsync(rest_angle_bar, 0, (size_t)B * sizeof(float), s));
  if (rest_length_bar && E)
    HIPCHK(hipMemsetAsync(rest_length_bar, 0, (size_t)E * sizeof(float), s));
  RodAdjArgs a{};
  a.c = derive(p, q, dt);
  a.bd = p->bd;
  a.L = p->L;
  a.N = N;
  a.rest_len = ed->rest_lengths;
  a.rest_ang = ed->bending_rest_angles;
  a.kb = ed->bending_stiffness;
  a.xb = xb;
  a.vb = vb;
  a.thb = thb;
  a.rest_bar = rest_angle_bar;
  a.len_bar = E ? rest_length_bar : nullptr;
  a.utape = p->adj_utape;
  const int64_t nck = (T + every - 1) / every;
  const bool steric = q->enable_steric != 0;
  float *fst = nullptr, *pb1 = nullptr, *vg = nullptr, *vbar = nullptr, *dbar = nullptr;
  if (steric) {
    MCCHK(ensure_adj_steric(p));
    fst = p->adj_steric;
    pb1 = fst + 2 * (size_t)N;
    vg = pb1 + 2 * (size_t)N;
    vbar = vg + p->grid_cells;
    dbar = vbar + p->grid_cells;
    a.fst = fst;
    a.pb1 = pb1;
  }
  const bool fluid = q->enable_fluid != 0;
  float *fbn = nullptr, *fbo = nullptr;
  if (fluid) {
    MCCHK(ensure_adj_fluid(p, q));
    MCCHK(ensure_geo(p, q));
    // the final lattice carries no cotangent (the losses score node states)
    fbn = p->adj_fb;
    fbo = fbn + 9 * (size_t)p->lat_cells;
    HIPCHK(hipMemsetAsync(fbn, 0, 9 * (size_t)p->lat_cells * sizeof(float), s));
  }
  for (int64_t seg = nck - 1; seg >= 0; --seg) {
    const int len = (int)std::min<int64_t>(every, T - seg * every);
    // the cotangents of the frames at the state this segment ends in
    for (int f = 0; f < nframes; ++f)
      if (frame_steps[f] == seg * every + len)
        HIPCHK(launch_add(xb, frame_bars + (size_t)f * 2 * N, 2 * (size_t)N, s));
    const float* ck = p->adj_ckpt + (size_t)seg * S;
    if (fluid) {
      // the step's rod half again from its checkpoint (positions before the
      // wrap, velocities before the drag), then the fluid half's reverse:
      // vb becomes the cotangent of those velocities, fbn that of the step's
      // input lattice
      float* x1 = p->adj_fx;
      float* v1 = x1 + 2 * (size_t)N;
      float* d1 = v1 + 2 * (size_t)N;
      float* t1 = d1 + 2 * (size_t)N;
      MCCHK(run_steric(p, a.c, ck, ck + 2 * (size_t)N, steric ? vg : nullptr, nullptr, s));
      HIPCHK(launch_rod(a.c, p->bd, p->L, N, p->rod_lds_bytes, p->rod_use_lds, p->rod_gscratch, 0,
                        ck, ck + 2 * (size_t)N, nullptr, ck + 4 * (size_t)N, ed->rest_lengths,
                        ed->bending_rest_angles, ed->bending_stiffness, nullptr,
                        steric ? vg : nullptr, p->pred_w, x1, v1, d1, t1, nullptr, s,
                        p->all_chains));
      HIPCHK(fluid_adjoint_step(a.c, p->geo_h[0], N, p->adj_fw, p->adj_iw,
                                p->adj_ftape + (size_t)seg * 9 * (size_t)p->lat_cells, x1, v1, vb,
                                fbn, fbo, s));
      std::swap(fbn, fbo);
    }
    const float *vt = vg, *pt = p->pred_w;
    if (steric && p->adj_stored) {
      // the step's field, force and prediction from the forward's tape
      const size_t slot = (size_t)p->grid_cells + 4 * (size_t)N;
      vt = p->adj_stape + (size_t)seg * slot;
      a.fst = vt + p->grid_cells;
      pt = a.fst + 2 * (size_t)N;
    } else if (steric) {
      // recompute the step's prediction, field and force from its state
      if (!fluid) MCCHK(run_steric(p, a.c, ck, ck + 2 * (size_t)N, vg, nullptr, s));
      HIPCHK(launch_steric_nodes(a.c, p->bd, N, vg, p->pred_w, fst, s));
    }
    HIPCHK(launch_rod_segment(a, true, len, const_cast<float*>(ck), p->adj_tape, lds, s));
    if (steric) {
      HIPCHK(launch_steric_bar(a.c, p->bd, pt, pb1, vbar, s));
      HIPCHK(run_filter(p, vbar, nullptr, nullptr, dbar, s));
      HIPCHK(launch_steric_adj_nodes(a.c, p->bd, N, pt, vt, dbar, pb1, xb, vb, s));
    }
  }
  if (position0_bar) HIPCHK(hipMemcpyAsync(position0_bar, xb, n2, hipMemcpyDeviceToDevice, s));
  if (velocity0_bar) HIPCHK(hipMemcpyAsync(velocity0_bar, vb, n2, hipMemcpyDeviceToDevice, s));
  if (theta0_bar && E)
    HIPCHK(hipMemcpyAsync(theta0_bar, thb, (size_t)E * sizeof(float), hipMemcpyDeviceToDevice, s));
  return MC_OK;
}

}  // namespace

extern "C" {

mc_status mc_rod_adjoint_backward(mc_plan* p, const mc_params* q, const mc_edge_params* ed,
                                  double dt, const float* position_bar, const float* velocity_bar,
                                  float* rest_angle_bar, float* position0_bar,
                                  float* velocity0_bar, float* theta0_bar, void* stream) {
  g_err.clear();
  if (!position_bar) return fail(MC_ERR_INVALID, "position_bar is NULL");
  const int64_t last = p ? p->adj_steps : 0;
  return adjoint_backward(p, q, ed, dt, 1, &last, position_bar, velocity_bar, rest_angle_bar,
                          nullptr, position0_bar, velocity0_bar, theta0_bar, stream);
}

mc_status mc_rod_adjoint_state(mc_plan* p, int64_t step, float* position_out,
                               float* velocity_out, float* theta_out, void* stream) {
  g_err.clear();
  if (!p) return fail(MC_ERR_INVALID, "plan is NULL");
  if (!p->adj_ok) return fail(MC_ERR_INVALID, "no rollout recorded by mc_rod_adjoint_forward");
  if (step < 0 || step >= p->adj_steps || step % p->adj_every != 0)
    return fail(MC_ERR_BAD_SHAPE, "step %lld: the recorded states are at multiples of %d below %lld",
                (long long)step, p->adj_every, (long long)p->adj_steps);
  hipStream_t s = (hipStream_t)stream;
  const size_t S = 4 * (size_t)p->N + (size_t)p->E, n2 = 2 * (size_t)p->N * sizeof(float);
  const float* ck = p->adj_ckpt + (size_t)(step / p->adj_every) * S;
  if (position_out) HIPCHK(hipMemcpyAsync(position_out, ck, n2, hipMemcpyDeviceToDevice, s));
  if (velocity_out)
    HIPCHK(hipMemcpyAsync(velocity_out, ck + 2 * (size_t)p->N, n2, hipMemcpyDeviceToDevice, s));
  if (theta_out && p->E)
    HIPCHK(hipMemcpyAsync(theta_out, ck + 4 * (size_t)p->N, (size_t)p->E * sizeof(float),
                          hipMemcpyDeviceToDevice, s));
  return MC_OK;
}

mc_status mc_rod_adjoint_backward_frames(mc_plan* p, const mc_params* q,
                                         const mc_edge_params* ed, double dt, int32_t num_frames,
                                         const int64_t* frame_steps,
                                         const float* frame_position_bars,
                                         const float* velocity_bar, float* rest_angle_bar,
                                         float* position0_bar, float* velocity0_bar,
                                         float* theta0_bar, void* stream) {
  g_err.clear();
  return adjoint_backward(p, q, ed, dt, num_frames, frame_steps, frame_position_bars,
                          velocity_bar, rest_angle_bar, nullptr, position0_bar, velocity0_bar,
                          theta0_bar, stream);
}

mc_status mc_rod_adjoint_backward_params(mc_plan* p, const mc_params* q,
                                         const mc_edge_params* ed, double dt, int32_t num_frames,
                                         const int64_t* frame_steps,
                                         const float* frame_position_bars,
                                         const float* velocity_bar, float* rest_angle_bar,
                                         float* rest_length_bar, float* position0_bar,
                                         float* velocity0_bar, float* theta0_bar, void* stream) {
  g_err.clear();
  if (p && p->E > 0 && !rest_length_bar) return fail(MC_ERR_INVALID, "rest_length_bar is NULL");
  return adjoint_backward(p, q, ed, dt, num_frames, frame_steps, frame_position_bars,
                          velocity_bar, rest_angle_bar, rest_length_bar, position0_bar,
                          velocity0_bar, theta0_bar, stream);
}


int32_t mc_filter_taps(int32_t n, double sigma, int32_t* lo, float* taps, int32_t capacity) {
  g_err.clear();
  if (n < 1 || !(sigma > 0.0) || !lo || (!taps && capacity > 0)) {
    fail(MC_ERR_INVALID, "mc_filter_taps: bad argument");
    return -1;
  }
  int l = 0;
  std::vector<float> t;
  if (!gauss_taps(n, sigma, &l, &t)) return 0;
  *lo = l;
  for (int i = 0; i < (int)t.size() && i < capacity; ++i) taps[i] = t[i];
  return (int32_t)t.size();
}

int32_t mc_math_f32(int32_t fn, const float* x, int64_t n, float* out0, float* out1) {
  g_err.clear();
  if (n < 0 || (n > 0 && (!x || !out0 || (fn == 0 && !out1))) || fn < 0 || fn > 1)
    return fail(MC_ERR_INVALID, "mc_math_f32: bad argument");
  for (int64_t i = 0; i < n; ++i) {
    if (fn == 0) mc::fsincos(x[i], &out0[i], &out1[i]);
    else out0[i] = mc::fexp(x[i]);
  }
  return MC_OK;
}

int32_t mc_resize_taps(int32_t n_in, int32_t n_out, int32_t* offsets, int32_t* index,
                       float* weight, int32_t capacity) {
  if (n_in < 1 || n_out < 1 || !offsets) return -1;
  std::vector<int> off, idx;
  std::vector<float> w;
  resize_taps(n_in, n_out, off, idx, w);
  for (int i = 0; i <= n_out; ++i) offsets[i] = off[i];
  for (int i = 0; i < (int)idx.size() && i < capacity; ++i) {
    if (index) index[i] = idx[i];
    if (weight) weight[i] = w[i];
  }
  return (int32_t)idx.size();
}

size_t mc_point_loss_workspace(int64_t n, int64_t m) {
  return n < 0 || m < 0 ? 0 : point_loss_workspace(n, m);
}

static mc_status check_point_loss(int32_t kind, const float* particles, int64_t n,
                                  const float* targets, int64_t m, double alpha, const float* loss) {
  if (kind != MC_LOSS_CHAMFER && kind != MC_LOSS_DCD && kind != MC_LOSS_SMOOTH_MAX)
    return fail(MC_ERR_INVALID, "unknown loss kind %d", kind);
  if (n < 1 || m < 1 || n > (1 << 24) || m > (1 << 24))
    return fail(MC_ERR_BAD_SHAPE, "particle and target counts must be in 1 .. 2^24");
  if (!particles || !targets || !loss) return fail(MC_ERR_INVALID, "NULL argument");
  if (kind == MC_LOSS_SMOOTH_MAX && !(alpha > 0.0))
    return fail(MC_ERR_BAD_SHAPE, "alpha must be positive");
  return MC_OK;
}

mc_status mc_point_loss(int32_t kind, const float* particles, int64_t n, const float* targets,
                        int64_t m, double alpha, float* loss, float* grad, void* workspace,
                        void* stream) {
  g_err.clear();
  const mc_status st = check_point_loss(kind, particles, n, targets, m, alpha, loss);
  if (st != MC_OK) return st;
  if (!workspace) return fail(MC_ERR_INVALID, "NULL argument");
  HIPCHK(launch_point_loss(kind, particles, (int)n, targets, (int)m, (float)alpha, loss, grad,
                           workspace, (hipStream_t)stream));
  return MC_OK;
}

mc_status mc_point_loss_static(int32_t kind, const float* particles, int64_t n,
                               const float* targets, int64_t m, double alpha, float* loss,
                               float* grad, void* stream) {
  g_err.clear();
  const mc_status st = check_point_loss(kind, particles, n, targets, m, alpha, loss);
  if (st != MC_OK) return st;
  if (n > MC_POINT_LOSS_STATIC_MAX || m > MC_POINT_LOSS_STATIC_MAX)
    return fail(MC_ERR_BAD_SHAPE,
                "%lld particles, %lld targets: the static workspace holds at most %d of each",
                (long long)n, (long long)m, MC_POINT_LOSS_STATIC_MAX);
  void* ws = nullptr;
  HIPCHK(point_loss_scratch(&ws));
  HIPCHK(launch_point_loss(kind, particles, (int)n, targets, (int)m, (float)alpha, loss, grad, ws,
                           (hipStream_t)stream));
  return MC_OK;
}

mc_status mc_frame_stability(const float* frames, int32_t num_frames, int64_t n,
                             const float* main_pos, const float* anchor_pos, int32_t height,
                             int32_t width, float* loss, float* frames_grad, float* main_grad,
                             float* anchor_grad, void* stream) {
  g_err.clear();
  if (num_frames < 1 || num_frames > 4096 || n < 1 || n > (1 << 24))
    return fail(MC_ERR_BAD_SHAPE, "need 1 .. 4096 frames of 1 .. 2^24 nodes");
  if (height < 1 || width < 1) return fail(MC_ERR_BAD_SHAPE, "grid must be at least 1 x 1");
  if (!frames || !main_pos || !loss) return fail(MC_ERR_INVALID, "NULL argument");
  if (frames_grad && (!main_grad || (anchor_pos && !anchor_grad)))
    return fail(MC_ERR_INVALID, "frames_grad needs main_grad (and anchor_grad with an anchor)");
  HIPCHK(launch_stability(frames, num_frames, (int)n, main_pos, anchor_pos, height, width, loss,
                          frames_grad, frames_grad ? main_grad : nullptr,
                          frames_grad ? anchor_grad : nullptr, (hipStream_t)stream));
  return MC_OK;
}

mc_status mc_plan_instances(const mc_plan* p, int32_t* num_instances, int64_t* offsets) {
  g_err.clear();
  if (!p || !num_instances) return fail(MC_ERR_INVALID, "NULL argument");
  *num_instances = p->n_inst;
  if (offsets)
    for (int i = 0; i < p->n_inst; ++i) {
      const InstGeo& g = p->geo_h[i];
      offsets[3 * i] = g.grid_off;
      offsets[3 * i + 1] = g.fl_off;
      offsets[3 * i + 2] = g.node0;
    }
  return MC_OK;
}

}  // extern "C"
