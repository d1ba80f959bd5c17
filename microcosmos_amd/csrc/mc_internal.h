// Internal (non-exported) declarations shared by the kernel translation units
// and the C-ABI layer.
#pragma once

#include <vector>
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace mc {

constexpr int IB_KMAX = 6;                 // largest supported ibm_kernel_size
constexpr int SORT_LDS_CAP = 2048;         // bitonic sort capacity in LDS (ids)
constexpr int FFT_MAX_RADICES = 16;        // radix stages of one FFT length
constexpr int IB_WIDE_MIN = 768;           // clusters with more nodes take the wide solve
constexpr int IB_CLUSTER_MAX = 1024;       // largest cluster the per-cluster solve can take
constexpr int IB_WIDE_WG = 256;            // workgroups of the wide solve launch (at most)
constexpr int IB_WIDE_NODES_PER_WG = 192;  // wide nodes per participating workgroup

// Float constants derived from ConstraintSolverConfig exactly as the reference
// derives them (Python double arithmetic, then a float32 weak scalar).  Shared
// by every instance of a plan; per-instance geometry is in InstGeo.
struct Consts {
  // step (simulate.py:37-100)
  float dt, damping;
  int enable_fluid, enable_steric;
  // steric (forces.py)
  float steric_strength, steric_sv, steric_2s2, steric_norm, steric_svs, fft_coef;
  int steric_nskip;
  // rod (pbd.py / stable_cosserat.py)
  int solver;                 // constraint kernel: 0 pbd, 1 cosserat (bending_constraint)
  int cos_target;             // solver_type == cosserat: y = debug_vector = prediction
  int cycles;
  float k_st, k_sh, cos_w;
  // IB + LBM (fluid.py)
  int K, half, iters, force_clamp;
  float half_width, relax, node_mass;
  float omega_p, omega_m, cp, cm;
  // synthetic IB points (fluid.py:246-263): on when synthetic_node_width > 0;
  // syn_d its float32 value
  int synthetic;
  float syn_d;
};

// One independent instance (lattice + filaments) of a plan.  A plan holds
// n_inst >= 1 instances; every per-instance array is the concatenation of the
// instances' arrays in instance order:
//   (H, W) grids (steric, density) start at grid_off;
//   a k-plane lattice array (f: 9, u* and force: 2) starts at k * fl_off,
//   each plane Hf * Wf cells;
//   nodes [node0, node0 + nodes).
struct InstGeo {
  int H, W, Hf, Wf;           // Fields.grid_shape, f_grid.shape[1:]
  float dom_hf, dom_wf;       // H, W as floats (periodic_boundary sizes)
  float scale_x, scale_y;     // fluid.py:225-234 (scale_x uses grid_shape[0])
  float umax;                 // max_fluid_velocity * mean scale (params-dependent)
  float mean_scale;           // float32 of 0.5 * (scale_x + scale_y) in double
  long long grid_off;         // first cell of the (H, W) grids
  long long fl_off;           // first cell of the lattice
  int row0, col0;             // first global steric row / column
  int lbm_blk0;               // first LBM workgroup
  int node0, nodes;           // node range
  int comp0, ncomp;           // filament (connected component) range
  int fft_w, fft_h;           // FFT plan of length W / of length H
  long long kf_off;           // transposed transfer-function table of the (H, W) shape
  unsigned long long wdiv;    // ceil(2^40 / Wf): y = (cell * wdiv) >> 40 (0: divide)
};

// Stockham plan of one FFT length: radix schedule, twiddle table offset, and
// the offset / length of the same twiddles laid out per stage.
struct FftDev {
  int n, tw_off, nrad;
  int stw_off, stw_len;
  int radix[FFT_MAX_RADICES];
};

// Device view of a plan's instances.
struct BatchDev {
  int n_inst;
  const InstGeo* geo;
  const int* node_inst;       // (N) instance of every node
  const int* comp_inst;       // (C) instance of every filament component
  const int* row_inst;        // (rows) instance of every global steric row
  const int* col_inst;        // (cols) instance of every global steric column
  const int* lbm_blk;         // (n_inst + 1) LBM workgroup offsets
  const int* ctile;           // (n_inst + 1) direct-filter column-pass tile offsets
  int ctiles;                 // total column-pass tiles
  const FftDev* fft;
  const float2* tw;           // concatenated twiddle tables
  int rows, cols, lbm_blocks; // totals over instances
  int max_w, max_h;           // largest steric grid sides
  int max_nodes;              // largest instance node count
};

constexpr int LBM_THREADS = 256;  // lattice cells per LBM workgroup
constexpr int ROLLOUT_GRAPH_STEPS = 16;  // steps of the replayed rollout graph (even)
// Output tile of the direct filter's column pass (columns x rows).
constexpr int FILTER_TILE_W = 64, FILTER_TILE_H = 32;

// Instance of item i, given per-instance offsets off[0..n] (off[0] = 0).
__device__ __forceinline__ int find_inst(const int* off, int n, int i) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= i) lo = mid; else hi = mid - 1;
  }
  return lo;
}

struct RodLayoutDev {
  int num_comp;
  const int *comp_node_off, *comp_nodes, *comp_edge_off, *comp_edges;
  const int *comp_hinge_off, *comp_hinges;
  const int *edge_src, *edge_tgt, *hinge_in, *hinge_out;
  const int *node_inc_off, *node_inc, *edge_inc_off, *edge_inc;
  // per component: the topology as one block (component-relative) in the LDS
  // order hin[nb] hout[nb] esrc[ne] etgt[ne] eoff[ne+1] einc[nei] noff[nn+1]
  // ninc[nni], and the first node / edge / hinge id when the component's ids
  // are one contiguous range (-1 otherwise)
  const int *blob, *blob_off, *node_first, *edge_first, *hinge_first;
  const struct CompMeta* meta;  // the per-component scalars in one record
};

// Per-component scalars of the rod layout, read by one 64-byte load.
struct CompMeta {
  int n0, nn, e0, ne, b0, nb, nni, nei;  // node / edge / hinge ranges, incidence counts
  int nf, ef, hf;                        // first ids of contiguous ranges (-1: listed)
  int blob_off, inst;
  int chain;                             // 1: a chain (the rod one-wavefront path)
  int pad1, pad2;
};

// Per-step rewrite of the rest values inside the rod launch.  kind 0:
// rest_b(t) = amp_b * sin(omega_b * t + phase_b); kind 1: rest lengths and
// rest angles = base + row t of the action arrays (a NULL array adds 0).
struct Actuation {
  const float *amp, *omega, *phase;
  const long long* t;        // device step counter (advanced by launch_tick)
  const long long* t_inst;   // per-instance offset added to *t (NULL: none)
  const float* d_rest;       // kind 1: (rows, e_stride) rest-length offsets or NULL
  const float* d_ang;        // kind 1: (rows, b_stride) rest-angle offsets or NULL
  long long e_stride, b_stride;
  int kind;
  int pad;                   // no tail padding: graph keys compare bytes
};

// Arguments of one LBM pass (fused collide-stream of every instance).
struct LbmArgs {
  BatchDev bd;
  const float* fin;
  float *fout, *ustar;
  uint8_t* mask;
  const float* force;
  float op, om, cp, cm;
  int force_clamp;
};

// ---- launchers (stream-ordered, no host synchronisation) ----
// steric
// Per node: wrapped prediction, CIC cell and weights; per filament: the band
// of steric rows its CIC entries touch (comp_rows); clears *arena_top.
hipError_t launch_predict(const Consts& c, const BatchDev& b, const RodLayoutDev& L,
                          const float* pos, const float* vel, float* pred_w, int* cic_cell,
                          float* cic_w, int4* comp_rows, int2* comp_cols, int* arena_top,
                          hipStream_t s);
// Raises the dynamic LDS limit of the steric kernels to lds_max - 1024 bytes.
hipError_t steric_prepare(int lds_max);
hipError_t launch_gauss_table(int H, int W, float coef, float* kfT, hipStream_t s);
// The inertial prediction of a step (launch_predict, or the IB launch that
// ends the previous step): state in, prediction and CIC set-up out.
struct PredictArgs {
  Consts c;
  BatchDev bd;
  RodLayoutDev L;
  const float* pos;
  const float* vel;
  float* pred_w;
  int* cic_cell;
  float* cic_w;
  int4* comp_rows;
  int2* comp_cols;  // per component: steric column band (first column, columns)
  int* arena_top;  // the deposits' overflow counters (rows, tiles), cleared
};
// Inputs of the CIC deposit (launch_predict's outputs and scratch).
// One CIC contribution staged by the fused deposit + filter launch: window
// cell << 2 | corner, node id, weight.
struct DepEntry {
  int pq, n;
  float w;
};
struct StericDep {
  RodLayoutDev L;
  const int4* comp_rows;
  const int2* comp_cols;
  const int* cic_cell;
  const float* cic_w;
  int* arena;                   // 2N ints
  int* arena_top;               // 2 ints (rows, tiles), cleared by launch_predict
  int* sort_tmp;                // 2N ints
  unsigned long long* key_tmp;  // 8N
  size_t key_half;              // 4N
  // the fused deposit + filter launch (NULL: the row deposit launch): its
  // windows beyond the LDS take slices of tile_arena (2 x 16N entries),
  // counted by tile_top (cleared with arena_top)
  DepEntry* tile_arena;
  int* tile_top;
};
// Steric filter (FFT path) over every instance whose table holds a transfer
// function: forward rows (CIC deposit from `dep` when real_in is NULL),
// transfer function on the transposed spectrum, inverse rows into `out`.
hipError_t launch_steric_fft_T(const BatchDev& b, const float* real_in, const StericDep* dep,
                               float2* specT, const float* kfT, float* density_out, float* out,
                               hipStream_t s);
// Steric filter (direct path) over every instance whose table holds taps
// (both axes' kernels at most FILTER_TAPS_MAX taps): the table is
// (int lo_x, n_x, lo_y, n_y), the x taps, the y taps; tmp is grid-sized
// scratch.
constexpr int FILTER_TAPS_MAX = 129;
// fused: every direct shape's kernels have at most 33 taps per axis; the
// deposit then goes to density_out (or tmp) and one launch filters both axes.
hipError_t launch_steric_conv(const BatchDev& b, const float* real_in, const StericDep* dep,
                              float* tmp, const float* taps, float* density_out, float* out,
                              bool fused, hipStream_t s);
size_t fft_rows_lds_bytes(int max_w);
size_t fft_cols_lds_bytes(int max_h);
// rod
// mode 0: full node phase of a step (prologue + constraint cycles + velocity)
// mode 1: constraint cycles only, positions in place, y = Cosserat target
// lbm (may be NULL): an LBM pass independent of this rod solve, run by extra
// workgroups of the same launch (it fills the CUs the filaments leave idle)
hipError_t launch_rod(const Consts& c, const BatchDev& b, const RodLayoutDev& L, int N,
                      size_t lds_bytes, int use_lds, float* gscratch, int mode,
                      const float* pos_in, const float* vel_in, const float* y_in,
                      const float* theta_in, const float* rest_len, const float* rest_ang,
                      const float* kb, const Actuation* act, const float* steric,
                      const float* pred_w, float* pos_out, float* vel_out, float* debug_out,
                      float* theta_out, const LbmArgs* lbm, hipStream_t s,
                      bool all_chains = false, hipEvent_t ev0 = nullptr,
                      hipEvent_t ev1 = nullptr);
// Raises the dynamic LDS limit of the rod kernels to lds_max - 1024 bytes.
hipError_t rod_prepare(int lds_max);
hipError_t launch_steric_nodes(const Consts& c, const BatchDev& b, int N, const float* steric,
                               const float* pred_w, float* force_out, hipStream_t s);
// One recorded frame of a rollout (launch_record): the frame arrays (NULL:
// not recorded), the live state, the step counter and the stride, the
// linear-resize taps of the fluid velocity (CSR per output row / column).
struct RecordDev {
  float *pos, *vel, *dbg, *steric, *fv;
  const float *s_pos, *s_vel, *s_dbg, *s_steric, *s_fv;
  const long long* counter;
  const int *ry_off, *ry_idx, *rx_off, *rx_idx;
  const float *ry_w, *rx_w;
  long long stride;
  int N, H, W, Hf, Wf, resize;
};
// Writes frame (counter + 1) / stride - 1 when (counter + 1) % stride == 0.
hipError_t launch_record(const RecordDev& r, hipStream_t s);
// jax.image.resize's linear weights from `in` to `out` samples along one axis.
void resize_taps(int in, int out, std::vector<int>& off, std::vector<int>& idx,
                 std::vector<float>& w);
// Reverse of the fluid half of one coupled step of a single-instance plan:
// from the step's input lattice fk, the rod's outputs x1 (before the wrap)
// and v1 (before the drag), and the cotangents vb (of the step's output
// velocities; replaced by that of v1) and fbn (of the output lattice), the
// cotangent fb_out of fk.  fw / iw: fluid_adj_floats / fluid_adj_ints words.
size_t fluid_adj_floats(int N, int K, int iters, long long HW);
size_t fluid_adj_ints(int N, int K, long long HW);
hipError_t fluid_adjoint_step(const Consts& c, const InstGeo& g, int N, float* fw, int* iw,
                              const float* fk, const float* x1, const float* v1, float* vb,
                              const float* fbn, float* fb_out, hipStream_t s);
hipError_t launch_tick(long long* t, hipStream_t s);
hipError_t launch_set_tick(long long* t, long long value, hipStream_t s);
size_t rod_scratch_floats(int nn, int ne, int nb);
size_t rod_lds_bytes(int nn, int ne, int nb, int nni, int nei);
// rod-only PBD step, one workgroup per filament running a whole checkpoint
// segment: forward (state advanced in place, start state to the checkpoint
// slot) or reverse (recompute from the checkpoint into the tape, then the
// adjoint, updated in place, rest-angle adjoint accumulated)
struct RodAdjArgs {
  Consts c;
  BatchDev bd;
  RodLayoutDev L;
  int N;
  const float *pos, *vel, *theta;              // forward: the state (in place)
  const float *rest_len, *rest_ang, *kb;
  float *xb, *vb, *thb;                        // adjoints: segment end in, start out
  float* rest_bar;                             // (B) += d loss / d rest angles
  // steric term (one-step segments only): fst = the step's steric force per
  // node, added to the prediction; pb1 (reverse) = output of d loss / d the
  // pushed prediction, whose steric and prologue adjoint launch_steric_bar,
  // the filter and launch_steric_adj_nodes finish (a.xb / a.vb are left to them)
  const float* fst;
  float* pb1;
  // rest-length adjoint (reverse, optional): len_bar (E) += d loss / d rest
  // lengths; utape = the Cosserat cycles' node updates (steps x cycles x N
  // float2) the reverse pass records for it
  float* len_bar;
  float* utape;
};
size_t rod_segment_lds_bytes(int cycles, int nn, int ne, int nb, bool reverse);
// forward: ck = checkpoint slot written, tape_or_dbg = debug_vector (N, 2);
// reverse: ck = checkpoint slot read, tape_or_dbg = tape (steps x cycles x E float4)
// Adjoint of the steric push p1 = p0 + f(p0) strength dt (forces.py:41-118)
// and of the prediction p0 = x + (v d) dt, given pb1 = d loss / d p1 and the
// wrapped prediction pred_w, in three launches: launch_steric_bar writes the
// field adjoint vbar (grid-sized) of the force's samples; the caller filters
// it (the Gaussian filter is its own adjoint) into dbar; then
// launch_steric_adj_nodes, with the step's field V, writes xb = d loss / d x
// and vb = d loss / d v (vb holds d loss / d v_new on entry).
// y += x over n floats.
hipError_t launch_add(float* y, const float* x, size_t n, hipStream_t s);
hipError_t launch_steric_bar(const Consts& c, const BatchDev& b, const float* pred_w,
                             const float* pb1, float* vbar, hipStream_t s);
hipError_t launch_steric_adj_nodes(const Consts& c, const BatchDev& b, int N, const float* pred_w,
                                   const float* V, const float* dbar, const float* pb1, float* xb,
                                   float* vb, hipStream_t s);
// Folding losses (mc_losses.hip): kind = mc_loss_kind; the workspace holds
// point_loss_workspace(n, m) bytes; grad may be NULL.
size_t point_loss_workspace(long long n, long long m);
// The current device's address of the library's static point-loss workspace
// (point_loss_workspace(MC_POINT_LOSS_STATIC_MAX, MC_POINT_LOSS_STATIC_MAX)
// bytes); resolved once per device, allocates nothing.
hipError_t point_loss_scratch(void** out);
hipError_t launch_point_loss(int kind, const float* particles, int n, const float* targets, int m,
                             float alpha, float* loss, float* grad, void* workspace,
                             hipStream_t s);
// Stability term of the folding loss over nf frames (nf, n, 2) against the
// main and (optional) anchor frames; g_frames NULL skips the gradients.
hipError_t launch_stability(const float* frames, int nf, int n, const float* main_pos,
                            const float* anchor_pos, int height, int width, float* loss,
                            float* g_frames, float* g_main, float* g_anchor, hipStream_t s);
hipError_t launch_rod_segment(const RodAdjArgs& a, bool reverse, int steps, float* ck,
                              float* tape_or_dbg, size_t lds_bytes, hipStream_t s);
// immersed boundary (solved per interaction cluster)
struct IBWork {
  // The IB runs on N points: the nodes, or with synthetic points (a second
  // IBWork of the plan) 2 x nodes points, point p < nodes the + point of node
  // p, point nodes + k the - point of node k.  Component layout of the points
  // (the rod layout, or each component's + points then its - points).
  int N, C;                    // IB points, filament components
  const int *comp_of_node, *comp_node_off, *comp_nodes;  // per point / component
  const struct CompMeta* meta;                           // per-component record (nn, nf, n0)
  // synthetic points: the node count, the node -> incident edge table (per
  // node its (src, tgt) pairs: every edge at its source in edge order, then at
  // its target in edge order), each point's instance, and the per-point drag
  // the solves leave for the node pass (NULL otherwise)
  int nodes;
  const int* syn_inc_off;
  const int2* syn_inc;
  const int* point_inst;
  float* syn_force;
  int *st_sx, *st_sy;          // per point wrapped stencil start cell
  float *st_w;                 // per node K*K weights
  unsigned long long* st_won;  // per node: bit t set when tap t claimed its cell
  float *vib;                  // per node (2) IB target velocity
  int *owner, *cell_seg;       // per lattice cell: claiming component (-1 idle), segment
  int2* links;                 // component pairs whose stencils share a cell
  int* nlinks;                 // links per bucket (0 between calls)
  // bucketed: one link bucket per instance (capacity: its nodes x K2, from
  // link_off[inst]) and membership by a union-find per solve workgroup; else
  // one list and the single-workgroup cluster build
  bool bucketed;
  int nbuckets;
  int* link_off;
  struct IbCtl* ictl;
  int *uf_parent, *label;      // per component: union-find parent (global fallback), cluster root
  int *cl_cnt, *cl_node_off, *cl_mem_cnt, *cl_mem_off, *cl_members;  // per cluster root
  unsigned char* cl_scratch;   // cl_bpn bytes per node, cluster-major
  size_t cl_bpn;
  size_t lds_bytes;            // dynamic LDS of the cluster solve
  // solve workgroup -> component (NULL: identity): the instances with the most
  // stencil taps per lattice cell, where the large clusters form, first
  const int* comp_order;
  bool dirty;                  // a call stopped between launches: idle state to restore
  long long cells;             // lattice cells over all instances
  float *force;                // (2, Hf, Wf) per instance: IB force, valid on the support U
  uint8_t* mask;               // U flag per lattice cell (1: the force is valid), lattice layout
  // wide clusters (more than IB_WIDE_MIN nodes) are solved together by a
  // launch of wide_wg workgroups after the cluster solve (0: no instance is
  // large enough to form one)
  int wide_wg;
  int wide_min;                // clusters with more nodes take the wide solve
  unsigned char* wide_scratch; // ib_wide_scratch_bytes(N, kmax)
  int* wide_count;             // per lattice cell entry counter (0 between calls)
  struct WideCtl* wide_ctl;
};
// Per-step counters of the cluster solve: the scratch allocator (nodes
// taken this step; cleared by the next claim launch) and the count of solve
// workgroups done with the link buckets (the last one clears the buckets).
struct IbCtl {
  int scratch_top;
  int done;
};
// Instances of at most this many filament components take the bucketed
// membership (a union-find over the instance's components in each solve
// workgroup's LDS).
constexpr int IB_UF_MAX = 8192;

// Control block of the wide solve: grid-barrier counter, wide-cluster and
// wide-node counts, segment and node counters, the per-workgroup partial sums
// of the grid scans, and the barrier-timeout record.  A barrier wait gives up
// after poll_limit polls; it then sets err (sticky until launch_ib_reset
// clears it) and *host_err (pinned host memory the host checks at its next
// call), and the workgroup leaves the solve.  While err is set the wide solve
// only wraps positions, so no later launch indexes through the counters the
// abandoned solve left behind.
struct WideCtl {
  unsigned int bar;
  int wide, wide_nodes, nseg, nq, err;
  int poll_limit;
  int* host_err;
  int bsum[IB_WIDE_WG];
};
size_t ib_wide_scratch_bytes(int N, int kmax);
size_t ib_scratch_bytes_per_node(int kmax);
size_t ib_lds_bytes(int lds_max);
hipError_t ib_prepare(int lds_max);
// Whole IB of a step: stencil + cluster links, cluster build, cluster solve
// (support, iterations, force into w.force / w.mask, node drag into vel_io,
// final wrap of pos into pos_out).  Only the solve reads f_in.  `tick` (may be
// NULL) is advanced once.  With synthetic points (w.syn_force), pos and
// vel_io are the nodes' and a last launch averages each node's two drag
// samples.  `pred` (may be NULL; not with synthetic points): the next step's inertial
// prediction from the final state (pos_out, vel_io), made by the launch that
// ends the IB (the wide solve's), so the next step needs no launch for it.
hipError_t launch_ib(const Consts& c, const BatchDev& b, IBWork& w, const float* f_in,
                     const float* pos, float* vel_io, float* pos_out, long long* tick,
                     hipStream_t s, const PredictArgs* pred = nullptr);
// Restores the idle state between calls (cell owners, union-find parents, the
// link count, the wide solve's cell counters) and clears WideCtl::err.
hipError_t launch_ib_reset(const IBWork& w, hipStream_t s);
// swimmer centre of mass / displacement tracking
struct TrackWork {
  int n_groups;       // 0: tracking off
  int start_step;     // steps up to this one do not accumulate
  int* off;           // (n_groups + 1) node-range offsets
  int2* dims;         // (n_groups) grid (H, W) of each group
  float *prev, *cum;  // (n_groups, 2) previous centre, cumulative displacement
  long long* count;   // (n_groups) steps since the reset
};
hipError_t launch_com(int n_groups, const int* off, const float* pos, int H, int W, int periodic,
                      float* com, hipStream_t s);
// mode 0: reset (prev = centre, cumulative = 0); mode 1: one step's update;
// mode 2: resume (prev = centre, step count 0, cumulative kept)
hipError_t launch_track(const TrackWork& t, const float* pos, int mode, hipStream_t s);
// LBM
LbmArgs make_lbm_args(const Consts& c, const BatchDev& b, const float* f_in, float* f_out,
                      float* ustar, uint8_t* mask, const float* force);
// start / stop (optional): events recorded with the dispatch, timing the kernel
hipError_t launch_lbm(const Consts& c, const BatchDev& b, const float* f_in, float* f_out,
                      float* ustar, uint8_t* mask, const float* force, hipStream_t s,
                      hipEvent_t start = nullptr, hipEvent_t stop = nullptr);

}  // namespace mc
