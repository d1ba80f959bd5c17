// Internal (non-exported) declarations shared by the kernel translation units
// and the C-ABI layer.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace mc {

constexpr int IB_KMAX = 6;                 // largest supported ibm_kernel_size
constexpr int SORT_LDS_CAP = 2048;         // bitonic sort capacity in LDS (ids)

// Float constants derived from ConstraintSolverConfig exactly as the reference
// derives them (Python double arithmetic, then a float32 weak scalar).
struct Consts {
  // step (simulate.py:37-100)
  float dt, damping;
  int enable_fluid, enable_steric;
  int dom_h, dom_w;           // Fields.grid_shape
  float dom_hf, dom_wf;       // as floats
  // steric (forces.py)
  float steric_strength, steric_sv, steric_2s2, steric_norm, steric_svs, fft_coef;
  int steric_nskip;
  // rod (pbd.py / stable_cosserat.py)
  int solver;                 // 0 pbd, 1 cosserat
  int cycles;
  float k_st, k_sh, cos_w;
  // IB + LBM (fluid.py)
  int K, half, iters, force_clamp;
  float half_width, relax, umax, scale_x, scale_y, sdt_x, sdt_y, node_mass;
  float omega_p, omega_m, cp, cm;
  int fl_h, fl_w;             // f_grid spatial shape
};

struct RodLayoutDev {
  int num_comp;
  const int *comp_node_off, *comp_nodes, *comp_edge_off, *comp_edges;
  const int *comp_hinge_off, *comp_hinges;
  const int *edge_src, *edge_tgt, *hinge_in, *hinge_out;
  const int *node_inc_off, *node_inc, *edge_inc_off, *edge_inc;
};

struct Actuation {           // rest_b(t) = amp_b * sin(omega_b * t + phase_b)
  const float *amp, *omega, *phase;
  const long long* t;        // device step counter (advanced by launch_tick)
};

// ---- launchers (stream-ordered, no host synchronisation) ----
// steric
hipError_t launch_predict(const Consts& c, int N, const float* pos, const float* vel,
                          float* pred_w, int* cic_cell, float* cic_w, int* row_cnt,
                          hipStream_t s);
hipError_t launch_bin_scan(int T, const int* cnt, int* off, int* fill, int* active,
                           int* n_active, hipStream_t s);
hipError_t launch_row_fill(const Consts& c, int N, const int* cic_cell, const int* row_off,
                           int* row_fill, int* row_list, hipStream_t s);
hipError_t launch_deposit_rows_fft(const Consts& c, int N, const int* row_off, int* row_list,
                                   const int* cic_cell, const float* cic_w, float2* spec,
                                   const float2* tw_w, const int* radix_w, int nrad_w,
                                   float* density_out, int* sort_scratch,
                                   unsigned long long* key_scratch, hipStream_t s);
hipError_t launch_fft_rows(int H, int W, const float* real_in, float2* spec,
                           const float2* tw_w, const int* radix_w, int nrad_w, hipStream_t s);
hipError_t launch_gauss_table(int H, int W, float coef, float* kf, hipStream_t s);
// Whole steric filter: forward rows (CIC deposit when real_in is NULL),
// transfer function on the transposed spectrum, inverse rows into `out`.
int fft_rows_per_group(int H, int W);
hipError_t launch_steric_fft_T(int H, int W, const float* real_in, const int* row_off, int* row_list,
                               int* sort_scratch, unsigned long long* key_scratch, size_t key_half,
                               const int* cic_cell, const float* cic_w, float2* specT,
                               const float2* tw_w, const int* radix_w, int nrad_w,
                               const float2* tw_h, const int* radix_h, int nrad_h, const float* kfT,
                               float* density_out, float* out, hipStream_t s);
hipError_t launch_fft_cols_filter(int H, int W, float2* spec, const float2* tw_h,
                                  const int* radix_h, int nrad_h, const float* kf, int lds_bytes_max,
                                  hipStream_t s);
hipError_t launch_ifft_rows_real(int H, int W, const float2* spec, float* out,
                                 const float2* tw_w, const int* radix_w, int nrad_w,
                                 hipStream_t s);
// rod
// mode 0: full node phase of a step (prologue + constraint cycles + velocity)
// mode 1: constraint cycles only, positions in place, y = Cosserat target
hipError_t launch_rod(const Consts& c, const RodLayoutDev& L, int N, size_t lds_bytes,
                      int use_lds, float* gscratch, int mode, const float* pos_in,
                      const float* vel_in, const float* y_in, const float* theta_in,
                      const float* rest_len, const float* rest_ang, const float* kb,
                      const Actuation* act, const float* steric, const float* pred_w,
                      float* pos_out, float* vel_out, float* debug_out, float* theta_out,
                      hipStream_t s);
hipError_t launch_steric_nodes(const Consts& c, int N, const float* steric, const float* pred_w,
                               float* force_out, hipStream_t s);
hipError_t launch_tick(long long* t, hipStream_t s);
size_t rod_scratch_floats(int nn, int ne, int nb);
size_t rod_lds_bytes(int nn, int ne, int nb, int nni, int nei);
// immersed boundary (solved per interaction cluster)
struct IBWork {
  int N, C;                    // nodes, filament components
  const int *comp_of_node, *comp_node_off, *comp_nodes;  // component layout (rod layout)
  int *st_sx, *st_sy;          // per node wrapped stencil start cell
  float *st_w;                 // per node K*K weights
  unsigned long long* st_won;  // per node: bit t set when tap t claimed its cell
  float *vib;                  // per node (2) IB target velocity
  int *owner, *cell_seg;       // per lattice cell: claiming component (-1 idle), segment
  int2* links;                 // component pairs whose stencils share a cell
  int* nlinks;                 // number of links (0 between calls)
  int *uf_parent, *label;      // per component: union-find parent (global fallback), cluster root
  int *cl_cnt, *cl_node_off, *cl_mem_cnt, *cl_mem_off, *cl_members;  // per cluster root
  unsigned char* cl_scratch;   // cl_bpn bytes per node, cluster-major
  size_t cl_bpn;
  size_t lds_bytes;            // dynamic LDS of the cluster solve
  bool dirty;                  // a call stopped between launches: idle state to restore
  float *force;                // (2, H, W) IB force, valid on the support U
  uint32_t* mask;              // per lattice row ceil(W/32) words: U bits
  int mask_words_per_row;
};
size_t ib_scratch_bytes_per_node(int kmax);
size_t ib_lds_bytes(int lds_max);
hipError_t ib_prepare(int lds_max);
// Whole IB of a step: stencil + cluster links, cluster build, cluster solve
// (support, iterations, force into w.force / w.mask, node drag into vel_io,
// final wrap of pos into pos_out).  Only the solve reads f_in.  `tick` (may be
// NULL) is advanced once.
hipError_t launch_ib(const Consts& c, IBWork& w, const float* f_in, const float* pos,
                     float* vel_io, float* pos_out, long long* tick, hipStream_t s);
// swimmer centre of mass / displacement tracking
struct TrackWork {
  int n_groups;       // 0: tracking off
  int start_step;     // steps up to this one do not accumulate
  int* off;           // (n_groups + 1) node-range offsets
  float *prev, *cum;  // (n_groups, 2) previous centre, cumulative displacement
  long long* count;   // (n_groups) steps since the reset
};
hipError_t launch_com(int n_groups, const int* off, const float* pos, int H, int W, int periodic,
                      float* com, hipStream_t s);
// mode 0: reset (prev = centre, cumulative = 0); mode 1: one step's update
hipError_t launch_track(const TrackWork& t, const float* pos, int H, int W, int mode,
                        hipStream_t s);
// LBM
hipError_t launch_lbm(const Consts& c, const float* f_in, float* f_out, float* ustar,
                      uint32_t* mask, int mask_words_per_row, const float* force, hipStream_t s);

}  // namespace mc
