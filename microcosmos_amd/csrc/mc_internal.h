// Internal (non-exported) declarations shared by the kernel translation units
// and the C-ABI layer.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace mc {

// Tile edge (cells) of the immersed-boundary cell phase: 16 x 16 cells, one
// thread per cell, one workgroup per tile.
constexpr int IB_TILE = 16;
constexpr int IB_KMAX = 6;                 // largest supported ibm_kernel_size
constexpr int IB_TAPS_MAX = IB_KMAX * IB_KMAX;
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
// immersed boundary
struct IBWork {
  int N, T, tiles_x, tiles_y;
  int *st_sx, *st_sy;          // per node wrapped stencil start cell
  float *st_w;                 // per node K*K weights
  float *vib;                  // per node (2) IB target velocity
  float *corr;                 // per node (2) correction of the current iteration
  int *tile_cnt, *tile_off, *tile_fill, *tile_active, *n_active, *tile_list;
  int *sort_scratch;
  float *rho, *u0, *ucorr, *inv, *force;   // dense (H*W) / (2*H*W) on the fluid grid
  uint32_t* mask;              // per fluid row ceil(W/32) words
  int mask_words_per_row;
  int2* cell_span;             // per covered cell: (first entry, entry count)
  int2* entries;               // ordered (node, weight bits) lists of covered cells
};
// Immersed boundary in two parts: the f-independent binning (stencil, tile
// scan, fill), then the support / iterations / force (reads f_in).  The drag
// and final wrap run separately so the LBM pass can overlap them; `tick`
// (may be NULL) is advanced once by the drag launch.
hipError_t launch_ib_bin(const Consts& c, IBWork& w, const float* pos, float* vel_io,
                         hipStream_t s);
hipError_t launch_ib_solve(const Consts& c, IBWork& w, const float* f_in, hipStream_t s);
hipError_t launch_ib_drag(const Consts& c, IBWork& w, const float* pos, float* vel_io,
                          float* pos_out, long long* tick, hipStream_t s);
// LBM
hipError_t launch_lbm(const Consts& c, const float* f_in, float* f_out, float* ustar,
                      uint32_t* mask, int mask_words_per_row, const float* force, hipStream_t s);

}  // namespace mc
