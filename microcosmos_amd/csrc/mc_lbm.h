// Fused D2Q9 TRT + Guo collide-and-stream update of one lattice cell
// (fluid.py:268-357), shared by the standalone LBM kernel and the rod launch
// that carries the previous step's LBM pass.
//
// One thread per cell: it reads the nine populations of its cell once
// (structure-of-arrays, one coalesced 256 B segment per wavefront per
// population), recomputes density and velocity, takes the immersed-boundary
// force from the sparse support U (or the clamp-only force off U), collides,
// and pushes the nine post-collision values to their streamed destinations.
// Workgroups are mapped so each XCD owns a contiguous band of rows: the +-1
// row and column shifts of the push then land in lines written by the same
// XCD, except at band edges.  The reference's state contract stores
// post-stream populations, so a push of locally collided values is the
// streaming pass with no redundant collision.
//
// HBM traffic per cell: 36 B read + 36 B written (f) + 8 B written (u*)
// (+ 1 B of the U flags, + 8 B read and 1 B written only on the ~6 % of
// cells in U).
#pragma once
#include "mc_common.h"
#include "mc_internal.h"

namespace mc {
namespace lbm_detail {

constexpr float W0 = (float)(4.0 / 9.0);    // fluid.py:20, as float32
constexpr float W1 = (float)(1.0 / 9.0);
constexpr float W5 = (float)(1.0 / 36.0);

__device__ __forceinline__ float lw(int i) { return i == 0 ? W0 : (i < 5 ? W1 : W5); }

// e_i . v with e components in {-1, 0, 1}: the same expression the reference
// evaluates ((0 + ex vx) + ey vy), written per direction.
__device__ __forceinline__ float edot(int i, float vx, float vy) {
  switch (i) {
    case 0: return 0.0f;
    case 1: return vx;
    case 2: return vy;
    case 3: return -vx;
    case 4: return -vy;
    case 5: return vx + vy;
    case 6: return -vx + vy;
    case 7: return -vx - vy;
    default: return vx - vy;
  }
}

}  // namespace lbm_detail

// Cell t (0 .. LBM_THREADS-1) of LBM workgroup b (instance-major numbering).
__device__ __forceinline__ void lbm_cell(const LbmArgs& a, int b, int t) {
  using namespace lbm_detail;
  const int inst = find_inst(a.bd.lbm_blk, a.bd.n_inst, b);
  const InstGeo& g = a.bd.geo[inst];
  const int H = g.Hf, W = g.Wf;
  const size_t HW = (size_t)H * W;
  // LBM_THREADS consecutive cells of the instance's lattice per workgroup
  const int ci = (b - g.lbm_blk0) * LBM_THREADS + t;
  if (ci >= (int)HW) return;
  // row of the cell: exact reciprocal multiply when cells * Wf < 2^40 (set on
  // the host), else a division
  const int y = g.wdiv ? (int)(((unsigned long long)ci * g.wdiv) >> 40) : ci / W;
  const int x = ci - y * W;
  const float umax = g.umax;
  const float* fin = a.fin + 9 * g.fl_off;
  float* fout = a.fout + 9 * g.fl_off;
  float* ustar = a.ustar + 2 * g.fl_off;
  const float* force = a.force + 2 * g.fl_off;
  uint8_t* mask = a.mask + g.fl_off;
  const int yp = (y + 1 == H) ? 0 : y + 1;
  const int ym = (y == 0) ? H - 1 : y - 1;
  const size_t cell = (size_t)y * W + x;
  // the populations are read once: non-temporal loads keep them from
  // displacing the lines the pushes write into
  float f[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) f[i] = fmax_nan(__builtin_nontemporal_load(&fin[i * HW + cell]), MC_EPSIL);
  // the cell's U flag (set by the IB solve that wrote its force)
  const uint8_t m = mask[cell];
  // density and clamped velocity (fluid.py:136-142,268-274); `cl1` tells
  // whether the clamp changed u
  float r = f[0];
#pragma unroll
  for (int i = 1; i < 9; ++i) r = r + f[i];
  const float rho = fmax_nan(r, 0.1f);
  float ux = (((((f[1] - f[3]) + f[5]) - f[6]) - f[7]) + f[8]) / rho;
  float uy = (((((f[2] - f[4]) + f[5]) + f[6]) - f[7]) - f[8]) / rho;
  const float mag1 = sqrtf((ux * ux + uy * uy) + MC_EPSIL);
  const float cl1 = fmin_nan(1.0f, umax / (mag1 + MC_EPSIL));
  ux = ux * cl1;
  uy = uy * cl1;
  float Fx, Fy;
  bool zero_force = false;
  if (m) {
    Fx = force[cell];
    Fy = force[HW + cell];
  } else if (cl1 == 1.0f) {
    // Off the support u_corrected == u; when the first clamp left u unchanged
    // the post-clamp sees the same magnitude, so its factor is 1 too and
    // F = (u - u) rho = +0 exactly (fluid.py:327-332).
    Fx = 0.0f;
    Fy = 0.0f;
    zero_force = true;
  } else {
    const float mag = sqrtf((ux * ux + uy * uy) + MC_EPSIL);
    const float cl = fmin_nan(1.0f, umax / (mag + MC_EPSIL));
    const float cx = ux * cl, cy = uy * cl;
    Fx = (cx - ux) * rho;
    Fy = (cy - uy) * rho;
    if (a.force_clamp) {
      const float fm = sqrtf((Fx * Fx + Fy * Fy) + MC_EPSIL);
      const float fc = fmin_nan(1.0f, umax / (fm + MC_EPSIL));
      Fx = Fx * fc;
      Fy = Fy * fc;
    }
  }
  if (m) mask[cell] = 0;  // consume the U flag
  // u* = u + 0.5 F / rho (fluid.py:346); with F = +0 the quotient is +0
  const float usx = zero_force ? ux + 0.0f : ux + (0.5f * Fx) / rho;
  const float usy = zero_force ? uy + 0.0f : uy + (0.5f * Fy) / rho;
  const float usq = usx * usx + usy * usy;
  float feq[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    const float eu = edot(i, usx, usy);
    feq[i] = (rho * lw(i)) * (((1.0f + 3.0f * eu) + 4.5f * (eu * eu)) - 1.5f * usq);
  }
  float post[9];
  if (zero_force) {
    // the TRT Guo source of a zero force is a signed zero: adding it is exact
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      const int o = OPP[i];
      const float coll = (f[i] - a.op * (0.5f * (f[i] + f[o]) - 0.5f * (feq[i] + feq[o]))) -
                         a.om * (0.5f * (f[i] - f[o]) - 0.5f * (feq[i] - feq[o]));
      post[i] = fmax_nan(coll, MC_EPSIL);
    }
  } else {
    const float udf = usx * Fx + usy * Fy;
    float raw[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      const float eu = edot(i, usx, usy);
      const float ef = edot(i, Fx, Fy);
      raw[i] = lw(i) * (3.0f * (ef - udf) + (9.0f * eu) * ef);
    }
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      const int o = OPP[i];
      const float coll = (f[i] - a.op * (0.5f * (f[i] + f[o]) - 0.5f * (feq[i] + feq[o]))) -
                         a.om * (0.5f * (f[i] - f[o]) - 0.5f * (feq[i] - feq[o]));
      const float src = a.cp * (0.5f * (raw[i] + raw[o])) + a.cm * (0.5f * (raw[i] - raw[o]));
      post[i] = fmax_nan(coll + src, MC_EPSIL);
    }
  }
  const int xp = (x + 1 == W) ? 0 : x + 1;
  const int xm = (x == 0) ? W - 1 : x - 1;
  const size_t ry = (size_t)y * W, ryp = (size_t)yp * W, rym = (size_t)ym * W;
  fout[0 * HW + ry + x] = post[0];
  fout[1 * HW + ry + xp] = post[1];
  fout[2 * HW + ryp + x] = post[2];
  fout[3 * HW + ry + xm] = post[3];
  fout[4 * HW + rym + x] = post[4];
  fout[5 * HW + ryp + xp] = post[5];
  fout[6 * HW + ryp + xm] = post[6];
  fout[7 * HW + rym + xm] = post[7];
  fout[8 * HW + rym + xp] = post[8];
  ustar[cell] = usx;
  ustar[HW + cell] = usy;
}

}  // namespace mc
