// Fused D2Q9 TRT + Guo collide-and-stream pass (fluid.py:268-357).
//
// One workgroup owns whole lattice rows: it reads the nine populations of each
// cell once (structure-of-arrays, one coalesced 256 B segment per wavefront per
// population), recomputes density and velocity, takes the immersed-boundary
// force from the sparse support U (or the clamp-only force off U), collides,
// and pushes the nine post-collision values to their streamed destinations.
// Because a workgroup owns entire rows, every destination row segment of every
// population is written by exactly one workgroup (the +-1 x shift wraps inside
// the row), so no cache line is shared between workgroups on different XCDs.
// The reference's state contract stores post-stream populations, so a push of
// locally collided values is the streaming pass with no redundant collision.
//
// HBM traffic per cell: 36 B read + 36 B written (f) + 8 B written (u*)
// (+ 1/8 B of the U bitmask, + 8 B only on the ~6 % of cells in U).
#include "mc_common.h"
#include "mc_internal.h"

namespace mc {
namespace {

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

}  // namespace

__global__ void __launch_bounds__(256) k_lbm(const float* __restrict__ fin, float* __restrict__ fout,
                                             float* __restrict__ ustar, uint32_t* __restrict__ mask,
                                             int mwpr, const float* __restrict__ force, int H,
                                             int W, int rows_per_block, float umax, float op,
                                             float om, float cp, float cm, int force_clamp) {
  const size_t HW = (size_t)H * W;
  const int y0 = blockIdx.x * rows_per_block;
  for (int yy = 0; yy < rows_per_block; ++yy) {
    const int y = y0 + yy;
    if (y >= H) break;
    const int yp = (y + 1 == H) ? 0 : y + 1;
    const int ym = (y == 0) ? H - 1 : y - 1;
    for (int x = threadIdx.x; x < W; x += blockDim.x) {
      const size_t cell = (size_t)y * W + x;
      float f[9];
#pragma unroll
      for (int i = 0; i < 9; ++i) f[i] = fmax_nan(fin[i * HW + cell], MC_EPSIL);
      float rho, ux, uy;
      lbm_macro(f, umax, rho, ux, uy);
      uint32_t* mw = mask + (size_t)y * mwpr + (x >> 5);
      const uint32_t m = *mw;
      float Fx, Fy;
      if ((m >> (x & 31)) & 1u) {
        Fx = force[cell];
        Fy = force[HW + cell];
      } else {
        // u_corrected == u off the support: post-clamp and force (fluid.py:327-339)
        float mag = sqrtf((ux * ux + uy * uy) + MC_EPSIL);
        float cl = fmin_nan(1.0f, umax / (mag + MC_EPSIL));
        float cx = ux * cl, cy = uy * cl;
        Fx = (cx - ux) * rho;
        Fy = (cy - uy) * rho;
        if (force_clamp) {
          float fm = sqrtf((Fx * Fx + Fy * Fy) + MC_EPSIL);
          float fc = fmin_nan(1.0f, umax / (fm + MC_EPSIL));
          Fx = Fx * fc;
          Fy = Fy * fc;
        }
      }
      if ((x & 31) == 0) *mw = 0u;  // consume the U bitmask word (every lane has read it)
      // u* = u + 0.5 F / rho (fluid.py:346)
      const float usx = ux + (0.5f * Fx) / rho;
      const float usy = uy + (0.5f * Fy) / rho;
      const float usq = usx * usx + usy * usy;
      const float udf = usx * Fx + usy * Fy;
      float feq[9], raw[9];
#pragma unroll
      for (int i = 0; i < 9; ++i) {
        const float eu = edot(i, usx, usy);
        const float ef = edot(i, Fx, Fy);
        feq[i] = (rho * lw(i)) * (((1.0f + 3.0f * eu) + 4.5f * (eu * eu)) - 1.5f * usq);
        raw[i] = lw(i) * (3.0f * (ef - udf) + (9.0f * eu) * ef);
      }
      float post[9];
#pragma unroll
      for (int i = 0; i < 9; ++i) {
        const int o = OPP[i];
        const float coll = (f[i] - op * (0.5f * (f[i] + f[o]) - 0.5f * (feq[i] + feq[o]))) -
                           om * (0.5f * (f[i] - f[o]) - 0.5f * (feq[i] - feq[o]));
        const float src = cp * (0.5f * (raw[i] + raw[o])) + cm * (0.5f * (raw[i] - raw[o]));
        post[i] = fmax_nan(coll + src, MC_EPSIL);
      }
      const int xp = (x + 1 == W) ? 0 : x + 1;
      const int xm = (x == 0) ? W - 1 : x - 1;
      const size_t ry = (size_t)y * W, ryp = (size_t)yp * W, rym = (size_t)ym * W;
      *(fout + 0 * HW + ry + x) = post[0];
      *(fout + 1 * HW + ry + xp) = post[1];
      *(fout + 2 * HW + ryp + x) = post[2];
      *(fout + 3 * HW + ry + xm) = post[3];
      *(fout + 4 * HW + rym + x) = post[4];
      *(fout + 5 * HW + ryp + xp) = post[5];
      *(fout + 6 * HW + ryp + xm) = post[6];
      *(fout + 7 * HW + rym + xm) = post[7];
      *(fout + 8 * HW + rym + xp) = post[8];
      *(ustar + cell) = usx;
      *(ustar + HW + cell) = usy;
    }
  }
}

hipError_t launch_lbm(const Consts& c, const float* f_in, float* f_out, float* ustar,
                      uint32_t* mask, int mwpr, const float* force, hipStream_t s) {
  const int H = c.fl_h, W = c.fl_w;
  int block = 256;
  if (W < 256) block = ((W + 63) / 64) * 64;
  // at least ~1024 cells per workgroup
  int rpb = 1;
  while (rpb * W < 1024 && rpb < H) rpb *= 2;
  const int grid = (H + rpb - 1) / rpb;
  hipLaunchKernelGGL(k_lbm, dim3(grid), dim3(block), 0, s, f_in, f_out, ustar, mask, mwpr, force,
                     H, W, rpb, c.umax, c.omega_p, c.omega_m, c.cp, c.cm, c.force_clamp);
  return hipGetLastError();
}

}  // namespace mc
