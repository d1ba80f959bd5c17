// Shared device helpers for the MI355X (gfx950) coupled fluid-filament step.
//
// Floating-point policy: every kernel is compiled with -ffp-contract=off and
// with IEEE division/sqrt, so each source-level operation rounds once, in the
// order written -- the order of the reference's JAX expressions.
// cos/sin/atan2/exp are evaluated in double and rounded to float (a correctly
// rounded float result).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define MC_EPSIL 1.1920928955078125e-07f  // utils.py:9, float32 machine epsilon
#define MC_PI_F 3.14159274101257324f        // float32(pi) (weak-typed jnp.pi)
#define MC_2PI_F 6.28318548202514648f       // float32(2 * pi)

namespace mc {

// numpy / jnp.maximum and minimum semantics (NaN propagates) where the second
// operand is a finite constant: one compare and one select each.
__device__ __forceinline__ float fmax_nan(float a, float c) { return (a < c) ? c : a; }
__device__ __forceinline__ float fmin_nan(float c, float a) { return (a >= c) ? c : a; }

// Floor-mod of floats (jnp.mod / np.mod): fmod, then add y when the signs differ.
__device__ __forceinline__ float fmodf_floor_slow(float x, float y) {
  float r = fmodf(x, y);
  if (r != 0.0f) {
    if ((r < 0.0f) != (y < 0.0f)) r = r + y;
  } else {
    r = copysignf(0.0f, y);
  }
  return r;
}

// Same result for a positive modulus y, without fmodf on the common ranges:
// for x in [0, y) fmod is x; for x in [y, 2y) it is x - y (exact, Sterbenz);
// for x in [-y, 0) fmod is x and the sign fix-up adds y (one rounding, as in
// np.mod).  A zero result is +0, as np.mod returns.
__device__ __forceinline__ float fmodf_floor(float x, float y) {
  float r;
  if (x >= 0.0f) {
    if (x < y) r = x;
    else if (x < 2.0f * y) r = x - y;
    else return fmodf_floor_slow(x, y);
  } else if (x >= -y) {
    r = x + y;
  } else {
    return fmodf_floor_slow(x, y);
  }
  return r == 0.0f ? 0.0f : r;
}

// Floor-mod of an int already known to lie in [-m, 2m).
__device__ __forceinline__ int wrap1(int a, int m) { return a < 0 ? a + m : (a >= m ? a - m : a); }

// Floor-mod of ints (Python/numpy semantics for a positive modulus).
__device__ __forceinline__ int imod(int a, int m) {
  int r = a % m;
  return r < 0 ? r + m : r;
}

// float -> int32 truncation of an integer-valued float; saturates and maps NaN
// to 0 so no index computed from a corrupt position can leave the grid after
// the floor-mod.
__device__ __forceinline__ int f2i(float v) {
  if (!(v == v)) return 0;
  v = fminf(fmaxf(v, -2.0e9f), 2.0e9f);
  return (int)v;
}

__device__ __forceinline__ float dcos(float x) { return (float)cos((double)x); }
__device__ __forceinline__ float dsin(float x) { return (float)sin((double)x); }
__device__ __forceinline__ float dexp(float x) { return (float)exp((double)x); }
__device__ __forceinline__ float datan2(float y, float x) { return (float)atan2((double)y, (double)x); }

// utils.py:18-23 minimum-image displacement along one axis of size s.
__device__ __forceinline__ float min_image(float a, float b, float s) {
  float half = s / 2.0f;
  return fmodf_floor(a - b + half, s) - half;
}

// pbd.py:38 angle wrap (d + pi) mod 2 pi - pi, floor-mod.
__device__ __forceinline__ float wrap_angle(float d) {
  return fmodf_floor(d + MC_PI_F, MC_2PI_F) - MC_PI_F;
}

// D2Q9 lattice (fluid.py:15-23).
__constant__ constexpr int EX[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
__constant__ constexpr int EY[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
__constant__ constexpr int OPP[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6};

// Density and clamped velocity of one cell from floored populations
// (fluid.py:136-142,268-274).  The LBM pass and the immersed-boundary kernels
// share this definition so both see bit-identical values.
__device__ __forceinline__ void lbm_macro(const float f[9], float umax, float& rho, float& ux,
                                          float& uy) {
  float r = f[0];
#pragma unroll
  for (int i = 1; i < 9; ++i) r = r + f[i];
  rho = fmax_nan(r, 0.1f);
  float mx = ((((f[1] - f[3]) + f[5]) - f[6]) - f[7]) + f[8];
  float my = ((((f[2] - f[4]) + f[5]) + f[6]) - f[7]) - f[8];
  ux = mx / rho;
  uy = my / rho;
  float mag = sqrtf((ux * ux + uy * uy) + MC_EPSIL);
  float cl = fmin_nan(1.0f, umax / (mag + MC_EPSIL));
  ux = ux * cl;
  uy = uy * cl;
}

}  // namespace mc
