// Shared device helpers for the MI355X (gfx950) coupled fluid-filament step.
//
// Floating-point policy: every kernel is compiled with -ffp-contract=off and
// with IEEE division/sqrt, so each source-level operation rounds once, in the
// order written -- the order of the reference's JAX expressions.
// The filament angles' sin/cos, the IB weights' cos and the steric window's
// exp use the float32 routines below (fsincos, fexp); atan2 and the other
// transcendentals are evaluated in double and rounded to float.
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
// np.mod).  A zero result is +0, as np.mod returns.  The common ranges are
// selects, so independent calls overlap; fmodf runs only for lanes outside
// them (and NaN).
__device__ __forceinline__ float fmodf_floor(float x, float y) {
  const float below = x + y, above = x - y;
  float r = x < 0.0f ? below : (x >= y ? above : x);
  r = r == 0.0f ? 0.0f : r;
  if (!(x >= -y && x < 2.0f * y)) r = fmodf_floor_slow(x, y);
  return r;
}

// fmodf_floor's common ranges alone: exact for x in [-y, 2y); sets `slow` for
// a lane outside them (or NaN), whose caller then recomputes with
// fmodf_floor.  Lets a chain of independent wraps run without a branch each.
__device__ __forceinline__ float fmodf_floor_fast(float x, float y, bool& slow) {
  const float below = x + y, above = x - y;
  const float r = x < 0.0f ? below : (x >= y ? above : x);
  slow = slow || !(x >= -y && x < 2.0f * y);
  return r == 0.0f ? 0.0f : r;
}

// Floor-mod of an int already known to lie in [-m, 2m).
__device__ __forceinline__ int wrap1(int a, int m) { return a < 0 ? a + m : (a >= m ? a - m : a); }

// Floor-mod of an int that every finite input of the caller keeps in
// [-m, 2m) (selects, no division); it is clamped there first, so a corrupt
// (non-finite) input still yields an index in [0, m).
__device__ __forceinline__ int wrap1c(int a, int m) { return wrap1(min(max(a, -m), 2 * m - 1), m); }

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

// Float32 sine and cosine for the latency-bound chains (the filament cycles,
// the IB weights): Cody-Waite reduction by pi/2 (a three-part split of pi/2
// whose products with the quadrant count are exact for |x| <= 8192) and the
// Taylor series of sin to x^9 and cos to x^10 on [-pi/4, pi/4], in Horner
// form.  Every step is one IEEE float32 operation (the build contracts
// nothing), so the NumPy restatement reproduces it bit for bit.  Absolute
// error below 1e-7 (within 2 ulp away from the zeros of the function).
// Beyond |x| = 8192 the double evaluation is used.
__host__ __device__ __forceinline__ void fsincos(float x, float* s, float* c) {
  const float k = rintf(x * 0x1.45f306p-1f);  // float32(2 / pi)
  const float r = ((x - k * 0x1.92p+0f) - k * 0x1.fb4p-12f) - k * 0x1.4442d2p-24f;
  const float z = r * r;
  const float ps =
      r + (r * z) * (-0x1.555556p-3f + z * (0x1.111112p-7f + z * (-0x1.a01a02p-13f + z * 0x1.71de3ap-19f)));
  const float pc = (1.0f - 0.5f * z) +
                   (z * z) * (0x1.555556p-5f + z * (-0x1.6c16c2p-10f + z * (0x1.a01a02p-16f + z * -0x1.27e4fcp-22f)));
  const int q = (k == k) ? ((int)k & 3) : 0;
  float sv = q == 0 ? ps : (q == 1 ? pc : (q == 2 ? -ps : -pc));
  float cv = q == 0 ? pc : (q == 1 ? -ps : (q == 2 ? -pc : ps));
  if (fabsf(x) > 8192.0f) {
    double sd, cd;
    sincos((double)x, &sd, &cd);
    sv = (float)sd;
    cv = (float)cd;
  }
  *s = sv;
  *c = cv;
}
__host__ __device__ __forceinline__ float fcos(float x) {
  float s, c;
  fsincos(x, &s, &c);
  return c;
}

// Float32 exp: Cody-Waite reduction by ln 2 (a two-part split exact for the
// clamped range), the Taylor series to x^7 on [-ln2 / 2, ln2 / 2], scaled by
// 2^n with one rounding (ldexp, denormal results included).  The argument is
// clamped to [-104, 88.5] (exp(-104) rounds to +0); NaN propagates.  Within
// 1.2 ulp; reproduced bit for bit by the NumPy restatement.
__host__ __device__ __forceinline__ float fexp(float x) {
  x = x < -104.0f ? -104.0f : (x > 88.5f ? 88.5f : x);
  const float n = rintf(x * 0x1.715476p+0f);  // float32(1 / ln 2)
  const float r = (x - n * 0x1.63p-1f) - n * -0x1.bd0106p-13f;
  const float p =
      1.0f + r * (1.0f + r * (0.5f + r * (0x1.555556p-3f + r * (0x1.555556p-5f +
                  r * (0x1.111112p-7f + r * (0x1.6c16c2p-10f + r * 0x1.a01a02p-13f))))));
  return ldexpf(p, (n == n) ? (int)n : 0);
}

// utils.py:18-23 minimum-image displacement along one axis of size s.
__device__ __forceinline__ float min_image(float a, float b, float s) {
  float half = s / 2.0f;
  return fmodf_floor(a - b + half, s) - half;
}

// pbd.py:38 angle wrap (d + pi) mod 2 pi - pi, floor-mod.
__device__ __forceinline__ float wrap_angle(float d) {
  return fmodf_floor(d + MC_PI_F, MC_2PI_F) - MC_PI_F;
}

// min_image / wrap_angle on the common ranges (see fmodf_floor_fast).
__device__ __forceinline__ float min_image_fast(float a, float b, float s, bool& slow) {
  float half = s / 2.0f;
  return fmodf_floor_fast(a - b + half, s, slow) - half;
}
__device__ __forceinline__ float wrap_angle_fast(float d, bool& slow) {
  return fmodf_floor_fast(d + MC_PI_F, MC_2PI_F, slow) - MC_PI_F;
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
