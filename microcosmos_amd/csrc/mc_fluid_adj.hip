// Adjoint of the fluid half of one coupled step (fluid.py:202-386): the
// immersed-boundary forcing with the node positions held constant (their
// stop_gradient, fluid.py:239) and the D2Q9 TRT + Guo collide-and-stream.
//
// Forward quantities are recomputed from the step's input lattice f_k and
// the rod's output (positions x', velocities v'):
//   f0 = max(f_k, eps); rho = max(sum f0, 0.1); ur = (sum e f0) / rho;
//   u = clamp(ur); uc_0 = u;
//   per iteration j:  U_n = sum_t w_nt uc_j[c_nt];  d_n = (vib_n - U_n) r;
//                     cc_n = clamp(d_n);  uc_{j+1}[c] = uc_j[c] + ikw[c] sum w cc;
//   F = [clamp]((clamp(uc_J) - u) rho);  v'' = v' - (sum_t w F) / m / (scale dt);
//   f_{k+1,i}[x + e_i] = max(collide_i(f0, rho, u, F)[x], eps)
// with clamp(x) = x min(1, A / (sqrt(|x|^2 + eps) + eps)), A = u_max.
// Given the cotangents of f_{k+1} and v'', the reverse pass produces those of
// f_k and v'.  Every sum over a cell's entries (node, tap) runs in ascending
// entry order through a per-cell list, and every per-node sum in tap order,
// so the result does not depend on scheduling; no float atomics.
#include "mc_common.h"
#include "mc_claim.h"
#include "mc_internal.h"

namespace mc {
namespace {

constexpr int FA_THREADS = 256;

__device__ __forceinline__ float lwt(int i) {
  return i == 0 ? (float)(4.0 / 9.0) : (i < 5 ? (float)(1.0 / 9.0) : (float)(1.0 / 36.0));
}
__device__ __forceinline__ int ex_of(int i) {
  return (i == 1 || i == 5 || i == 8) ? 1 : ((i == 3 || i == 6 || i == 7) ? -1 : 0);
}
__device__ __forceinline__ int ey_of(int i) {
  return (i == 2 || i == 5 || i == 6) ? 1 : ((i == 4 || i == 7 || i == 8) ? -1 : 0);
}

// d max(a, b) / d a: 1 above, 0 below, 1/2 at a tie (jax's rule).
__device__ __forceinline__ float dmax(float a, float b) {
  return a > b ? 1.0f : (a == b ? 0.5f : 0.0f);
}

__device__ __forceinline__ float2 clamp2(float2 x, float A) {
  const float m = sqrtf((x.x * x.x + x.y * x.y) + MC_EPSIL);
  const float s = fmin_nan(1.0f, A / (m + MC_EPSIL));
  return make_float2(x.x * s, x.y * s);
}

// Cotangent of x for y = clamp(x) given that of y.
__device__ __forceinline__ float2 clamp2_bar(float2 x, float2 yb, float A) {
  const float m = sqrtf((x.x * x.x + x.y * x.y) + MC_EPSIL);
  const float q = A / (m + MC_EPSIL);
  if (q >= 1.0f) return yb;
  const float dot = yb.x * x.x + yb.y * x.y;
  const float coef = A / ((m + MC_EPSIL) * (m + MC_EPSIL) * m);
  return make_float2(q * yb.x - coef * dot * x.x, q * yb.y - coef * dot * x.y);
}

__device__ __forceinline__ float2 ld2(const float* p, long long i, long long HW) {
  return make_float2(p[i], p[HW + i]);
}
__device__ __forceinline__ void st2(float* p, long long i, long long HW, float2 v) {
  p[i] = v.x;
  p[HW + i] = v.y;
}

// Per node: its K^2 taps' cells and normalised weights (as the claim
// computes them) and the cells' entry counts.
__global__ void k_fa_taps(Consts c, InstGeo g, int N, const float* __restrict__ pos,
                          int* __restrict__ tapc, float* __restrict__ tapw, int* __restrict__ cnt) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  const int K = c.K, K2 = K * K, W = g.Wf, H = g.Hf;
  const float gx = pos[2 * n] * g.scale_x;
  const float gy = pos[2 * n + 1] * g.scale_y;
  const float sxf = floorf(gx) - (float)c.half;
  const float syf = floorf(gy) - (float)c.half;
  const int sx = imod(f2i(sxf), W);
  const int sy = imod(f2i(syf), H);
  float prod[IB_KMAX * IB_KMAX];
  float total = 0.0f;
  for (int t = 0; t < K2; ++t) {
    const int dx = t / K, dy = t - dx * K;
    const float wxv = peskin((sxf + (float)dx) - gx, c.half_width);
    const float wyv = peskin((syf + (float)dy) - gy, c.half_width);
    prod[t] = wxv * wyv;
    total = total + prod[t];
  }
  const float den = total + MC_EPSIL;
  for (int t = 0; t < K2; ++t) {
    const int dx = t / K, dy = t - dx * K;
    const int cell = wrap1(sy + dy, H) * W + wrap1(sx + dx, W);
    tapc[(size_t)n * K2 + t] = cell;
    tapw[(size_t)n * K2 + t] = prod[t] / den;
    atomicAdd(&cnt[cell], 1);
  }
}

// Exclusive prefix of the counts (one workgroup, a contiguous chunk per
// thread); off[HW] = the total.
__global__ void __launch_bounds__(1024) k_fa_scan(const int* __restrict__ cnt,
                                                  int* __restrict__ off, long long HW) {
  __shared__ int part[1024];
  const int tid = threadIdx.x;
  const long long chunk = (HW + 1023) / 1024;
  const long long lo = tid * chunk, hi = lo + chunk < HW ? lo + chunk : HW;
  int s = 0;
  for (long long i = lo; i < hi; ++i) s += cnt[i];
  part[tid] = s;
  __syncthreads();
  if (tid == 0) {
    int run = 0;
    for (int k = 0; k < 1024; ++k) {
      const int v = part[k];
      part[k] = run;
      run += v;
    }
    off[HW] = run;
  }
  __syncthreads();
  int run = part[tid];
  for (long long i = lo; i < hi; ++i) {
    off[i] = run;
    run += cnt[i];
  }
}

__global__ void k_fa_fill(int NE, const int* __restrict__ tapc, const int* __restrict__ off,
                          int* __restrict__ cur, int* __restrict__ list) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= NE) return;
  const int cell = tapc[e];
  list[off[cell] + atomicAdd(&cur[cell], 1)] = e;
}

// Per cell: its entries in ascending order, the kernel weight and 1/max(kw,1)
// (fluid.py:123-133,289-290), and the macroscopic state (fluid.py:268-276).
__global__ void k_fa_cells(Consts c, InstGeo g, long long HW, const int* __restrict__ off,
                           int* __restrict__ list, const float* __restrict__ tapw,
                           const float* __restrict__ fk, float* __restrict__ ikw,
                           float* __restrict__ rraw, float* __restrict__ rho,
                           float* __restrict__ ur, float* __restrict__ u, float* __restrict__ uc) {
  const long long x = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= HW) return;
  const int lo = off[x], hi = off[x + 1];
  for (int j = lo + 1; j < hi; ++j) {
    const int v = list[j];
    int k = j - 1;
    while (k >= lo && list[k] > v) {
      list[k + 1] = list[k];
      --k;
    }
    list[k + 1] = v;
  }
  float kw = 0.0f;
  for (int j = lo; j < hi; ++j) kw = kw + tapw[list[j]];
  ikw[x] = 1.0f / fmax_nan(kw, 1.0f);
  float f[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) f[i] = fmax_nan(fk[i * HW + x], MC_EPSIL);
  float r = f[0];
#pragma unroll
  for (int i = 1; i < 9; ++i) r = r + f[i];
  const float rh = fmax_nan(r, 0.1f);
  const float mx = ((((f[1] - f[3]) + f[5]) - f[6]) - f[7]) + f[8];
  const float my = ((((f[2] - f[4]) + f[5]) + f[6]) - f[7]) - f[8];
  const float2 v = make_float2(mx / rh, my / rh);
  const float2 vc = clamp2(v, g.umax);
  rraw[x] = r;
  rho[x] = rh;
  st2(ur, x, HW, v);
  st2(u, x, HW, vc);
  st2(uc, x, HW, vc);
}

// Iteration j, nodes: interpolation, deficit, per-node clamp (fluid.py:292-303).
__global__ void k_fa_interp(Consts c, InstGeo g, int N, long long HW, const int* __restrict__ tapc,
                            const float* __restrict__ tapw, const float* __restrict__ uc,
                            const float* __restrict__ vel, float* __restrict__ dj,
                            float* __restrict__ cc) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  const int K2 = c.K * c.K;
  float ux = 0.0f, uy = 0.0f;
  for (int t = 0; t < K2; ++t) {
    const int cell = tapc[(size_t)n * K2 + t];
    const float w = tapw[(size_t)n * K2 + t];
    ux = ux + uc[cell] * w;
    uy = uy + uc[HW + cell] * w;
  }
  const float vbx = (vel[2 * n] * g.scale_x) * c.dt;
  const float vby = (vel[2 * n + 1] * g.scale_y) * c.dt;
  const float2 d = make_float2((vbx - ux) * c.relax, (vby - uy) * c.relax);
  dj[2 * n] = d.x;
  dj[2 * n + 1] = d.y;
  const float2 q = clamp2(d, g.umax);
  cc[2 * n] = q.x;
  cc[2 * n + 1] = q.y;
}

// Iteration j, cells: the spread of the clamped corrections times 1/kw
// (fluid.py:305-312).
__global__ void k_fa_spread(Consts c, long long HW, const int* __restrict__ off,
                            const int* __restrict__ list, const float* __restrict__ tapw,
                            const float* __restrict__ ikw, const float* __restrict__ cc,
                            float* __restrict__ uc) {
  const long long x = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= HW) return;
  const int lo = off[x], hi = off[x + 1];
  if (lo == hi) return;
  const int K2 = c.K * c.K;
  float ax = 0.0f, ay = 0.0f;
  for (int j = lo; j < hi; ++j) {
    const int e = list[j];
    const int n = e / K2;
    const float w = tapw[e];
    ax = ax + cc[2 * n] * w;
    ay = ay + cc[2 * n + 1] * w;
  }
  const float iv = ikw[x];
  uc[x] = uc[x] + ax * iv;
  uc[HW + x] = uc[HW + x] + ay * iv;
}

// The force on every cell (fluid.py:327-339): F = [clamp]((clamp(uc) - u) rho).
__global__ void k_fa_force(Consts c, InstGeo g, long long HW, const float* __restrict__ uc,
                           const float* __restrict__ u, const float* __restrict__ rho,
                           float* __restrict__ F) {
  const long long x = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= HW) return;
  const float2 ucl = clamp2(ld2(uc, x, HW), g.umax);
  const float2 u0 = ld2(u, x, HW);
  const float r = rho[x];
  float2 G = make_float2((ucl.x - u0.x) * r, (ucl.y - u0.y) * r);
  if (c.force_clamp) G = clamp2(G, g.umax);
  st2(F, x, HW, G);
}

// Reverse of collide-and-stream per cell x: the cotangents of f0, rho, u and
// F from those of the streamed populations f'[x + e_i].
__global__ void k_fa_collide_bar(Consts c, InstGeo g, long long HW, const float* __restrict__ fk,
                                 const float* __restrict__ rho, const float* __restrict__ u,
                                 const float* __restrict__ F, const float* __restrict__ fbn,
                                 float* __restrict__ f0b, float* __restrict__ rhob,
                                 float* __restrict__ ub, float* __restrict__ Fb) {
  const long long x = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= HW) return;
  const int W = g.Wf, H = g.Hf;
  const int y = (int)(x / W), xx = (int)(x - (long long)y * W);
  float f[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) f[i] = fmax_nan(fk[i * HW + x], MC_EPSIL);
  const float r = rho[x];
  const float2 u0 = ld2(u, x, HW), Fx = ld2(F, x, HW);
  const float usx = u0.x + (0.5f * Fx.x) / r, usy = u0.y + (0.5f * Fx.y) / r;
  const float usq = usx * usx + usy * usy;
  const float udf = usx * Fx.x + usy * Fx.y;
  float feq[9], raw[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    const float eu = (float)ex_of(i) * usx + (float)ey_of(i) * usy;
    const float ef = (float)ex_of(i) * Fx.x + (float)ey_of(i) * Fx.y;
    feq[i] = (r * lwt(i)) * (((1.0f + 3.0f * eu) + 4.5f * (eu * eu)) - 1.5f * usq);
    raw[i] = lwt(i) * (3.0f * (ef - udf) + (9.0f * eu) * ef);
  }
  // cotangent of each post-collision value: that of its streamed destination
  // through the floor at eps
  float gp[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    const int o = OPP[i];
    const float coll = (f[i] - c.omega_p * (0.5f * (f[i] + f[o]) - 0.5f * (feq[i] + feq[o]))) -
                       c.omega_m * (0.5f * (f[i] - f[o]) - 0.5f * (feq[i] - feq[o]));
    const float src = c.cp * (0.5f * (raw[i] + raw[o])) + c.cm * (0.5f * (raw[i] - raw[o]));
    const int yd = wrap1(y + ey_of(i), H), xd = wrap1(xx + ex_of(i), W);
    gp[i] = fbn[i * HW + (long long)yd * W + xd] * dmax(coll + src, MC_EPSIL);
  }
  const float op = c.omega_p, om = c.omega_m, cp = c.cp, cm = c.cm;
  float feqb[9], rawb[9];
#pragma unroll
  for (int j = 0; j < 9; ++j) {
    const int o = OPP[j];
    f0b[j * HW + x] = gp[j] * (1.0f - 0.5f * op - 0.5f * om) + gp[o] * (0.5f * om - 0.5f * op);
    feqb[j] = gp[j] * (0.5f * op + 0.5f * om) + gp[o] * (0.5f * op - 0.5f * om);
    rawb[j] = gp[j] * (0.5f * cp + 0.5f * cm) + gp[o] * (0.5f * cp - 0.5f * cm);
  }
  float rb = 0.0f, usbx = 0.0f, usby = 0.0f, fbx = 0.0f, fby = 0.0f;
#pragma unroll
  for (int j = 0; j < 9; ++j) {
    const float exf = (float)ex_of(j), eyf = (float)ey_of(j);
    const float eu = exf * usx + eyf * usy;
    const float ef = exf * Fx.x + eyf * Fx.y;
    const float wj = lwt(j);
    // feq_j = rho w (1 + 3 eu + 4.5 eu^2 - 1.5 |u*|^2)
    rb += feqb[j] * wj * (1.0f + 3.0f * eu + 4.5f * eu * eu - 1.5f * usq);
    usbx += feqb[j] * r * wj * (3.0f * exf + 9.0f * eu * exf - 3.0f * usx);
    usby += feqb[j] * r * wj * (3.0f * eyf + 9.0f * eu * eyf - 3.0f * usy);
    // raw_j = w (3 (e.F - u*.F) + 9 (e.u*) (e.F))
    fbx += rawb[j] * wj * (3.0f * exf - 3.0f * usx + 9.0f * eu * exf);
    fby += rawb[j] * wj * (3.0f * eyf - 3.0f * usy + 9.0f * eu * eyf);
    usbx += rawb[j] * wj * (-3.0f * Fx.x + 9.0f * ef * exf);
    usby += rawb[j] * wj * (-3.0f * Fx.y + 9.0f * ef * eyf);
  }
  // u* = u + 0.5 F / rho
  fbx += 0.5f * usbx / r;
  fby += 0.5f * usby / r;
  rb -= 0.5f * (usbx * Fx.x + usby * Fx.y) / (r * r);
  rhob[x] = rb;
  st2(ub, x, HW, make_float2(usbx, usby));
  st2(Fb, x, HW, make_float2(fbx, fby));
}

// Reverse of the drag (fluid.py:364-382): d fn = -d v'' / m / (scale dt).
__global__ void k_fa_drag_bar(Consts c, InstGeo g, int N, const float* __restrict__ vb,
                              float* __restrict__ fnb, float* __restrict__ vibb) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  fnb[2 * n] = -(vb[2 * n] / c.node_mass) / (g.scale_x * c.dt);
  fnb[2 * n + 1] = -(vb[2 * n + 1] / c.node_mass) / (g.scale_y * c.dt);
  vibb[2 * n] = 0.0f;
  vibb[2 * n + 1] = 0.0f;
}

// Reverse of the force per cell: the drag's share of F's cotangent (a sum
// over the cell's entries), the force clamp, F = (clamp(uc) - u) rho and the
// post clamp.
__global__ void k_fa_force_bar(Consts c, InstGeo g, long long HW, const int* __restrict__ off,
                               const int* __restrict__ list, const float* __restrict__ tapw,
                               const float* __restrict__ fnb, const float* __restrict__ uc,
                               const float* __restrict__ u, const float* __restrict__ rho,
                               float* __restrict__ Fb, float* __restrict__ ub,
                               float* __restrict__ rhob, float* __restrict__ ucb) {
  const long long x = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= HW) return;
  const int K2 = c.K * c.K;
  float2 fb = ld2(Fb, x, HW);
  for (int j = off[x]; j < off[x + 1]; ++j) {
    const int e = list[j];
    const int n = e / K2;
    fb.x += tapw[e] * fnb[2 * n];
    fb.y += tapw[e] * fnb[2 * n + 1];
  }
  const float2 ucv = ld2(uc, x, HW), u0 = ld2(u, x, HW);
  const float r = rho[x];
  const float2 ucl = clamp2(ucv, g.umax);
  const float2 G = make_float2((ucl.x - u0.x) * r, (ucl.y - u0.y) * r);
  const float2 gb = c.force_clamp ? clamp2_bar(G, fb, g.umax) : fb;
  const float2 uclb = make_float2(gb.x * r, gb.y * r);
  float2 ubv = ld2(ub, x, HW);
  ubv.x -= gb.x * r;
  ubv.y -= gb.y * r;
  st2(ub, x, HW, ubv);
  rhob[x] += gb.x * (ucl.x - u0.x) + gb.y * (ucl.y - u0.y);
  st2(ucb, x, HW, clamp2_bar(ucv, uclb, g.umax));
}

// Reverse of iteration j, nodes: the spread's cotangent gathered from the
// node's taps, the per-node clamp, the deficit.
__global__ void k_fa_iter_node(Consts c, InstGeo g, int N, long long HW,
                               const int* __restrict__ tapc, const float* __restrict__ tapw,
                               const float* __restrict__ ikw, const float* __restrict__ ucb,
                               const float* __restrict__ dj, float* __restrict__ vibb,
                               float* __restrict__ Ub) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  const int K2 = c.K * c.K;
  float2 ccb = make_float2(0.0f, 0.0f);
  for (int t = 0; t < K2; ++t) {
    const int cell = tapc[(size_t)n * K2 + t];
    const float w = tapw[(size_t)n * K2 + t] * ikw[cell];
    ccb.x += w * ucb[cell];
    ccb.y += w * ucb[HW + cell];
  }
  const float2 db = clamp2_bar(make_float2(dj[2 * n], dj[2 * n + 1]), ccb, g.umax);
  vibb[2 * n] += c.relax * db.x;
  vibb[2 * n + 1] += c.relax * db.y;
  Ub[2 * n] = -c.relax * db.x;
  Ub[2 * n + 1] = -c.relax * db.y;
}

// Reverse of iteration j, cells: the interpolation's cotangent.
__global__ void k_fa_iter_cell(Consts c, long long HW, const int* __restrict__ off,
                               const int* __restrict__ list, const float* __restrict__ tapw,
                               const float* __restrict__ Ub, float* __restrict__ ucb) {
  const long long x = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= HW) return;
  const int lo = off[x], hi = off[x + 1];
  if (lo == hi) return;
  const int K2 = c.K * c.K;
  float ax = 0.0f, ay = 0.0f;
  for (int j = lo; j < hi; ++j) {
    const int e = list[j];
    const int n = e / K2;
    ax += tapw[e] * Ub[2 * n];
    ay += tapw[e] * Ub[2 * n + 1];
  }
  ucb[x] += ax;
  ucb[HW + x] += ay;
}

// The populations' cotangent: u's (uc_0 = u, the collision and the force)
// through the first clamp and ur = m / rho, rho's through max(., 0.1), and
// the floor f0 = max(f_k, eps).
__global__ void k_fa_final_cell(Consts c, InstGeo g, long long HW, const float* __restrict__ fk,
                                const float* __restrict__ rraw, const float* __restrict__ rho,
                                const float* __restrict__ ur, const float* __restrict__ ub,
                                const float* __restrict__ ucb, const float* __restrict__ rhob,
                                const float* __restrict__ f0b, float* __restrict__ fb_out) {
  const long long x = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= HW) return;
  const float2 ubt = make_float2(ub[x] + ucb[x], ub[HW + x] + ucb[HW + x]);
  const float2 urv = ld2(ur, x, HW);
  const float2 urb = clamp2_bar(urv, ubt, g.umax);
  const float r = rho[x];
  const float mbx = urb.x / r, mby = urb.y / r;
  const float rb = rhob[x] - (urb.x * urv.x + urb.y * urv.y) / r;
  const float rrb = rb * dmax(rraw[x], 0.1f);
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    const float fbi = f0b[i * HW + x] + rrb + (float)ex_of(i) * mbx + (float)ey_of(i) * mby;
    fb_out[i * HW + x] = fbi * dmax(fk[i * HW + x], MC_EPSIL);
  }
}

// v' cotangent: that of v'' plus the targets' share, vib = v' scale dt.
__global__ void k_fa_final_node(Consts c, InstGeo g, int N, const float* __restrict__ vibb,
                                float* __restrict__ vb) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  vb[2 * n] += (vibb[2 * n] * g.scale_x) * c.dt;
  vb[2 * n + 1] += (vibb[2 * n + 1] * g.scale_y) * c.dt;
}

unsigned blocks(long long n) { return (unsigned)((n + FA_THREADS - 1) / FA_THREADS); }

}  // namespace

size_t fluid_adj_floats(int N, int K, int iters, long long HW) {
  const size_t K2 = (size_t)K * K;
  // per cell: ikw, rraw, rho, ur 2, u 2, uc 2, F 2, ub 2, ucb 2, rhob, Fb 2, f0b 9
  // per node: d (2 iters), cc 2, vibb 2, Ub 2, fnb 2; per entry: weight
  return (size_t)HW * 29 + (size_t)N * (2 * (size_t)iters + 8) + (size_t)N * K2;
}

size_t fluid_adj_ints(int N, int K, long long HW) {
  const size_t K2 = (size_t)K * K;
  // cells of the taps, entry list; counts, offsets (+1), cursors
  return 2 * (size_t)N * K2 + 3 * (size_t)HW + 1;
}

hipError_t fluid_adjoint_step(const Consts& c, const InstGeo& g, int N, float* fw, int* iw,
                              const float* fk, const float* x1, const float* v1, float* vb,
                              const float* fbn, float* fb_out, hipStream_t s) {
  const long long HW = (long long)g.Hf * g.Wf;
  const int K2 = c.K * c.K;
  const size_t NE = (size_t)N * K2;
  float* ikw = fw;
  float* rraw = ikw + HW;
  float* rho = rraw + HW;
  float* ur = rho + HW;
  float* u = ur + 2 * HW;
  float* uc = u + 2 * HW;
  float* F = uc + 2 * HW;
  float* ub = F + 2 * HW;
  float* ucb = ub + 2 * HW;
  float* rhob = ucb + 2 * HW;
  float* Fb = rhob + HW;
  float* f0b = Fb + 2 * HW;
  float* dj = f0b + 9 * HW;                 // iters x (N, 2)
  float* cc = dj + 2 * (size_t)N * c.iters;
  float* vibb = cc + 2 * (size_t)N;
  float* Ub = vibb + 2 * (size_t)N;
  float* fnb = Ub + 2 * (size_t)N;
  float* tapw = fnb + 2 * (size_t)N;
  int* tapc = iw;
  int* list = tapc + NE;
  int* cnt = list + NE;
  int* off = cnt + HW;
  int* cur = off + HW + 1;
  hipError_t e;
  if ((e = hipMemsetAsync(cnt, 0, (size_t)HW * sizeof(int), s)) != hipSuccess) return e;
  if ((e = hipMemsetAsync(cur, 0, (size_t)HW * sizeof(int), s)) != hipSuccess) return e;
  if (N > 0) hipLaunchKernelGGL(k_fa_taps, dim3(blocks(N)), dim3(FA_THREADS), 0, s, c, g, N, x1, tapc, tapw, cnt);
  hipLaunchKernelGGL(k_fa_scan, dim3(1), dim3(1024), 0, s, cnt, off, HW);
  if (N > 0)
    hipLaunchKernelGGL(k_fa_fill, dim3(blocks((long long)NE)), dim3(FA_THREADS), 0, s, (int)NE, tapc, off, cur, list);
  hipLaunchKernelGGL(k_fa_cells, dim3(blocks(HW)), dim3(FA_THREADS), 0, s, c, g, HW, off, list, tapw, fk, ikw,
                     rraw, rho, ur, u, uc);
  for (int j = 0; j < c.iters; ++j) {
    if (N > 0)
      hipLaunchKernelGGL(k_fa_interp, dim3(blocks(N)), dim3(FA_THREADS), 0, s, c, g, N, HW, tapc, tapw, uc, v1,
                         dj + 2 * (size_t)N * j, cc);
    hipLaunchKernelGGL(k_fa_spread, dim3(blocks(HW)), dim3(FA_THREADS), 0, s, c, HW, off, list, tapw, ikw, cc, uc);
  }
  hipLaunchKernelGGL(k_fa_force, dim3(blocks(HW)), dim3(FA_THREADS), 0, s, c, g, HW, uc, u, rho, F);
  // reverse
  hipLaunchKernelGGL(k_fa_collide_bar, dim3(blocks(HW)), dim3(FA_THREADS), 0, s, c, g, HW, fk, rho, u, F, fbn,
                     f0b, rhob, ub, Fb);
  if (N > 0) hipLaunchKernelGGL(k_fa_drag_bar, dim3(blocks(N)), dim3(FA_THREADS), 0, s, c, g, N, vb, fnb, vibb);
  hipLaunchKernelGGL(k_fa_force_bar, dim3(blocks(HW)), dim3(FA_THREADS), 0, s, c, g, HW, off, list, tapw, fnb, uc,
                     u, rho, Fb, ub, rhob, ucb);
  for (int j = c.iters - 1; j >= 0; --j) {
    if (N > 0)
      hipLaunchKernelGGL(k_fa_iter_node, dim3(blocks(N)), dim3(FA_THREADS), 0, s, c, g, N, HW, tapc, tapw, ikw,
                         ucb, dj + 2 * (size_t)N * j, vibb, Ub);
    hipLaunchKernelGGL(k_fa_iter_cell, dim3(blocks(HW)), dim3(FA_THREADS), 0, s, c, HW, off, list, tapw, Ub, ucb);
  }
  hipLaunchKernelGGL(k_fa_final_cell, dim3(blocks(HW)), dim3(FA_THREADS), 0, s, c, g, HW, fk, rraw, rho, ur, ub,
                     ucb, rhob, f0b, fb_out);
  if (N > 0) hipLaunchKernelGGL(k_fa_final_node, dim3(blocks(N)), dim3(FA_THREADS), 0, s, c, g, N, vibb, vb);
  return hipGetLastError();
}

}  // namespace mc
