// Adjoint (reverse-mode derivative) of the rod-only coupled step: simulate.py
// :37-100 with the fluid and the steric term off and the PBD constraint
// (pbd.py:14-91).  The reference differentiates this step with jax.grad
// through lax.scan (experiments/filament_folding/experiment.py:985-1018).
//
// One workgroup per filament (connected component), like the forward rod
// kernel.  The workgroup first recomputes the step's constraint cycles from
// the step's input state, with the forward's arithmetic, keeping every
// cycle's positions (before the position pass) and angles (after the angle
// pass) in LDS; then it runs the cycles backwards:
//
//   position pass, per edge (s, g):  Fb = pb[s] / deg_s - pb[g] / deg_g
//       rb = J Fb,  J = k_st/2 ((1 - k_sh)((1 - l0/L) I + l0/L rhat rhat^T) + k_sh I)
//       pb[g] += rb, pb[s] -= rb   (per node: an ordered gather of its edges)
//       thb'_e += -k_st/2 k_sh l0 (-sin th' Fb_x + cos th' Fb_y)
//   angle pass, per hinge: corrb = 1.8 (thb[e_in] / deg_in - thb[e_out] / deg_out)
//       rest_bar -= corrb kb / 2;  thb[e_out] += corrb kb / 2, thb[e_in] -= corrb kb / 2
//
// and finally the prologue p0 = x + (v d) dt and the epilogue v = (p_C - x)/dt.
// The minimum image, the angle wrap and the final periodic wrap have
// derivative 1 (their jumps are not differentiated, as in jax.grad).
#include "mc_common.h"
#include "mc_internal.h"

namespace mc {
namespace {

__global__ void __launch_bounds__(256) k_rod_adjoint(RodAdjArgs a) {
  extern __shared__ float sm[];
  const Consts& c = a.c;
  const RodLayoutDev& L = a.L;
  const int comp = blockIdx.x;
  const CompMeta m = L.meta[comp];
  const InstGeo& gi = a.bd.geo[m.inst];
  const int n0 = m.n0, nn = m.nn, e0 = m.e0, ne = m.ne, b0 = m.b0, nb = m.nb;
  const int C = c.cycles;
  const int tid = threadIdx.x, bs = blockDim.x;
  float* Xs = sm;                    // C x 2nn: positions before each position pass
  float* THs = Xs + (size_t)C * 2 * nn;  // C x ne: angles after each angle pass
  float* X = THs + (size_t)C * ne;   // 2nn
  float* TH = X + 2 * nn;            // ne
  float* CO = TH + ne;               // nb
  float* FE = CO + nb;               // 2ne: edge forces, then edge adjoints rb
  float* XB = FE + 2 * ne;           // 2nn
  float* THB = XB + 2 * nn;          // ne
  float* DB = THB + ne;              // nb
  float* RB = DB + nb;               // nb
  const int* esrc = L.edge_src + e0;
  const int* etgt = L.edge_tgt + e0;
  const int* hin = L.hinge_in + b0;
  const int* hout = L.hinge_out + b0;
  const int* einc_off = L.edge_inc_off + e0;
  const int* ninc_off = L.node_inc_off + n0;
  const float Wd = gi.dom_wf, Hd = gi.dom_hf;

  // prologue (simulate.py:46-50) and the step's angles
  for (int i = tid; i < nn; i += bs) {
    const int g = L.comp_nodes[n0 + i];
    const float dvx = a.vel[2 * g] * c.damping, dvy = a.vel[2 * g + 1] * c.damping;
    X[2 * i] = a.pos[2 * g] + dvx * c.dt;
    X[2 * i + 1] = a.pos[2 * g + 1] + dvy * c.dt;
  }
  for (int e = tid; e < ne; e += bs) TH[e] = a.theta[L.comp_edges[e0 + e]];
  __syncthreads();
  // forward cycles, as the rod kernel computes them (PBD branch)
  for (int cyc = 0; cyc < C; ++cyc) {
    float* Xc = Xs + (size_t)cyc * 2 * nn;
    for (int b = tid; b < nb; b += bs) {
      const int gb = L.comp_hinges[b0 + b];
      const float diff = wrap_angle(TH[hout[b]] - TH[hin[b]]);
      CO[b] = ((diff - a.rest_ang[gb]) * a.kb[gb]) * 0.5f;
    }
    for (int i = tid; i < 2 * nn; i += bs) Xc[i] = X[i];
    __syncthreads();
    for (int e = tid; e < ne; e += bs) {
      const int i0 = einc_off[e], i1 = einc_off[e + 1];
      float d = 0.0f;
      for (int q = i0; q < i1; ++q) {
        const int code = L.edge_inc[q];
        const float v = CO[code >> 1];
        d = d + ((code & 1) ? -v : v);
      }
      const float deg = fmax_nan((float)(i1 - i0), 1.0f);
      const float th = TH[e] + (d / deg) * 1.8f;
      TH[e] = th;
      THs[(size_t)cyc * ne + e] = th;
      const int sn = esrc[e], tn = etgt[e];
      const float rx = min_image(X[2 * tn], X[2 * sn], Wd);
      const float ry = min_image(X[2 * tn + 1], X[2 * sn + 1], Hd);
      const float l = a.rest_len[L.comp_edges[e0 + e]];
      double sd, cd;
      sincos((double)th, &sd, &cd);
      const float tx = (float)cd * l, ty = (float)sd * l;
      const float len = sqrtf(rx * rx + ry * ry);
      const float md = fmax_nan(len, 1e-8f);
      const float cdx = rx / md, cdy = ry / md;
      const float stx = cdx * (len - l), sty = cdy * (len - l);
      const float shx = (rx - tx) - stx, shy = (ry - ty) - sty;
      FE[2 * e] = ((stx + shx * c.k_sh) * c.k_st) * 0.5f;
      FE[2 * e + 1] = ((sty + shy * c.k_sh) * c.k_st) * 0.5f;
    }
    __syncthreads();
    for (int i = tid; i < nn; i += bs) {
      const int i0 = ninc_off[i], i1 = ninc_off[i + 1];
      float ux = 0.0f, uy = 0.0f;
      for (int q = i0; q < i1; ++q) {
        const int code = L.node_inc[q];
        const int e = code >> 1;
        ux = ux + ((code & 1) ? -FE[2 * e] : FE[2 * e]);
        uy = uy + ((code & 1) ? -FE[2 * e + 1] : FE[2 * e + 1]);
      }
      const float deg = fmax_nan((float)(i1 - i0), 1.0f);
      X[2 * i] = X[2 * i] + ux / deg;
      X[2 * i + 1] = X[2 * i + 1] + uy / deg;
    }
    __syncthreads();
  }
  // adjoint of the epilogue: pb = xb + vb / dt (the direct -vb / dt term is
  // added at the end)
  for (int i = tid; i < nn; i += bs) {
    const int g = L.comp_nodes[n0 + i];
    XB[2 * i] = a.xb[2 * g] + a.vb[2 * g] / c.dt;
    XB[2 * i + 1] = a.xb[2 * g + 1] + a.vb[2 * g + 1] / c.dt;
  }
  for (int e = tid; e < ne; e += bs) THB[e] = a.thb[L.comp_edges[e0 + e]];
  for (int b = tid; b < nb; b += bs) RB[b] = 0.0f;
  __syncthreads();
  const float kst = c.k_st, ksh = c.k_sh;
  for (int cyc = C - 1; cyc >= 0; --cyc) {
    const float* Xc = Xs + (size_t)cyc * 2 * nn;
    const float* THc = THs + (size_t)cyc * ne;
    // position pass: per edge
    for (int e = tid; e < ne; e += bs) {
      const int sn = esrc[e], tn = etgt[e];
      const float ds = fmax_nan((float)(ninc_off[sn + 1] - ninc_off[sn]), 1.0f);
      const float dg = fmax_nan((float)(ninc_off[tn + 1] - ninc_off[tn]), 1.0f);
      const float fbx = XB[2 * sn] / ds - XB[2 * tn] / dg;
      const float fby = XB[2 * sn + 1] / ds - XB[2 * tn + 1] / dg;
      const float rx = min_image(Xc[2 * tn], Xc[2 * sn], Wd);
      const float ry = min_image(Xc[2 * tn + 1], Xc[2 * sn + 1], Hd);
      const float l = a.rest_len[L.comp_edges[e0 + e]];
      const float len = sqrtf(rx * rx + ry * ry);
      const float ls = fmax_nan(len, 1e-8f);
      const float hx = rx / ls, hy = ry / ls;
      const float al = 1.0f - l / ls, bl = l / ls;
      const float pr = hx * fbx + hy * fby;
      const float sx = al * fbx + bl * pr * hx, sy = al * fby + bl * pr * hy;
      FE[2 * e] = 0.5f * kst * ((1.0f - ksh) * sx + ksh * fbx);
      FE[2 * e + 1] = 0.5f * kst * ((1.0f - ksh) * sy + ksh * fby);
      double sd, cd;
      sincos((double)THc[e], &sd, &cd);
      THB[e] = THB[e] + (-0.5f * kst * ksh * l) * (-(float)sd * fbx + (float)cd * fby);
    }
    __syncthreads();
    // per node: gather of its edges' rb (+ at the edge's target, - at its source)
    for (int i = tid; i < nn; i += bs) {
      const int i0 = ninc_off[i], i1 = ninc_off[i + 1];
      float ux = 0.0f, uy = 0.0f;
      for (int q = i0; q < i1; ++q) {
        const int code = L.node_inc[q];
        const int e = code >> 1;
        ux = ux + ((code & 1) ? FE[2 * e] : -FE[2 * e]);
        uy = uy + ((code & 1) ? FE[2 * e + 1] : -FE[2 * e + 1]);
      }
      XB[2 * i] = XB[2 * i] + ux;
      XB[2 * i + 1] = XB[2 * i + 1] + uy;
    }
    // angle pass: per hinge
    for (int b = tid; b < nb; b += bs) {
      const int ei = hin[b], eo = hout[b];
      const float di = fmax_nan((float)(einc_off[ei + 1] - einc_off[ei]), 1.0f);
      const float dout = fmax_nan((float)(einc_off[eo + 1] - einc_off[eo]), 1.0f);
      const float corrb = 1.8f * (THB[ei] / di - THB[eo] / dout);
      const float db = corrb * a.kb[L.comp_hinges[b0 + b]] * 0.5f;
      DB[b] = db;
      RB[b] = RB[b] - db;
    }
    __syncthreads();
    for (int e = tid; e < ne; e += bs) {
      const int i0 = einc_off[e], i1 = einc_off[e + 1];
      float d = 0.0f;
      for (int q = i0; q < i1; ++q) {
        const int code = L.edge_inc[q];
        d = d + ((code & 1) ? DB[code >> 1] : -DB[code >> 1]);
      }
      THB[e] = THB[e] + d;
    }
    __syncthreads();
  }
  // adjoint of the prologue
  for (int i = tid; i < nn; i += bs) {
    const int g = L.comp_nodes[n0 + i];
    const float vbx = a.vb[2 * g], vby = a.vb[2 * g + 1];
    a.xb[2 * g] = XB[2 * i] - vbx / c.dt;
    a.xb[2 * g + 1] = XB[2 * i + 1] - vby / c.dt;
    a.vb[2 * g] = (XB[2 * i] * c.damping) * c.dt;
    a.vb[2 * g + 1] = (XB[2 * i + 1] * c.damping) * c.dt;
  }
  for (int e = tid; e < ne; e += bs) a.thb[L.comp_edges[e0 + e]] = THB[e];
  for (int b = tid; b < nb; b += bs) {
    const int gb = L.comp_hinges[b0 + b];
    a.rest_bar[gb] = a.rest_bar[gb] + RB[b];
  }
}

}  // namespace

size_t rod_adjoint_lds_bytes(int cycles, int nn, int ne, int nb) {
  return sizeof(float) * ((size_t)cycles * (2 * (size_t)nn + ne) + 4 * (size_t)nn + 4 * (size_t)ne +
                          3 * (size_t)nb);
}

hipError_t launch_rod_adjoint(const RodAdjArgs& a, size_t lds_bytes, hipStream_t s) {
  if (a.L.num_comp == 0) return hipSuccess;
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_rod_adjoint),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_rod_adjoint, dim3(a.L.num_comp), dim3(256), lds_bytes, s, a);
  return hipGetLastError();
}

}  // namespace mc
