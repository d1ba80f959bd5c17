// Adjoint (reverse-mode derivative) of the coupled step without the fluid:
// simulate.py:37-100 with the PBD (pbd.py:14-91) or stable Cosserat
// (stable_cosserat.py:13-119) constraint, with or without the steric term.
// The reference differentiates this step with jax.grad through lax.scan
// (experiments/filament_folding/experiment.py:985-1018).
//
// With the steric term every filament feels every other through the field,
// so each step is its own segment: the forward runs the coupled step's
// steric pipeline (prediction, deposit, filter, per-node force) and then a
// one-step segment that adds the force to the prediction; the reverse
// recomputes that step's field and force from the kept state, runs the
// one-step segment adjoint up to d loss / d (pushed prediction), and
// launch_steric_bar, the Gaussian filter and launch_steric_adj_nodes finish
// the step (see the steric section below).
//
// Without the fluid and the steric term the filaments do not interact, so one
// workgroup (one wavefront: its barriers are cheap) per filament (connected
// component) runs a whole checkpoint segment in one launch, its state in LDS:
//   * forward segment: K steps of the rod-only step with the forward rod
//     kernel's arithmetic (bit-identical results), the segment's start state
//     written to the checkpoint slot;
//   * reverse segment: the K steps recomputed once from the checkpoint, each
//     cycle's per-edge minimum-image vector and director (cos, sin) written
//     to the filament's part of a global tape, then the steps' cycles run
//     backwards from the tape (no trigonometry or recomputation there):
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

// Per-component view of the rod layout and the LDS arrays of a segment.
struct SegCtx {
  int nn, ne, nb, n0, e0, b0, C;
  const int *esrc, *etgt, *hin, *hout, *einc_off, *ninc_off, *edge_inc, *node_inc;
  const int *nodes, *edges, *hinges;   // component order -> global ids
  float Wd, Hd;
  // LDS
  float *X, *V, *TH, *P, *CO, *FE, *RS, *KB, *LE;
  float *Y, *HN;                       // Cosserat: inertial target, per-node diagonal h
  float *XB, *VB, *THB, *DB, *RB, *YB; // adjoints (reverse only)
  float* LB;                           // rest-length adjoint (reverse only)
};

__device__ __forceinline__ void seg_init(const RodAdjArgs& a, SegCtx& s, float* sm, bool reverse) {
  const RodLayoutDev& L = a.L;
  const CompMeta m = L.meta[blockIdx.x];
  const InstGeo& gi = a.bd.geo[m.inst];
  s.nn = m.nn; s.ne = m.ne; s.nb = m.nb; s.n0 = m.n0; s.e0 = m.e0; s.b0 = m.b0;
  s.C = a.c.cycles;
  s.esrc = L.edge_src + m.e0; s.etgt = L.edge_tgt + m.e0;
  s.hin = L.hinge_in + m.b0; s.hout = L.hinge_out + m.b0;
  s.einc_off = L.edge_inc_off + m.e0; s.ninc_off = L.node_inc_off + m.n0;
  s.edge_inc = L.edge_inc; s.node_inc = L.node_inc;
  s.nodes = L.comp_nodes + m.n0; s.edges = L.comp_edges + m.e0; s.hinges = L.comp_hinges + m.b0;
  s.Wd = gi.dom_wf; s.Hd = gi.dom_hf;
  const int nn = s.nn, ne = s.ne, nb = s.nb;
  float* q = sm;
  auto take = [&](size_t n) { float* r = q; q += n; return r; };
  s.X = take(2 * nn); s.V = take(2 * nn); s.P = take(2 * nn); s.TH = take(ne);
  s.CO = take(nb); s.FE = take(2 * ne); s.RS = take(nb); s.KB = take(nb); s.LE = take(ne);
  s.Y = take(2 * nn); s.HN = take(nn);
  if (reverse) {
    s.XB = take(2 * nn); s.VB = take(2 * nn); s.THB = take(ne); s.DB = take(nb); s.RB = take(nb);
    s.YB = take(2 * nn); s.LB = take(ne);
  }
  for (int b = threadIdx.x; b < nb; b += blockDim.x) {
    s.RS[b] = a.rest_ang[s.hinges[b]];
    s.KB[b] = a.kb[s.hinges[b]];
  }
  for (int e = threadIdx.x; e < ne; e += blockDim.x) s.LE[e] = a.rest_len[s.edges[e]];
  __syncthreads();
  // Cosserat diagonal h_n = w + sum of k_st / l^2 over the node's edges, in
  // the forward's incidence order (stable_cosserat.py:197-207)
  if (a.c.solver == 1)
    for (int i = threadIdx.x; i < nn; i += blockDim.x) {
      float h = a.c.cos_w;
      for (int q = s.ninc_off[i]; q < s.ninc_off[i + 1]; ++q) {
        const float l = s.LE[s.node_inc[q] >> 1];
        h = h + a.c.k_st / (l * l);
      }
      s.HN[i] = h;
    }
}

// One rod-only step on the LDS state (X, V, TH) -> (X, V, TH), with the
// arithmetic of the forward rod kernel (PBD branch, steric and fluid off):
// prologue (simulate.py:46-50), the cycles (pbd.py:14-91), velocity and the
// final wrap (simulate.py:83-98).  With `rec`, every cycle's per-edge
// minimum-image vector and the direction (cos, sin) of the updated angle --
// all the adjoint of the position pass needs -- go to rec[cyc * ne + e].
// With `urec` (Cosserat), each cycle's node update (w (y - x) + sum F) / h
// goes to urec[cyc * nn + i]: the rest-length adjoint's share through h.
__device__ __forceinline__ void seg_step(const Consts& c, SegCtx& s, float4* rec,
                                         const float* fst, float2* urec = nullptr) {
  const int tid = threadIdx.x, bs = blockDim.x;
  const int nn = s.nn, ne = s.ne, nb = s.nb;
  const bool is_cos = c.solver == 1;
  for (int i = tid; i < nn; i += bs) {
    const float dvx = s.V[2 * i] * c.damping, dvy = s.V[2 * i + 1] * c.damping;
    float x = s.X[2 * i] + dvx * c.dt, y = s.X[2 * i + 1] + dvy * c.dt;
    if (fst) {  // the steric push (simulate.py:52-63)
      const int g = s.nodes[i];
      x = x + (fst[2 * g] * c.steric_strength) * c.dt;
      y = y + (fst[2 * g + 1] * c.steric_strength) * c.dt;
    }
    s.P[2 * i] = x;
    s.P[2 * i + 1] = y;
    s.Y[2 * i] = x;
    s.Y[2 * i + 1] = y;
  }
  __syncthreads();
  for (int cyc = 0; cyc < s.C; ++cyc) {
    for (int b = tid; b < nb; b += bs) {
      const float diff = wrap_angle(s.TH[s.hout[b]] - s.TH[s.hin[b]]);
      s.CO[b] = ((diff - s.RS[b]) * s.KB[b]) * 0.5f;
    }
    __syncthreads();
    for (int e = tid; e < ne; e += bs) {
      const int i0 = s.einc_off[e], i1 = s.einc_off[e + 1];
      float d = 0.0f;
      for (int q = i0; q < i1; ++q) {
        const int code = s.edge_inc[q];
        const float v = s.CO[code >> 1];
        d = d + ((code & 1) ? -v : v);
      }
      const float deg = fmax_nan((float)(i1 - i0), 1.0f);
      const int sn = s.esrc[e], tn = s.etgt[e];
      const float rx = min_image(s.P[2 * tn], s.P[2 * sn], s.Wd);
      const float ry = min_image(s.P[2 * tn + 1], s.P[2 * sn + 1], s.Hd);
      float th;
      if (!is_cos) {
        th = s.TH[e] + (d / deg) * 1.8f;
      } else {
        // stable_cosserat.py:55-80: angle pass without SOR, then the 0.05 pull
        th = s.TH[e] + d / deg;
        const float alpha = datan2(ry, rx);
        const float pull = wrap_angle(alpha - th);
        th = th + pull * 0.05f;
      }
      s.TH[e] = th;
      const float l = s.LE[e];
      float sf, cf;
      fsincos(th, &sf, &cf);
      const double sd = sf, cd = cf;
      if (rec) rec[(size_t)cyc * ne + e] = make_float4(rx, ry, cf, sf);
      if (!is_cos) {
        const float tx = (float)cd * l, ty = (float)sd * l;
        const float len = sqrtf(rx * rx + ry * ry);
        const float md = fmax_nan(len, 1e-8f);
        const float cdx = rx / md, cdy = ry / md;
        const float stx = cdx * (len - l), sty = cdy * (len - l);
        const float shx = (rx - tx) - stx, shy = (ry - ty) - sty;
        s.FE[2 * e] = ((stx + shx * c.k_sh) * c.k_st) * 0.5f;
        s.FE[2 * e + 1] = ((sty + shy * c.k_sh) * c.k_st) * 0.5f;
      } else {
        const float cx = rx / l - (float)cd, cy = ry / l - (float)sd;
        const float kl = c.k_st / l;
        s.FE[2 * e] = kl * cx;
        s.FE[2 * e + 1] = kl * cy;
      }
    }
    __syncthreads();
    for (int i = tid; i < nn; i += bs) {
      const int i0 = s.ninc_off[i], i1 = s.ninc_off[i + 1];
      if (!is_cos) {
        float ux = 0.0f, uy = 0.0f;
        for (int q = i0; q < i1; ++q) {
          const int code = s.node_inc[q];
          const int e = code >> 1;
          ux = ux + ((code & 1) ? -s.FE[2 * e] : s.FE[2 * e]);
          uy = uy + ((code & 1) ? -s.FE[2 * e + 1] : s.FE[2 * e + 1]);
        }
        const float deg = fmax_nan((float)(i1 - i0), 1.0f);
        s.P[2 * i] = s.P[2 * i] + ux / deg;
        s.P[2 * i + 1] = s.P[2 * i + 1] + uy / deg;
      } else {
        const float w = c.cos_w;
        const float xx = s.P[2 * i], xy = s.P[2 * i + 1];
        float fx = w * (s.Y[2 * i] - xx), fy = w * (s.Y[2 * i + 1] - xy);
        for (int q = i0; q < i1; ++q) {
          const int code = s.node_inc[q];
          const int e = code >> 1;
          fx = fx + ((code & 1) ? -s.FE[2 * e] : s.FE[2 * e]);
          fy = fy + ((code & 1) ? -s.FE[2 * e + 1] : s.FE[2 * e + 1]);
        }
        s.P[2 * i] = xx + fx / s.HN[i];
        s.P[2 * i + 1] = xy + fy / s.HN[i];
        if (urec) urec[(size_t)cyc * nn + i] = make_float2(fx / s.HN[i], fy / s.HN[i]);
      }
    }
    __syncthreads();
  }
  for (int i = tid; i < nn; i += bs) {
    const float xx = s.P[2 * i], xy = s.P[2 * i + 1];
    s.V[2 * i] = (xx - s.X[2 * i]) / c.dt;
    s.V[2 * i + 1] = (xy - s.X[2 * i + 1]) / c.dt;
    s.X[2 * i] = fmodf_floor(xx, s.Wd);
    s.X[2 * i + 1] = fmodf_floor(xy, s.Hd);
  }
  __syncthreads();
}

// Packed state slot: position (2N), velocity (2N), theta (E) by global ids.
__device__ __forceinline__ void seg_load(const SegCtx& s, const float* pos, const float* vel,
                                         const float* th) {
  for (int i = threadIdx.x; i < s.nn; i += blockDim.x) {
    const int g = s.nodes[i];
    s.X[2 * i] = pos[2 * g];
    s.X[2 * i + 1] = pos[2 * g + 1];
    s.V[2 * i] = vel[2 * g];
    s.V[2 * i + 1] = vel[2 * g + 1];
  }
  for (int e = threadIdx.x; e < s.ne; e += blockDim.x) s.TH[e] = th[s.edges[e]];
  __syncthreads();
}

__device__ __forceinline__ void seg_store(const SegCtx& s, float* pos, float* vel, float* th) {
  for (int i = threadIdx.x; i < s.nn; i += blockDim.x) {
    const int g = s.nodes[i];
    pos[2 * g] = s.X[2 * i];
    pos[2 * g + 1] = s.X[2 * i + 1];
    vel[2 * g] = s.V[2 * i];
    vel[2 * g + 1] = s.V[2 * i + 1];
  }
  for (int e = threadIdx.x; e < s.ne; e += blockDim.x) th[s.edges[e]] = s.TH[e];
}

// Forward segment: the state (a.pos / a.vel / a.theta, in place) is written
// to the checkpoint slot, then advanced `steps` steps; debug_vector is zero
// (PBD).
__global__ void __launch_bounds__(64) k_rod_fwd_segment(RodAdjArgs a, int steps, float* ck,
                                                         float* dbg) {
  extern __shared__ float sm[];
  SegCtx s;
  seg_init(a, s, sm, false);
  float* pos = const_cast<float*>(a.pos);
  float* vel = const_cast<float*>(a.vel);
  float* th = const_cast<float*>(a.theta);
  seg_load(s, pos, vel, th);
  const size_t N2 = 2 * (size_t)a.N;
  seg_store(s, ck, ck + N2, ck + 2 * N2);
  for (int k = 0; k < steps; ++k) seg_step(a.c, s, nullptr, a.fst);
  seg_store(s, pos, vel, th);
  const bool is_cos = a.c.solver == 1;
  for (int i = threadIdx.x; i < s.nn; i += blockDim.x) {
    const int g = s.nodes[i];
    dbg[2 * g] = (is_cos && steps > 0) ? s.Y[2 * i] : 0.0f;
    dbg[2 * g + 1] = (is_cos && steps > 0) ? s.Y[2 * i + 1] : 0.0f;
  }
}

// Reverse of one step's stable Cosserat cycles (stable_cosserat.py:13-119)
// from its tape rec[cyc * ne + e] = (r, cos th2, sin th2):
//   Fb = pb[s] / h_s - pb[g] / h_g
//   rb = k / l^2 Fb + 0.05 th2b (-r_y, r_x) / |r|^2   (the pull's atan2)
//   th2b = thb - k / l (-sin th2 Fb_x + cos th2 Fb_y),  th1b = 0.95 th2b
//   pb' = pb (1 - w / h) + sum (+rb at the target, -rb at the source);
//   yb += pb w / h;  corrb = th1b[e_in] / deg_in - th1b[e_out] / deg_out
// With `urec` (the cycles' node updates u = (x' - x)), the rest-length
// adjoint: lb += Fb . dF/dl + (hb_s + hb_g) dh/dl with dF/dl = -2 k r / l^3 +
// k / l^2 (cos th2, sin th2), dh/dl = -2 k / l^3, hb_n = -(pb_n . u_n) / h_n.
__device__ __forceinline__ void cosserat_cycles_adjoint(const Consts& c, SegCtx& s,
                                                        const float4* rec,
                                                        const float2* urec) {
  const int tid = threadIdx.x, bs = blockDim.x;
  const int nn = s.nn, ne = s.ne, nb = s.nb;
  const float k = c.k_st, w = c.cos_w;
  for (int cyc = s.C - 1; cyc >= 0; --cyc) {
    const float4* rc = rec + (size_t)cyc * ne;
    for (int e = tid; e < ne; e += bs) {
      const float4 q = rc[e];
      const int sn = s.esrc[e], tn = s.etgt[e];
      const float fbx = s.XB[2 * sn] / s.HN[sn] - s.XB[2 * tn] / s.HN[tn];
      const float fby = s.XB[2 * sn + 1] / s.HN[sn] - s.XB[2 * tn + 1] / s.HN[tn];
      const float l = s.LE[e];
      if (urec) {
        const float2 us = urec[(size_t)cyc * nn + sn], ut = urec[(size_t)cyc * nn + tn];
        const float hb = -(s.XB[2 * sn] * us.x + s.XB[2 * sn + 1] * us.y) / s.HN[sn] -
                         (s.XB[2 * tn] * ut.x + s.XB[2 * tn + 1] * ut.y) / s.HN[tn];
        const float m3 = -2.0f * k / (l * l * l);
        s.LB[e] = s.LB[e] + m3 * ((q.x * fbx + q.y * fby) + hb) +
                  (k / (l * l)) * (q.z * fbx + q.w * fby);
      }
      const float th2b = s.THB[e] + (-(k / l)) * (-q.w * fbx + q.z * fby);
      const float ab = 0.05f * th2b / (q.x * q.x + q.y * q.y);
      s.FE[2 * e] = (k / (l * l)) * fbx - ab * q.y;
      s.FE[2 * e + 1] = (k / (l * l)) * fby + ab * q.x;
      s.THB[e] = 0.95f * th2b;
    }
    __syncthreads();
    for (int i = tid; i < nn; i += bs) {
      const int i0 = s.ninc_off[i], i1 = s.ninc_off[i + 1];
      float ux = 0.0f, uy = 0.0f;
      for (int q = i0; q < i1; ++q) {
        const int code = s.node_inc[q];
        const int e = code >> 1;
        ux = ux + ((code & 1) ? s.FE[2 * e] : -s.FE[2 * e]);
        uy = uy + ((code & 1) ? s.FE[2 * e + 1] : -s.FE[2 * e + 1]);
      }
      const float r = w / s.HN[i];
      const float pbx = s.XB[2 * i], pby = s.XB[2 * i + 1];
      s.YB[2 * i] = s.YB[2 * i] + pbx * r;
      s.YB[2 * i + 1] = s.YB[2 * i + 1] + pby * r;
      s.XB[2 * i] = pbx * (1.0f - r) + ux;
      s.XB[2 * i + 1] = pby * (1.0f - r) + uy;
    }
    for (int b = tid; b < nb; b += bs) {
      const int ei = s.hin[b], eo = s.hout[b];
      const float di = fmax_nan((float)(s.einc_off[ei + 1] - s.einc_off[ei]), 1.0f);
      const float dout = fmax_nan((float)(s.einc_off[eo + 1] - s.einc_off[eo]), 1.0f);
      const float corrb = s.THB[ei] / di - s.THB[eo] / dout;
      const float db = corrb * s.KB[b] * 0.5f;
      s.DB[b] = db;
      s.RB[b] = s.RB[b] - db;
    }
    __syncthreads();
    for (int e = tid; e < ne; e += bs) {
      const int i0 = s.einc_off[e], i1 = s.einc_off[e + 1];
      float d = 0.0f;
      for (int q = i0; q < i1; ++q) {
        const int code = s.edge_inc[q];
        d = d + ((code & 1) ? s.DB[code >> 1] : -s.DB[code >> 1]);
      }
      s.THB[e] = s.THB[e] + d;
    }
    __syncthreads();
  }
}

// Reverse segment from checkpoint slot `ck` over `steps` steps: one forward
// pass records each cycle's per-edge vector and direction in the filament's
// part of the tape (steps x C x ne float4), then the step adjoints run
// backwards from the tape; adjoints (a.xb / a.vb / a.thb) and the rest-angle
// adjoint (a.rest_bar) are updated in place.
__global__ void __launch_bounds__(64) k_rod_adj_segment(RodAdjArgs a, int steps, const float* ck,
                                                         float* tape) {
  extern __shared__ float sm[];
  SegCtx s;
  seg_init(a, s, sm, true);
  const Consts& c = a.c;
  const int tid = threadIdx.x, bs = blockDim.x;
  const int nn = s.nn, ne = s.ne, nb = s.nb, C = s.C;
  const size_t N2 = 2 * (size_t)a.N;
  float4* tp = reinterpret_cast<float4*>(tape) + (size_t)steps * C * s.e0;
  const bool lbar = a.len_bar != nullptr;
  float2* up = (lbar && c.solver == 1)
                   ? reinterpret_cast<float2*>(a.utape) + (size_t)steps * C * s.n0
                   : nullptr;
  seg_load(s, ck, ck + N2, ck + 2 * N2);
  for (int k = 0; k < steps; ++k)
    seg_step(c, s, tp + (size_t)k * C * ne, a.fst, up ? up + (size_t)k * C * nn : nullptr);
  for (int i = tid; i < nn; i += bs) {
    const int g = s.nodes[i];
    s.XB[2 * i] = a.xb[2 * g];
    s.XB[2 * i + 1] = a.xb[2 * g + 1];
    s.VB[2 * i] = a.vb[2 * g];
    s.VB[2 * i + 1] = a.vb[2 * g + 1];
  }
  for (int e = tid; e < ne; e += bs) s.THB[e] = a.thb[s.edges[e]];
  // the rest-angle adjoint continues the running sum, so the summation order
  // does not depend on where the checkpoints fall
  for (int b = tid; b < nb; b += bs) s.RB[b] = a.rest_bar[s.hinges[b]];
  for (int e = tid; e < ne; e += bs) s.LB[e] = lbar ? a.len_bar[s.edges[e]] : 0.0f;
  __syncthreads();
  const float kst = c.k_st, ksh = c.k_sh;
  const bool is_cos = c.solver == 1;
  for (int k = steps - 1; k >= 0; --k) {
    // adjoint of the epilogue: pb = xb + vb / dt
    for (int i = tid; i < 2 * nn; i += bs) {
      s.XB[i] = s.XB[i] + s.VB[i] / c.dt;
      s.YB[i] = 0.0f;
    }
    __syncthreads();
    if (is_cos) {
      cosserat_cycles_adjoint(c, s, tp + (size_t)k * C * ne,
                              up ? up + (size_t)k * C * nn : nullptr);
    } else
    for (int cyc = C - 1; cyc >= 0; --cyc) {
      const float4* rc = tp + ((size_t)k * C + cyc) * ne;
      for (int e = tid; e < ne; e += bs) {
        const float4 q = rc[e];  // (rx, ry, cos th', sin th')
        const int sn = s.esrc[e], tn = s.etgt[e];
        const float ds = fmax_nan((float)(s.ninc_off[sn + 1] - s.ninc_off[sn]), 1.0f);
        const float dg = fmax_nan((float)(s.ninc_off[tn + 1] - s.ninc_off[tn]), 1.0f);
        const float fbx = s.XB[2 * sn] / ds - s.XB[2 * tn] / dg;
        const float fby = s.XB[2 * sn + 1] / ds - s.XB[2 * tn + 1] / dg;
        const float rx = q.x, ry = q.y;
        const float l = s.LE[e];
        const float len = sqrtf(rx * rx + ry * ry);
        const float ls = fmax_nan(len, 1e-8f);
        const float hx = rx / ls, hy = ry / ls;
        const float al = 1.0f - l / ls, bl = l / ls;
        const float pr = hx * fbx + hy * fby;
        const float sx = al * fbx + bl * pr * hx, sy = al * fby + bl * pr * hy;
        s.FE[2 * e] = 0.5f * kst * ((1.0f - ksh) * sx + ksh * fbx);
        s.FE[2 * e + 1] = 0.5f * kst * ((1.0f - ksh) * sy + ksh * fby);
        s.THB[e] = s.THB[e] + (-0.5f * kst * ksh * l) * (-q.w * fbx + q.z * fby);
        // rest length: dF/dl0 = -k_st/2 ((1 - k_sh) rhat + k_sh (cos th', sin th'))
        if (lbar)
          s.LB[e] = s.LB[e] - 0.5f * kst * ((1.0f - ksh) * pr + ksh * (q.z * fbx + q.w * fby));
      }
      __syncthreads();
      // per node: + rb at the edge's target, - rb at its source
      for (int i = tid; i < nn; i += bs) {
        const int i0 = s.ninc_off[i], i1 = s.ninc_off[i + 1];
        float ux = 0.0f, uy = 0.0f;
        for (int q = i0; q < i1; ++q) {
          const int code = s.node_inc[q];
          const int e = code >> 1;
          ux = ux + ((code & 1) ? s.FE[2 * e] : -s.FE[2 * e]);
          uy = uy + ((code & 1) ? s.FE[2 * e + 1] : -s.FE[2 * e + 1]);
        }
        s.XB[2 * i] = s.XB[2 * i] + ux;
        s.XB[2 * i + 1] = s.XB[2 * i + 1] + uy;
      }
      // angle pass, per hinge
      for (int b = tid; b < nb; b += bs) {
        const int ei = s.hin[b], eo = s.hout[b];
        const float di = fmax_nan((float)(s.einc_off[ei + 1] - s.einc_off[ei]), 1.0f);
        const float dout = fmax_nan((float)(s.einc_off[eo + 1] - s.einc_off[eo]), 1.0f);
        const float corrb = 1.8f * (s.THB[ei] / di - s.THB[eo] / dout);
        const float db = corrb * s.KB[b] * 0.5f;
        s.DB[b] = db;
        s.RB[b] = s.RB[b] - db;
      }
      __syncthreads();
      for (int e = tid; e < ne; e += bs) {
        const int i0 = s.einc_off[e], i1 = s.einc_off[e + 1];
        float d = 0.0f;
        for (int q = i0; q < i1; ++q) {
          const int code = s.edge_inc[q];
          d = d + ((code & 1) ? s.DB[code >> 1] : -s.DB[code >> 1]);
        }
        s.THB[e] = s.THB[e] + d;
      }
      __syncthreads();
    }
    // adjoint of the prologue: xb = pb - vb / dt, vb = pb d dt (pb includes
    // the Cosserat inertial target's share, y = the prediction).  With the
    // steric term, pb (of the pushed prediction) goes to a.pb1 instead.
    if (a.pb1) {
      for (int i = tid; i < nn; i += bs) {
        const int g = s.nodes[i];
        a.pb1[2 * g] = s.XB[2 * i] + s.YB[2 * i];
        a.pb1[2 * g + 1] = s.XB[2 * i + 1] + s.YB[2 * i + 1];
      }
    } else {
      for (int i = tid; i < 2 * nn; i += bs) {
        const float pb = s.XB[i] + s.YB[i];
        s.XB[i] = pb - s.VB[i] / c.dt;
        s.VB[i] = (pb * c.damping) * c.dt;
      }
    }
    __syncthreads();
  }
  if (!a.pb1)
    for (int i = tid; i < nn; i += bs) {
      const int g = s.nodes[i];
      a.xb[2 * g] = s.XB[2 * i];
      a.xb[2 * g + 1] = s.XB[2 * i + 1];
      a.vb[2 * g] = s.VB[2 * i];
      a.vb[2 * g + 1] = s.VB[2 * i + 1];
    }
  for (int e = tid; e < ne; e += bs) a.thb[s.edges[e]] = s.THB[e];
  for (int b = tid; b < nb; b += bs) a.rest_bar[s.hinges[b]] = s.RB[b];
  if (lbar)
    for (int e = tid; e < ne; e += bs) a.len_bar[s.edges[e]] = s.LB[e];
}

// ---- steric term ----
// f = -grad_FD V(p) - c sum_j K(r_pj) r_pj with V = G * D(p) (forces.py:20-118):
// G the Gaussian filter (symmetric, so its adjoint is G itself), D the CIC
// deposit, the gradient by central differences of bilinear samples at
// p +- 0.5 e_x, p +- 0.5 e_y, c = sv / sigma^2.  Given g = pb1 strength dt
// per node, the adjoint f^T g has three parts:
//   field samples through V:  Vbar = sum_i sum_samples -sgn g_comp w_corner
//   (a deposit of the samples' bilinear weights), Dbar = G Vbar, and per node
//   sv sum_corners Dbar dw_corner/dp;
//   field samples directly:   -J_i^T g_i with J the 2x2 derivative of the
//   central differences (bilinear slopes of V);
//   window term:              c sum_j K (I - r r^T / sigma^2)(g_j - g_i).

constexpr int BAR_THREADS = 256;
constexpr int BAR_LIST = 1024;  // staged nodes per flush

// v mod n for v in [-n, 2n) -- the cells of samples within one cell of a
// wrapped position (in [0, n], rounding included) -- without the integer
// divisions of imod; clamped so that no input (a NaN or saturated position)
// indexes outside [0, n).
__device__ __forceinline__ int wrap1(int v, int n) {
  v = v < 0 ? v + n : (v >= n ? v - n : v);
  return v < 0 ? 0 : (v >= n ? n - 1 : v);
}

__device__ __forceinline__ bool touches_row(float py, int H, int y) {
  const int a = f2i(floorf(py - 0.5f)), b = f2i(floorf(py)), d = f2i(floorf(py + 0.5f));
  return wrap1(a, H) == y || wrap1(a + 1, H) == y || wrap1(b, H) == y || wrap1(b + 1, H) == y ||
         wrap1(d, H) == y || wrap1(d + 1, H) == y;
}

// acc += q * (bilinear weight of cell (y, x) in the sample at (yy, xx))
__device__ __forceinline__ float bar_sample(float acc, float q, float yy, float xx, int H, int W,
                                            int y, int x) {
  const float ly = floorf(yy), lx = floorf(xx);
  const float wyh = yy - ly, wxh = xx - lx;
  const float wyl = 1.0f - wyh, wxl = 1.0f - wxh;
  const int iy = f2i(ly), ix = f2i(lx);
  const int y0 = wrap1(iy, H), y1 = wrap1(iy + 1, H);
  const int x0 = wrap1(ix, W), x1 = wrap1(ix + 1, W);
  if (y0 == y && x0 == x) acc = acc + q * (wyl * wxl);
  if (y0 == y && x1 == x) acc = acc + q * (wyl * wxh);
  if (y1 == y && x0 == x) acc = acc + q * (wyh * wxl);
  if (y1 == y && x1 == x) acc = acc + q * (wyh * wxh);
  return acc;
}

// One workgroup per steric row: the nodes whose samples reach the row are
// staged in node order, then each thread gathers its columns' Vbar from them
// in that order (deterministic, no atomics).
__global__ void __launch_bounds__(BAR_THREADS) k_steric_bar_rows(Consts c, BatchDev bd,
    const float* __restrict__ pred_w, const float* __restrict__ pb1, float* __restrict__ vbar) {
  __shared__ float4 lst[BAR_LIST];  // (px, py, gx, gy)
  __shared__ int wcnt[BAR_THREADS / 64];
  __shared__ int ltop;
  const int r = blockIdx.x;
  const InstGeo gi = bd.geo[bd.row_inst[r]];
  const int y = r - gi.row0, H = gi.H, W = gi.W;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const float sdt = c.steric_strength * c.dt;
  float* out = vbar + gi.grid_off + (size_t)y * W;
  for (int x0 = 0; x0 < W; x0 += BAR_THREADS) {
    const int x = x0 + tid;
    float acc = 0.0f;
    if (tid == 0) ltop = 0;
    __syncthreads();
    for (int i0 = 0;; i0 += BAR_THREADS) {
      const bool more = i0 < gi.nodes;
      if (more) {
        const int i = i0 + tid;
        float4 e = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        bool hit = false;
        if (i < gi.nodes) {
          const int g = gi.node0 + i;
          const float2 p = *reinterpret_cast<const float2*>(pred_w + 2 * g);
          hit = touches_row(p.y, H, y);
          e = make_float4(p.x, p.y, pb1[2 * g] * sdt, pb1[2 * g + 1] * sdt);
        }
        const unsigned long long m = __ballot(hit);
        if (lane == 0) wcnt[wv] = __popcll(m);
        __syncthreads();
        int base = ltop;
        for (int w = 0; w < wv; ++w) base += wcnt[w];
        if (hit) lst[base + __popcll(m & ((1ull << lane) - 1ull))] = e;
        __syncthreads();
        if (tid == 0) {
          int t = ltop;
          for (int w = 0; w < BAR_THREADS / 64; ++w) t += wcnt[w];
          ltop = t;
        }
        __syncthreads();
      }
      // gather from the staged nodes when the list is nearly full or complete
      if (!more || ltop > BAR_LIST - BAR_THREADS) {
        const int n = ltop;
        if (x < W)
          for (int k0 = 0; k0 < n; k0 += 8) {
            // eight staged nodes read ahead of their use (LDS latency overlaps)
            float4 qv[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) qv[u] = lst[min(k0 + u, n - 1)];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
              if (k0 + u >= n) break;
              const float4 q = qv[u];
              // the samples' cells lie in columns floor(px) - 1 .. floor(px) + 2
              int d = x - f2i(floorf(q.x)) + 1;
              d = d < 0 ? d + W : (d >= W ? d - W : d);
              if (d > 3) continue;
              acc = bar_sample(acc, -q.z, q.y, q.x + 0.5f, H, W, y, x);
              acc = bar_sample(acc, q.z, q.y, q.x - 0.5f, H, W, y, x);
              acc = bar_sample(acc, -q.w, q.y + 0.5f, q.x, H, W, y, x);
              acc = bar_sample(acc, q.w, q.y - 0.5f, q.x, H, W, y, x);
            }
          }
        __syncthreads();
        if (tid == 0) ltop = 0;
        __syncthreads();
      }
      if (!more) break;
    }
    if (x < W) out[x] = acc;
    __syncthreads();
  }
}

// Bilinear slopes (d/dx, d/dy) of the wrapped sample of V at (yy, xx).
__device__ __forceinline__ void sample_slopes(const float* V, int H, int W, float yy, float xx,
                                              float& sx, float& sy) {
  const float ly = floorf(yy), lx = floorf(xx);
  const float wyh = yy - ly, wxh = xx - lx;
  const float wyl = 1.0f - wyh, wxl = 1.0f - wxh;
  const int iy = f2i(ly), ix = f2i(lx);
  const int y0 = wrap1(iy, H), y1 = wrap1(iy + 1, H);
  const int x0 = wrap1(ix, W), x1 = wrap1(ix + 1, W);
  const float v00 = V[(size_t)y0 * W + x0], v01 = V[(size_t)y0 * W + x1];
  const float v10 = V[(size_t)y1 * W + x0], v11 = V[(size_t)y1 * W + x1];
  sx = wyl * (v01 - v00) + wyh * (v11 - v10);
  sy = wxl * (v10 - v00) + wxh * (v11 - v01);
}

// Per node: the direct sample term, the deposit term (from Dbar = G Vbar) and
// the window term; then the prologue adjoint xb = pb0 - vb / dt, vb = pb0 d dt.
__global__ void k_steric_adj_nodes(Consts c, BatchDev bd, int N, const float* __restrict__ pred_w,
                                   const float* __restrict__ V_all, const float* __restrict__ D_all,
                                   const float* __restrict__ pb1, float* __restrict__ xb,
                                   float* __restrict__ vb) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  const InstGeo gi = bd.geo[bd.node_inst[n]];
  const int H = gi.H, W = gi.W;
  const float* V = V_all + gi.grid_off;
  const float* D = D_all + gi.grid_off;
  const float sdt = c.steric_strength * c.dt;
  const float px = pred_w[2 * n], py = pred_w[2 * n + 1];
  const float gx = pb1[2 * n] * sdt, gy = pb1[2 * n + 1] * sdt;
  float ox = 0.0f, oy = 0.0f;
  // direct: gx_f = S(y, x+.5) - S(y, x-.5), gy_f = S(y+.5, x) - S(y-.5, x)
  float ax, ay, bx, by;
  sample_slopes(V, H, W, py, px + 0.5f, ax, ay);
  sample_slopes(V, H, W, py, px - 0.5f, bx, by);
  const float j00 = ax - bx, j01 = ay - by;
  sample_slopes(V, H, W, py + 0.5f, px, ax, ay);
  sample_slopes(V, H, W, py - 0.5f, px, bx, by);
  const float j10 = ax - bx, j11 = ay - by;
  ox = ox - (j00 * gx + j10 * gy);
  oy = oy - (j01 * gx + j11 * gy);
  // deposit: sv sum_c Dbar[c] dw_c / dp
  const float fx = floorf(px), fy = floorf(py);
  const float dx = px - fx, dy = py - fy;
  const int x0 = wrap1(f2i(fx), W), y0 = wrap1(f2i(fy), H);
  const int x1 = wrap1(x0 + 1, W), y1 = wrap1(y0 + 1, H);
  const float d00 = D[(size_t)y0 * W + x0], d10 = D[(size_t)y0 * W + x1];
  const float d01 = D[(size_t)y1 * W + x0], d11 = D[(size_t)y1 * W + x1];
  ox = ox + c.steric_sv * ((1.0f - dy) * (d10 - d00) + dy * (d11 - d01));
  oy = oy + c.steric_sv * ((1.0f - dx) * (d01 - d00) + dx * (d11 - d10));
  // window (flat node index within the instance)
  const int lo = gi.node0, hi = gi.node0 + gi.nodes;
  const float inv_s2 = 2.0f / c.steric_2s2;
  float wx = 0.0f, wy = 0.0f;
  for (int k = -c.steric_nskip; k <= c.steric_nskip; ++k) {
    const int j = n + k;
    if (j < lo || j >= hi) continue;
    const float rx = min_image(px, pred_w[2 * j], gi.dom_wf);
    const float ry = min_image(py, pred_w[2 * j + 1], gi.dom_hf);
    const float Kv = fexp(-(rx * rx + ry * ry) / c.steric_2s2) / c.steric_norm;
    const float dgx = pb1[2 * j] * sdt - gx, dgy = pb1[2 * j + 1] * sdt - gy;
    const float rd = (rx * dgx + ry * dgy) * inv_s2;
    wx = wx + Kv * (dgx - rx * rd);
    wy = wy + Kv * (dgy - ry * rd);
  }
  ox = ox + c.steric_svs * wx;
  oy = oy + c.steric_svs * wy;
  const float p0x = pb1[2 * n] + ox, p0y = pb1[2 * n + 1] + oy;
  const float vbx = vb[2 * n], vby = vb[2 * n + 1];
  xb[2 * n] = p0x - vbx / c.dt;
  xb[2 * n + 1] = p0y - vby / c.dt;
  vb[2 * n] = (p0x * c.damping) * c.dt;
  vb[2 * n + 1] = (p0y * c.damping) * c.dt;
}

}  // namespace

namespace {
__global__ void k_add(float* __restrict__ y, const float* __restrict__ x, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = y[i] + x[i];
}
}  // namespace

hipError_t launch_add(float* y, const float* x, size_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_add, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, y, x, n);
  return hipGetLastError();
}

hipError_t launch_steric_bar(const Consts& c, const BatchDev& b, const float* pred_w,
                             const float* pb1, float* vbar, hipStream_t s) {
  if (b.rows == 0) return hipSuccess;
  hipLaunchKernelGGL(k_steric_bar_rows, dim3(b.rows), dim3(BAR_THREADS), 0, s, c, b, pred_w, pb1,
                     vbar);
  return hipGetLastError();
}

hipError_t launch_steric_adj_nodes(const Consts& c, const BatchDev& b, int N, const float* pred_w,
                                   const float* V, const float* dbar, const float* pb1, float* xb,
                                   float* vb, hipStream_t s) {
  if (N == 0) return hipSuccess;
  hipLaunchKernelGGL(k_steric_adj_nodes, dim3((N + 255) / 256), dim3(256), 0, s, c, b, N, pred_w,
                     V, dbar, pb1, xb, vb);
  return hipGetLastError();
}

size_t rod_segment_lds_bytes(int cycles, int nn, int ne, int nb, bool reverse) {
  (void)cycles;
  size_t f = 9 * (size_t)nn + 4 * (size_t)ne + 3 * (size_t)nb;
  if (reverse) f += 6 * (size_t)nn + 2 * (size_t)ne + 2 * (size_t)nb;
  return f * sizeof(float);
}

hipError_t launch_rod_segment(const RodAdjArgs& a, bool reverse, int steps, float* ck,
                              float* tape_or_dbg, size_t lds_bytes, hipStream_t s) {
  if (a.L.num_comp == 0) return hipSuccess;
  const void* fn = reverse ? reinterpret_cast<const void*>(&k_rod_adj_segment)
                           : reinterpret_cast<const void*>(&k_rod_fwd_segment);
  hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
  if (e != hipSuccess) return e;
  if (reverse)
    hipLaunchKernelGGL(k_rod_adj_segment, dim3(a.L.num_comp), dim3(64), lds_bytes, s, a, steps,
                       (const float*)ck, tape_or_dbg);
  else
    hipLaunchKernelGGL(k_rod_fwd_segment, dim3(a.L.num_comp), dim3(64), lds_bytes, s, a, steps,
                       ck, tape_or_dbg);
  return hipGetLastError();
}

}  // namespace mc
