// Rod solve of one coupled step (simulate.py:41-84 with pbd.py:14-91 or
// stable_cosserat.py:13-119): one workgroup per connected component of the
// edge graph (a filament), all constraint cycles in one launch.
//
// Per component the node positions, edge angles and the per-edge / per-hinge
// intermediates live in LDS (global scratch when a component exceeds the LDS
// budget).  The reference's Jacobi scatters become ordered gathers through the
// incidence lists built on the host (src adds then tgt adds for nodes; e_in
// adds then e_out adds for edges), so every accumulation is the reference's
// left fold, with no atomics and no colouring needed: Jacobi reads only the
// previous iterate, and each node / edge is owned by one thread.
//
// The kernel also performs the step's node prologue -- damping, inertial
// prediction, steric sampling (forces.py:20-81) and the steric push -- and
// its epilogue, v = (x - x_prev) / dt.
#include <hip/hip_ext.h>

#include "mc_common.h"
#include "mc_internal.h"
#include "mc_lbm.h"

namespace mc {
namespace {

struct RodArgs {
  Consts c;
  BatchDev bd;
  RodLayoutDev L;
  int N;
  int mode;
  int use_lds;
  float* gscratch;
  const float *pos_in, *vel_in, *y_in, *theta_in, *rest_len, *rest_ang, *kb;
  Actuation act;
  int has_act;   // sine actuation of the rest angles (act.kind 0)
  int has_rows;  // per-step action rows added to the rest values (act.kind 1)
  const float* steric;
  const float* pred_w;
  float *pos_out, *vel_out, *debug_out, *theta_out;
};

// The step's action row of an instance (kind 1 actuation), and the rest
// values with that row added: base + action in float32 (gym/base.py:169-173).
__device__ __forceinline__ long long action_row(const RodArgs& a, int inst) {
  return a.has_rows ? *a.act.t + (a.act.t_inst ? a.act.t_inst[inst] : 0LL) : 0LL;
}
__device__ __forceinline__ float rest_length(const RodArgs& a, int ge, long long row) {
  const float l = a.rest_len[ge];
  return (a.has_rows && a.act.d_rest) ? l + a.act.d_rest[row * a.act.e_stride + ge] : l;
}
__device__ __forceinline__ float rest_angle(const RodArgs& a, int gb, long long row) {
  const float r = a.rest_ang[gb];
  return (a.has_rows && a.act.d_ang) ? r + a.act.d_ang[row * a.act.b_stride + gb] : r;
}

// Bilinear sample with index wrap (map_coordinates order 1, mode "wrap").
// The sample points are wrapped predictions +- 0.5, so the taps lie in
// [-1, side + 1] and wrap with selects.
__device__ __forceinline__ float sample_bilinear_wrap(const float* V, int H, int W, float yy, float xx) {
  const float ly = floorf(yy), lx = floorf(xx);
  const float wyh = yy - ly, wxh = xx - lx;
  const float wyl = 1.0f - wyh, wxl = 1.0f - wxh;
  const int iy = f2i(ly), ix = f2i(lx);
  const int y0 = wrap1c(iy, H), y1 = wrap1c(iy + 1, H);
  const int x0 = wrap1c(ix, W), x1 = wrap1c(ix + 1, W);
  float out = (wyl * wxl) * V[(size_t)y0 * W + x0];
  out = out + (wyl * wxh) * V[(size_t)y0 * W + x1];
  out = out + (wyh * wxl) * V[(size_t)y1 * W + x0];
  out = out + (wyh * wxh) * V[(size_t)y1 * W + x1];
  return out;
}

// The analytic window term of forces.py:66-81 for node g at (px, py): the
// (2n+1)-window over the flat node index (within the node's instance), summed
// in window order.  Its terms are independent (their displacements, kernel
// values and exponentials overlap, four at a time, with no branch between
// them): EXACT = false takes the minimum images on their common range and
// returns whether a lane left it (the caller then repeats with EXACT = true);
// FROM_LDS reads the window's predictions from pw (staged from node pw0).
template <bool EXACT, bool FROM_LDS>
__device__ __forceinline__ bool steric_window(const Consts& c, const InstGeo& gi,
                                              const float* pred_w, const float2* pw, int pw0,
                                              int g, float px, float py, float& ax, float& ay) {
  const int lo = gi.node0, hi = gi.node0 + gi.nodes;
  const int nsk = c.steric_nskip;
  bool slow = false;
  ax = 0.0f;
  ay = 0.0f;
  constexpr int WU = 4;
  for (int k0 = -nsk; k0 <= nsk; k0 += WU) {
    float rxv[WU], ryv[WU], kvv[WU];
#pragma unroll
    for (int u = 0; u < WU; ++u) {
      const int j = g + min(k0 + u, nsk);  // past the window: a repeat, not summed
      const float valid = (j >= lo && j < hi) ? 1.0f : 0.0f;
      const int js = j < lo ? lo : (j > hi - 1 ? hi - 1 : j);
      const float2 q =
          FROM_LDS ? pw[js - pw0] : *reinterpret_cast<const float2*>(pred_w + 2 * js);
      if (EXACT) {
        rxv[u] = min_image(px, q.x, gi.dom_wf);
        ryv[u] = min_image(py, q.y, gi.dom_hf);
      } else {
        rxv[u] = min_image_fast(px, q.x, gi.dom_wf, slow);
        ryv[u] = min_image_fast(py, q.y, gi.dom_hf, slow);
      }
      const float d2 = rxv[u] * rxv[u] + ryv[u] * ryv[u];
      const float Kv = fexp(-d2 / c.steric_2s2) / c.steric_norm;
      kvv[u] = Kv * valid;
    }
#pragma unroll
    for (int u = 0; u < WU; ++u)
      if (k0 + u <= nsk) {
        ax = ax + kvv[u] * rxv[u];
        ay = ay + kvv[u] * ryv[u];
      }
  }
  return slow;
}

// forces.py:20-81 for one node: -grad V by central differences minus the
// analytic (2n+1)-window self/neighbour contribution over the flat node index
// (within the node's instance: an instance is one reference simulation).
// pw (may be NULL): the predicted positions of nodes pw0 .. in LDS, covering
// every window node of g after clamping to the instance.
__device__ __forceinline__ void steric_force(const Consts& c, const InstGeo& gi, const float* V_all,
                                             const float* pred_w, int g, float& fx, float& fy,
                                             const float2* pw = nullptr, int pw0 = 0) {
  const int H = gi.H, W = gi.W;
  const float* V = V_all + gi.grid_off;
  // the node's own prediction: from the staged window when there is one (an
  // LDS read instead of a dependent global load before the field samples)
  const float2 pg = pw ? pw[g - pw0] : *reinterpret_cast<const float2*>(pred_w + 2 * g);
  const float px = pg.x, py = pg.y;
  const float gx = (sample_bilinear_wrap(V, H, W, py, px + 0.5f) -
                    sample_bilinear_wrap(V, H, W, py, px - 0.5f)) / 1.0f;
  const float gy = (sample_bilinear_wrap(V, H, W, py + 0.5f, px) -
                    sample_bilinear_wrap(V, H, W, py - 0.5f, px)) / 1.0f;
  float ax, ay;
  bool slow = pw ? steric_window<false, true>(c, gi, pred_w, pw, pw0, g, px, py, ax, ay)
                 : steric_window<false, false>(c, gi, pred_w, pw, pw0, g, px, py, ax, ay);
  if (slow) {
    if (pw) steric_window<true, true>(c, gi, pred_w, pw, pw0, g, px, py, ax, ay);
    else steric_window<true, false>(c, gi, pred_w, pw, pw0, g, px, py, ax, ay);
  }
  fx = -gx - c.steric_svs * ax;
  fy = -gy - c.steric_svs * ay;
}

constexpr int ROD_WIN_MAX = 8;  // steric windows up to this half-width are staged in LDS

__device__ __forceinline__ void rod_component(const RodArgs& a, int comp, float* smem) {
  const Consts& c = a.c;
  const RodLayoutDev& L = a.L;
  const CompMeta m = L.meta[comp];
  const InstGeo& gi = a.bd.geo[m.inst];
  const int n0 = m.n0, nn = m.nn, e0 = m.e0, ne = m.ne, b0 = m.b0, nb = m.nb;
  const int nni = m.nni, nei = m.nei;
  // LDS: the steric window's predicted positions pw[nn + 2 ROD_WIN_MAX], then
  // float scratch: x[2nn] y[2nn] th[ne] l[ne] F[2ne] He[ne] corr[nb] rest[nb] kb[nb]
  float2* pw = reinterpret_cast<float2*>(smem);
  float* base = a.use_lds ? smem + 2 * (nn + 2 * ROD_WIN_MAX)
                          : (a.gscratch + (size_t)4 * n0 + (size_t)5 * e0 + (size_t)3 * b0);
  float* X = base;
  float* Y = X + 2 * nn;
  float* TH = Y + 2 * nn;
  float* Le = TH + ne;
  float* Fe = Le + ne;
  float* He = Fe + 2 * ne;
  float* CO = He + ne;
  float* RS = CO + nb;
  float* KB = RS + nb;
  const bool is_cos = c.solver == 1;
  const int tid = threadIdx.x, bs = blockDim.x;
  // ids of this thread's first node / edge / hinge (no load for components
  // whose ids are one contiguous range)
  const int nf = m.nf, ef = m.ef, hf = m.hf;
  auto node_id = [&](int i) { return nf >= 0 ? nf + i : L.comp_nodes[n0 + i]; };
  auto edge_id = [&](int e) { return ef >= 0 ? ef + e : L.comp_edges[e0 + e]; };
  auto hinge_id = [&](int b) { return hf >= 0 ? hf + b : L.comp_hinges[b0 + b]; };

  // Every global load that depends on nothing else is issued first, then
  // the topology copy, then the dependent steric samples.  A contiguous
  // component stages the predicted positions of its steric window (its nodes
  // and nsk on either side, within the instance) in LDS: one coalesced load
  // instead of 2 nsk + 1 dependent ones per node.
  const int nsk = c.steric_nskip;
  const bool stage_pw = a.use_lds && a.mode == 0 && c.enable_steric && nf >= 0 &&
                        nsk <= ROD_WIN_MAX && nn + 2 * nsk <= 2 * bs;
  const int pw0 = nf - nsk;
  float2 pwv[2];  // entries tid and tid + bs, stored after the other loads are issued
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int i = tid + h * bs, j = pw0 + i;
    pwv[h] = (stage_pw && i < nn + 2 * nsk && j >= gi.node0 && j < gi.node0 + gi.nodes)
                 ? *reinterpret_cast<const float2*>(a.pred_w + 2 * (size_t)j)
                 : make_float2(0.0f, 0.0f);
  }
  const double tact =
      a.has_act ? (double)(*a.act.t + (a.act.t_inst ? a.act.t_inst[m.inst] : 0LL)) : 0.0;
  const long long arow = action_row(a, m.inst);
  int g = -1;
  float px = 0.0f, py = 0.0f, vx = 0.0f, vy = 0.0f, yx = 0.0f, yy = 0.0f;
  if (tid < nn) {
    g = node_id(tid);
    px = a.pos_in[2 * g];
    py = a.pos_in[2 * g + 1];
    if (a.mode == 1) {
      yx = a.y_in ? a.y_in[2 * g] : px;
      yy = a.y_in ? a.y_in[2 * g + 1] : py;
    } else {
      vx = a.vel_in[2 * g];
      vy = a.vel_in[2 * g + 1];
    }
  }
  float th0 = 0.0f, le0 = 0.0f;
  if (tid < ne) {
    const int ge = edge_id(tid);
    th0 = a.theta_in[ge];
    le0 = rest_length(a, ge, arow);
  }
  float am = 0.0f, om = 0.0f, ph = 0.0f, rs0 = 0.0f, kb0 = 0.0f;
  if (tid < nb) {
    const int gb = hinge_id(tid);
    if (a.has_act) {
      am = a.act.amp[gb];
      om = a.act.omega[gb];
      ph = a.act.phase[gb];
    } else {
      rs0 = rest_angle(a, gb, arow);
    }
    kb0 = a.kb[gb];
  }
  // topology: one contiguous block per component, staged into LDS after the
  // floats (component-relative offsets), or read in place from global memory
  // for components beyond the LDS budget
  const int *hin, *hout, *esrc, *etgt, *einc_off, *einc, *ninc_off, *ninc;
  if (a.use_lds) {
    int* ib = reinterpret_cast<int*>(KB + nb);
    // blocks start on 16 bytes: 16-byte loads, all in flight before the stores
    const int4* blob4 = reinterpret_cast<const int4*>(L.blob + m.blob_off);
    const int len = 2 * nb + 3 * ne + 1 + nei + nn + 1 + nni;
    const int len4 = (len + 3) >> 2;
    for (int q0 = 0; q0 < len4; q0 += 4 * bs) {
      int4 v[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int q = q0 + k * bs + tid;
        v[k] = q < len4 ? blob4[q] : make_int4(0, 0, 0, 0);
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int q = 4 * (q0 + k * bs + tid);
        if (q < len) ib[q] = v[k].x;
        if (q + 1 < len) ib[q + 1] = v[k].y;
        if (q + 2 < len) ib[q + 2] = v[k].z;
        if (q + 3 < len) ib[q + 3] = v[k].w;
      }
    }
    hin = ib;
    hout = hin + nb;
    esrc = hout + nb;
    etgt = esrc + ne;
    einc_off = etgt + ne;
    einc = einc_off + ne + 1;
    ninc_off = einc + nei;
    ninc = ninc_off + nn + 1;
  } else {
    hin = L.hinge_in + b0; hout = L.hinge_out + b0; esrc = L.edge_src + e0; etgt = L.edge_tgt + e0;
    einc_off = L.edge_inc_off + e0; einc = L.edge_inc; ninc_off = L.node_inc_off + n0; ninc = L.node_inc;
  }

  // prologue: damping, prediction, steric push (simulate.py:41-72)
  auto node_prologue = [&](int i, int gg, float qx, float qy, float ux, float uy, float zx,
                           float zy) {
    if (a.mode == 1) {
      X[2 * i] = qx;
      X[2 * i + 1] = qy;
      Y[2 * i] = zx;
      Y[2 * i + 1] = zy;
      return;
    }
    const float dvx = ux * c.damping, dvy = uy * c.damping;
    float x = qx + dvx * c.dt, y = qy + dvy * c.dt;
    if (c.enable_steric) {
      float fx, fy;
      steric_force(c, gi, a.steric, a.pred_w, gg, fx, fy, stage_pw ? pw : nullptr, pw0);
      x = x + (fx * c.steric_strength) * c.dt;
      y = y + (fy * c.steric_strength) * c.dt;
    }
    // the inertial target y = debug_vector is the prediction only under
    // solver_type "cosserat" (simulate.py:41,71-72), else it stays zero
    const float tx = c.cos_target ? x : 0.0f, ty = c.cos_target ? y : 0.0f;
    X[2 * i] = x;
    X[2 * i + 1] = y;
    Y[2 * i] = tx;
    Y[2 * i + 1] = ty;
    a.debug_out[2 * gg] = tx;
    a.debug_out[2 * gg + 1] = ty;
  };
  auto hinge_rest = [&](float amp_, float om_, float ph_, float rs_) {
    if (!a.has_act) return rs_;
    const double arg = (double)om_ * tact + (double)ph_;
    return (float)((double)amp_ * sin(arg));
  };
  if (stage_pw) {
#pragma unroll
    for (int h = 0; h < 2; ++h)
      if (tid + h * bs < nn + 2 * nsk) pw[tid + h * bs] = pwv[h];
    __syncthreads();  // the staged window is complete
  }
  if (tid < nn) node_prologue(tid, g, px, py, vx, vy, yx, yy);
  if (tid < ne) {
    TH[tid] = th0;
    Le[tid] = le0;
  }
  if (tid < nb) {
    RS[tid] = hinge_rest(am, om, ph, rs0);
    KB[tid] = kb0;
  }
  // components with more nodes / edges / hinges than threads
  for (int i = tid + bs; i < nn; i += bs) {
    const int gg = node_id(i);
    const float qx = a.pos_in[2 * gg], qy = a.pos_in[2 * gg + 1];
    if (a.mode == 1)
      node_prologue(i, gg, qx, qy, 0.0f, 0.0f, a.y_in ? a.y_in[2 * gg] : qx,
                    a.y_in ? a.y_in[2 * gg + 1] : qy);
    else
      node_prologue(i, gg, qx, qy, a.vel_in[2 * gg], a.vel_in[2 * gg + 1], 0.0f, 0.0f);
  }
  for (int e = tid + bs; e < ne; e += bs) {
    const int ge = edge_id(e);
    TH[e] = a.theta_in[ge];
    Le[e] = rest_length(a, ge, arow);
  }
  for (int b = tid + bs; b < nb; b += bs) {
    const int gb = hinge_id(b);
    RS[b] = a.has_act ? hinge_rest(a.act.amp[gb], a.act.omega[gb], a.act.phase[gb], 0.0f)
                      : rest_angle(a, gb, arow);
    KB[b] = a.kb[gb];
  }
  __syncthreads();

  const float Wd = gi.dom_wf, Hd = gi.dom_hf;
  for (int cyc = 0; cyc < c.cycles; ++cyc) {
    // theta pass: hinge corrections (pbd.py:32-40)
    for (int b = threadIdx.x; b < nb; b += blockDim.x) {
      const float diff = wrap_angle(TH[hout[b]] - TH[hin[b]]);
      CO[b] = ((diff - RS[b]) * KB[b]) * 0.5f;
    }
    __syncthreads();
    // per edge: ordered gather of +-corr and the angle update (pbd.py:44-49 /
    // stable_cosserat.py:158-171), then the edge force with the new angle
    // (pbd.py:60-81 / stable_cosserat.py:185-195) -- one thread owns the edge
    for (int e = threadIdx.x; e < ne; e += blockDim.x) {
      const int i0 = einc_off[e], i1 = einc_off[e + 1];
      float d = 0.0f;
      for (int q = i0; q < i1; ++q) {
        const int code = einc[q];
        const float v = CO[code >> 1];
        d = d + ((code & 1) ? -v : v);
      }
      const float deg = fmax_nan((float)(i1 - i0), 1.0f);
      const int sn = esrc[e], tn = etgt[e];
      const float rx = min_image(X[2 * tn], X[2 * sn], Wd);
      const float ry = min_image(X[2 * tn + 1], X[2 * sn + 1], Hd);
      float th;
      if (!is_cos) {
        th = TH[e] + (d / deg) * 1.8f;
      } else {
        th = TH[e] + d / deg;
        const float alpha = datan2(ry, rx);
        const float pull = wrap_angle(alpha - th);
        th = th + pull * 0.05f;
      }
      TH[e] = th;
      const float l = Le[e];
      float dx, dy;
      fsincos(th, &dy, &dx);
      if (!is_cos) {
        const float tx = dx * l, ty = dy * l;
        const float len = sqrtf(rx * rx + ry * ry);
        const float md = fmax_nan(len, 1e-8f);
        const float cdx = rx / md, cdy = ry / md;
        const float stx = cdx * (len - l), sty = cdy * (len - l);
        const float shx = (rx - tx) - stx, shy = (ry - ty) - sty;
        Fe[2 * e] = ((stx + shx * c.k_sh) * c.k_st) * 0.5f;
        Fe[2 * e + 1] = ((sty + shy * c.k_sh) * c.k_st) * 0.5f;
      } else {
        const float cx = rx / l - dx, cy = ry / l - dy;
        const float kl = c.k_st / l;
        Fe[2 * e] = kl * cx;
        Fe[2 * e + 1] = kl * cy;
        He[e] = c.k_st / (l * l);
      }
    }
    __syncthreads();
    // per node ordered gather (pbd.py:84-89 / stable_cosserat.py:197-207)
    for (int i = threadIdx.x; i < nn; i += blockDim.x) {
      const int i0 = ninc_off[i], i1 = ninc_off[i + 1];
      if (!is_cos) {
        float ux = 0.0f, uy = 0.0f;
        for (int q = i0; q < i1; ++q) {
          const int code = ninc[q];
          const int e = code >> 1;
          const float fx = Fe[2 * e], fy = Fe[2 * e + 1];
          ux = ux + ((code & 1) ? -fx : fx);
          uy = uy + ((code & 1) ? -fy : fy);
        }
        const float deg = fmax_nan((float)(i1 - i0), 1.0f);
        X[2 * i] = X[2 * i] + ux / deg;
        X[2 * i + 1] = X[2 * i + 1] + uy / deg;
      } else {
        const float w = c.cos_w;
        const float xx = X[2 * i], xy = X[2 * i + 1];
        float fx = w * (Y[2 * i] - xx), fy = w * (Y[2 * i + 1] - xy);
        float h = w;
        for (int q = i0; q < i1; ++q) {
          const int code = ninc[q];
          const int e = code >> 1;
          const float gx = Fe[2 * e], gy = Fe[2 * e + 1];
          fx = fx + ((code & 1) ? -gx : gx);
          fy = fy + ((code & 1) ? -gy : gy);
        }
        // H: src adds then tgt adds, all +H_edge (same incidence order)
        for (int q = i0; q < i1; ++q) h = h + He[ninc[q] >> 1];
        X[2 * i] = xx + fx / h;
        X[2 * i + 1] = xy + fy / h;
      }
    }
    __syncthreads();
  }

  // epilogue: v = (x - x_prev) / dt (simulate.py:83-84); unwrapped x feeds the
  // immersed boundary, or is wrapped here when the fluid is off.
  for (int i = threadIdx.x; i < nn; i += blockDim.x) {
    const int g = L.comp_nodes[n0 + i];
    const float xx = X[2 * i], xy = X[2 * i + 1];
    if (a.mode == 1) {
      a.pos_out[2 * g] = xx;
      a.pos_out[2 * g + 1] = xy;
      continue;
    }
    const float qx = a.pos_in[2 * g], qy = a.pos_in[2 * g + 1];
    a.vel_out[2 * g] = (xx - qx) / c.dt;
    a.vel_out[2 * g + 1] = (xy - qy) / c.dt;
    if (c.enable_fluid) {
      a.pos_out[2 * g] = xx;
      a.pos_out[2 * g + 1] = xy;
    } else {
      a.pos_out[2 * g] = fmodf_floor(xx, Wd);
      a.pos_out[2 * g + 1] = fmodf_floor(xy, Hd);
    }
  }
  for (int e = threadIdx.x; e < ne; e += blockDim.x) a.theta_out[L.comp_edges[e0 + e]] = TH[e];
}


// DPP wavefront shifts (GFX9 wave_shl:1 / wave_shr:1, every lane active): lane
// i receives the value of lane i + 1 / i - 1; the end lane receives 0.
__device__ __forceinline__ float from_next(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130, 0xf, 0xf, false));
}
__device__ __forceinline__ float from_prev(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xf, 0xf, false));
}

// Rod solve of a chain component (a line filament of at most 64 nodes whose
// node, edge and hinge ids are contiguous, edge e = (node e, node e + 1),
// hinge b = (edge b, edge b + 1)): one wavefront, lane i holding node i,
// edge i and hinge i in registers.  The incidence lists of a chain are fixed:
// edge e takes +corr(e) (as e_in of hinge e) then -corr(e - 1) (as e_out of
// hinge e - 1); node i takes +F(i) (as src of edge i) then -F(i - 1) (as tgt
// of edge i - 1).  The neighbour values come from the adjacent lanes by DPP
// shifts, so a cycle needs no LDS traffic and no barrier.  Every result is
// the general path's, operation for operation (a division by a degree of 1
// or 2 is the exact multiplication by 1 or 0.5).
__device__ __forceinline__ void rod_chain(const RodArgs& a, const CompMeta& m, float* smem) {
  const Consts& c = a.c;
  const InstGeo& gi = a.bd.geo[m.inst];
  const int lane = threadIdx.x;
  const int nn = m.nn, ne = m.ne, nb = m.nb;
  const bool isn = lane < nn, ise = lane < ne, isb = lane < nb;
  const int g = m.nf + lane, ge = m.ef + lane, gb = m.hf + lane;
  const bool is_cos = c.solver == 1;
  const int nsk = c.steric_nskip;
  const bool stage_pw = c.enable_steric && nsk <= ROD_WIN_MAX;
  const int pw0 = m.nf - nsk;
  // every independent global load first
  float2 pwv[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int i = lane + 64 * h, j = pw0 + i;
    pwv[h] = (stage_pw && i < nn + 2 * nsk && j >= gi.node0 && j < gi.node0 + gi.nodes)
                 ? *reinterpret_cast<const float2*>(a.pred_w + 2 * (size_t)j)
                 : make_float2(0.0f, 0.0f);
  }
  float2 q = make_float2(0.0f, 0.0f), v = make_float2(0.0f, 0.0f);
  if (isn) {
    q = *reinterpret_cast<const float2*>(a.pos_in + 2 * (size_t)g);
    v = *reinterpret_cast<const float2*>(a.vel_in + 2 * (size_t)g);
  }
  const long long arow = action_row(a, m.inst);
  float th = 0.0f, l = 1.0f;
  if (ise) {
    th = a.theta_in[ge];
    l = rest_length(a, ge, arow);
  }
  float am = 0.0f, om = 0.0f, ph = 0.0f, rs = 0.0f, kb = 0.0f;
  if (isb) {
    if (a.has_act) {
      am = a.act.amp[gb];
      om = a.act.omega[gb];
      ph = a.act.phase[gb];
    } else {
      rs = rest_angle(a, gb, arow);
    }
    kb = a.kb[gb];
  }
  const double tact =
      a.has_act ? (double)(*a.act.t + (a.act.t_inst ? a.act.t_inst[m.inst] : 0LL)) : 0.0;
  float2* pw = reinterpret_cast<float2*>(smem);
  if (stage_pw) {
#pragma unroll
    for (int h = 0; h < 2; ++h)
      if (lane + 64 * h < nn + 2 * nsk) pw[lane + 64 * h] = pwv[h];
    __syncthreads();
  }
  // prologue: damping, prediction, steric push, Cosserat target (simulate.py:41-72)
  float X = 0.0f, Y = 0.0f, TX = 0.0f, TY = 0.0f;
  if (isn) {
    const float dvx = v.x * c.damping, dvy = v.y * c.damping;
    X = q.x + dvx * c.dt;
    Y = q.y + dvy * c.dt;
    if (c.enable_steric) {
      float fx, fy;
      steric_force(c, gi, a.steric, a.pred_w, g, fx, fy, stage_pw ? pw : nullptr, pw0);
      X = X + (fx * c.steric_strength) * c.dt;
      Y = Y + (fy * c.steric_strength) * c.dt;
    }
    TX = c.cos_target ? X : 0.0f;
    TY = c.cos_target ? Y : 0.0f;
    a.debug_out[2 * (size_t)g] = TX;
    a.debug_out[2 * (size_t)g + 1] = TY;
  }
  if (isb && a.has_act) rs = (float)((double)am * sin((double)om * tact + (double)ph));
  // per-edge constants: degrees of the edge and of the node, the Cosserat
  // stiffness terms (stable_cosserat.py:186-195)
  const float deg_e = (lane < nb && lane >= 1) ? 0.5f : 1.0f;
  const float deg_n = (lane < ne && lane >= 1) ? 0.5f : 1.0f;
  const float kl = c.k_st / l, he = ise ? c.k_st / (l * l) : 0.0f;
  const float Wd = gi.dom_wf, Hd = gi.dom_hf;
  for (int cyc = 0; cyc < c.cycles; ++cyc) {
    // theta pass: hinge corrections (pbd.py:32-40); the cycle's wraps on
    // their common ranges, one check for the wavefront, exact otherwise
    const float th_next = from_next(th);
    const float xn = from_next(X), yn = from_next(Y);
    bool slow_b = false, slow_e = false;
    float diff = wrap_angle_fast(th_next - th, slow_b);
    float rx = min_image_fast(xn, X, Wd, slow_e);
    float ry = min_image_fast(yn, Y, Hd, slow_e);
    if ((slow_b && isb) || (slow_e && ise)) {
      diff = wrap_angle(th_next - th);
      rx = min_image(xn, X, Wd);
      ry = min_image(yn, Y, Hd);
    }
    float corr = 0.0f;
    if (isb) corr = ((diff - rs) * kb) * 0.5f;
    const float corr_prev = from_prev(corr);
    float fex = 0.0f, fey = 0.0f;
    if (ise) {
      float d = 0.0f;
      if (lane < nb) d = d + corr;
      if (lane >= 1) d = d + (-corr_prev);
      if (!is_cos) {
        th = th + (d * deg_e) * 1.8f;
      } else {
        th = th + d * deg_e;
        const float alpha = datan2(ry, rx);
        const float pull = wrap_angle(alpha - th);
        th = th + pull * 0.05f;
      }
      float dx, dy;
      fsincos(th, &dy, &dx);
      if (!is_cos) {
        const float tx = dx * l, ty = dy * l;
        const float len = sqrtf(rx * rx + ry * ry);
        const float md = fmax_nan(len, 1e-8f);
        const float cdx = rx / md, cdy = ry / md;
        const float stx = cdx * (len - l), sty = cdy * (len - l);
        const float shx = (rx - tx) - stx, shy = (ry - ty) - sty;
        fex = ((stx + shx * c.k_sh) * c.k_st) * 0.5f;
        fey = ((sty + shy * c.k_sh) * c.k_st) * 0.5f;
      } else {
        fex = kl * (rx / l - dx);
        fey = kl * (ry / l - dy);
      }
    }
    // per node: +F(i) as src, then -F(i - 1) as tgt (pbd.py:84-89 /
    // stable_cosserat.py:197-207)
    const float fpx = from_prev(fex), fpy = from_prev(fey);
    const float hp = from_prev(he);
    if (isn) {
      if (!is_cos) {
        float ux = 0.0f, uy = 0.0f;
        if (lane < ne) {
          ux = ux + fex;
          uy = uy + fey;
        }
        if (lane >= 1) {
          ux = ux + (-fpx);
          uy = uy + (-fpy);
        }
        X = X + ux * deg_n;
        Y = Y + uy * deg_n;
      } else {
        const float w = c.cos_w;
        float fx = w * (TX - X), fy = w * (TY - Y);
        float h = w;
        if (lane < ne) {
          fx = fx + fex;
          fy = fy + fey;
        }
        if (lane >= 1) {
          fx = fx + (-fpx);
          fy = fy + (-fpy);
        }
        if (lane < ne) h = h + he;
        if (lane >= 1) h = h + hp;
        X = X + fx / h;
        Y = Y + fy / h;
      }
    }
  }
  // epilogue: v = (x - x_prev) / dt (simulate.py:83-84); unwrapped x feeds the
  // immersed boundary, or is wrapped here when the fluid is off
  if (isn) {
    const float2 vo = make_float2((X - q.x) / c.dt, (Y - q.y) / c.dt);
    *reinterpret_cast<float2*>(a.vel_out + 2 * (size_t)g) = vo;
    const float2 po = c.enable_fluid ? make_float2(X, Y)
                                     : make_float2(fmodf_floor(X, Wd), fmodf_floor(Y, Hd));
    *reinterpret_cast<float2*>(a.pos_out + 2 * (size_t)g) = po;
  }
  if (ise) a.theta_out[ge] = th;
}

// The chain path runs the whole node phase of a step (mode 0).  CHAINS: every
// component is a chain, so the general path (and its registers) is left out.
template <bool CHAINS>
__device__ __forceinline__ void rod_dispatch(const RodArgs& a, int comp, float* smem) {
  const CompMeta m = a.L.meta[comp];
  if (CHAINS || (m.chain && a.mode == 0 && a.use_lds)) rod_chain(a, m, smem);
  else rod_component(a, comp, smem);
}

template <bool CHAINS>
__global__ void __launch_bounds__(256) k_rod(RodArgs a) {
  extern __shared__ float smem[];
  rod_dispatch<CHAINS>(a, blockIdx.x, smem);
}

// Rod components in workgroups [0, rod_blocks), then an LBM pass in 64-cell
// workgroups (a quarter of an LBM_THREADS block each, one wavefront; more
// cells per thread measured slower).  rod_blocks is a multiple of 8:
// workgroup i runs on XCD i % 8, and the LBM part keeps the standalone
// kernel's XCD-banded order.
template <bool CHAINS>
__global__ void __launch_bounds__(256) k_rod_lbm(RodArgs a, LbmArgs l, int rod_blocks) {
  extern __shared__ float smem[];
  const int bid = blockIdx.x;
  if (bid < rod_blocks) {
    if (bid < a.L.num_comp) rod_dispatch<CHAINS>(a, bid, smem);
    return;
  }
  int v = bid - rod_blocks;
  const int G = gridDim.x - rod_blocks;
  if ((G & 7) == 0) v = (v & 7) * (G >> 3) + (v >> 3);
  lbm_cell(l, v >> 2, (v & 3) * 64 + threadIdx.x);
}

}  // namespace

size_t rod_scratch_floats(int nn, int ne, int nb) {
  return (size_t)4 * nn + (size_t)5 * ne + (size_t)3 * nb;
}

size_t rod_lds_bytes(int nn, int ne, int nb, int nni, int nei) {
  const size_t ints = (size_t)2 * nb + (size_t)3 * ne + 1 + nei + (size_t)nn + 1 + nni;
  return ((size_t)nn + 2 * ROD_WIN_MAX) * sizeof(float2) +
         rod_scratch_floats(nn, ne, nb) * sizeof(float) + ints * sizeof(int);
}

__global__ void k_steric_nodes(Consts c, BatchDev bd, int N, const float* __restrict__ steric,
                               const float* __restrict__ pred_w, float* __restrict__ force_out) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  float fx, fy;
  steric_force(c, bd.geo[bd.node_inst[n]], steric, pred_w, n, fx, fy);
  force_out[2 * n] = fx;
  force_out[2 * n + 1] = fy;
}

__global__ void k_tick(long long* t) { *t = *t + 1; }
__global__ void k_set_tick(long long* t, long long v) { *t = v; }

hipError_t launch_steric_nodes(const Consts& c, const BatchDev& b, int N, const float* steric,
                               const float* pred_w, float* force_out, hipStream_t s) {
  if (N == 0) return hipSuccess;
  hipLaunchKernelGGL(k_steric_nodes, dim3((N + 255) / 256), dim3(256), 0, s, c, b, N, steric,
                     pred_w, force_out);
  return hipGetLastError();
}

hipError_t launch_tick(long long* t, hipStream_t s) {
  hipLaunchKernelGGL(k_tick, dim3(1), dim3(1), 0, s, t);
  return hipGetLastError();
}

hipError_t launch_set_tick(long long* t, long long value, hipStream_t s) {
  hipLaunchKernelGGL(k_set_tick, dim3(1), dim3(1), 0, s, t, value);
  return hipGetLastError();
}

hipError_t rod_prepare(int lds_max) {
  // components beyond 64 KB of LDS (about 2000 nodes) need the raised limit
  const int lds = lds_max - 1024;
  const void* fns[] = {reinterpret_cast<const void*>(&k_rod<false>),
                       reinterpret_cast<const void*>(&k_rod<true>),
                       reinterpret_cast<const void*>(&k_rod_lbm<false>),
                       reinterpret_cast<const void*>(&k_rod_lbm<true>)};
  for (const void* f : fns) {
    const hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_rod(const Consts& c, const BatchDev& b, const RodLayoutDev& L, int N,
                      size_t lds_bytes, int use_lds, float* gscratch, int mode, const float* pos_in, const float* vel_in,
                      const float* y_in, const float* theta_in, const float* rest_len,
                      const float* rest_ang, const float* kb, const Actuation* act,
                      const float* steric, const float* pred_w, float* pos_out, float* vel_out,
                      float* debug_out, float* theta_out, const LbmArgs* lbm, hipStream_t s,
                      bool all_chains, hipEvent_t ev0, hipEvent_t ev1) {
  if (L.num_comp == 0 && !lbm) return hipSuccess;
  RodArgs a;
  a.c = c;
  a.bd = b;
  a.L = L;
  a.N = N;
  a.mode = mode;
  a.use_lds = use_lds;
  a.gscratch = gscratch;
  a.pos_in = pos_in;
  a.vel_in = vel_in;
  a.y_in = y_in;
  a.theta_in = theta_in;
  a.rest_len = rest_len;
  a.rest_ang = rest_ang;
  a.kb = kb;
  a.has_act = act != nullptr && act->kind == 0;
  a.has_rows = act != nullptr && act->kind == 1;
  if (act) a.act = *act; else a.act = Actuation{};
  a.steric = steric;
  a.pred_w = pred_w;
  a.pos_out = pos_out;
  a.vel_out = vel_out;
  a.debug_out = debug_out;
  a.theta_out = theta_out;
  const size_t lds = use_lds ? lds_bytes : 0;
  const bool chains = all_chains && mode == 0 && use_lds;

  if (lbm) {
    const int rb = (L.num_comp + 7) & ~7;
    // events carried by the dispatch itself (measurement) bracket its execution
    if (chains)
      hipExtLaunchKernelGGL(k_rod_lbm<true>, dim3(rb + 4 * b.lbm_blocks), dim3(64), lds, s, ev0,
                            ev1, 0, a, *lbm, rb);
    else
      hipExtLaunchKernelGGL(k_rod_lbm<false>, dim3(rb + 4 * b.lbm_blocks), dim3(64), lds, s, ev0,
                            ev1, 0, a, *lbm, rb);
  } else {
    if (chains)
      hipLaunchKernelGGL(k_rod<true>, dim3(L.num_comp), dim3(64), lds, s, a);
    else
      hipLaunchKernelGGL(k_rod<false>, dim3(L.num_comp), dim3(64), lds, s, a);
  }
  return hipGetLastError();
}

}  // namespace mc
