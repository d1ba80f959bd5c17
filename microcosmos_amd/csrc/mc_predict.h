// The inertial prediction that opens a coupled step (simulate.py:41-50 with
// the CIC set-up of forces.py:95-110), one wavefront per filament.  Run by
// its own launch (k_predict) or, inside a rollout, by the immersed-boundary
// launch that ends the previous step, right after the node velocities and
// positions are final.
#pragma once
#include "mc_common.h"
#include "mc_internal.h"

namespace mc {

// Wrapped inertial prediction of node n (simulate.py:46-50, forces.py:95);
// without vel, the positions as given.
__device__ __forceinline__ void predict_xy(const Consts& c, const InstGeo& gi,
                                           const float* __restrict__ pos,
                                           const float* __restrict__ vel, int n, float& x,
                                           float& y) {
  float px = pos[2 * n], py = pos[2 * n + 1];
  if (vel) {
    const float vx = vel[2 * n] * c.damping, vy = vel[2 * n + 1] * c.damping;
    px = px + vx * c.dt;
    py = py + vy * c.dt;
  }
  x = fmodf_floor(px, gi.dom_wf);
  y = fmodf_floor(py, gi.dom_hf);
}

// One wavefront (lane = threadIdx.x & 63) predicts filament component
// `comp`: each node's wrapped prediction, CIC cell and weights
// (forces.py:95-110), and the band of instance rows the component's CIC
// entries touch, rows start .. start + len - 1 (mod H), with the component's
// node count and ids.  A node's row y0 is the first node's row plus a
// minimum-image offset d in [-H/2, H/2); the band runs from min d to max d + 1,
// so it holds every y0 and y1 = y0 + 1 of the component whatever its shape (a
// loose band only costs the row pass extra checks); the same for its columns
// when comp_cols is given.
__device__ __forceinline__ void predict_component(const PredictArgs& a, int comp, int lane) {
  const Consts& c = a.c;
  const CompMeta m = a.L.meta[comp];
  const InstGeo& gi = a.bd.geo[m.inst];
  const int H = gi.H, W = gi.W;
  auto node_id = [&](int i) { return m.nf >= 0 ? m.nf + i : a.L.comp_nodes[m.n0 + i]; };
  int yref = 0, xref = 0;
  if (c.enable_steric) {
    float x, y;
    predict_xy(c, gi, a.pos, a.vel, node_id(0), x, y);
    yref = imod(f2i(floorf(y)), H);
    xref = imod(f2i(floorf(x)), W);
  }
  int lo = 0x7fffffff, hi = -0x7fffffff, clo = 0x7fffffff, chi = -0x7fffffff;
  for (int i = lane; i < m.nn; i += 64) {
    const int n = node_id(i);
    float x, y;
    predict_xy(c, gi, a.pos, a.vel, n, x, y);
    a.pred_w[2 * n] = x;
    a.pred_w[2 * n + 1] = y;
    if (!c.enable_steric) continue;
    const float fx = floorf(x), fy = floorf(y);
    const int x0 = imod(f2i(fx), W), y0 = imod(f2i(fy), H);
    const int x1 = imod(x0 + 1, W), y1 = imod(y0 + 1, H);
    const float dx = x - fx, dy = y - fy;
    const float sv = c.steric_sv;
    // columns within the instance, rows as global steric rows
    *reinterpret_cast<int4*>(a.cic_cell + 4 * (size_t)n) =
        make_int4(x0, x1, gi.row0 + y0, gi.row0 + y1);
    *reinterpret_cast<float4*>(a.cic_w + 4 * (size_t)n) =
        make_float4(((1.0f - dx) * (1.0f - dy)) * sv, (dx * (1.0f - dy)) * sv,
                    ((1.0f - dx) * dy) * sv, (dx * dy) * sv);
    const int d = imod(y0 - yref + H / 2, H) - H / 2;
    lo = min(lo, d);
    hi = max(hi, d);
    const int dc = imod(x0 - xref + W / 2, W) - W / 2;
    clo = min(clo, dc);
    chi = max(chi, dc);
  }
  if (!c.enable_steric) return;
  for (int o = 32; o > 0; o >>= 1) {
    lo = min(lo, __shfl_xor(lo, o, 64));
    hi = max(hi, __shfl_xor(hi, o, 64));
    clo = min(clo, __shfl_xor(clo, o, 64));
    chi = max(chi, __shfl_xor(chi, o, 64));
  }
  if (lane == 0) {
    const int len = hi + 2 - lo;  // rows lo .. hi + 1 relative to yref
    // with the component's node count and first id (contiguous ids) or the
    // complement of its node-list offset, for the row pass
    const int first = m.nf >= 0 ? m.nf : ~m.n0;
    a.comp_rows[comp] = len >= H ? make_int4(0, H, m.nn, first)
                                 : make_int4(imod(yref + lo, H), len, m.nn, first);
    // the columns its CIC entries touch, the same way (x0 and x0 + 1)
    const int clen = chi + 2 - clo;
    if (a.comp_cols)
      a.comp_cols[comp] = clen >= W ? make_int2(0, W) : make_int2(imod(xref + clo, W), clen);
  }
}

}  // namespace mc
