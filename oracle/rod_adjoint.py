"""Reverse-mode derivative of the rod-only coupled step (CPU restatement).

TEST INFRASTRUCTURE ONLY, like ``step_ref``: imported by ``tests/`` as the
checker of the device adjoint, never by the product path.

Scope: ``step`` (simulate.py:37-100) with ``enable_fluid = False``, the PBD
constraint (pbd.py:14-91) or the stable Cosserat constraint
(stable_cosserat.py:13-119, see ``_cosserat_cycles_vjp``), with or without
the steric term (forces.py:20-118, see ``steric_vjp``): the configurations of
the reference's differentiable folding runs
(experiments/filament_folding/experiment.py:52-53, 925-941, 985-1018 take
``jax.value_and_grad`` through ``lax.scan`` of this step; its default physics
is stable Cosserat with the steric term).  The derivative is written out by
hand from the forward equations of ``step_ref``:

forward, one step (x, v, theta; rest angles rho):
    p0 = x + (v * d) * dt                                    simulate.py:46-50
    cycle c:  theta pass (pbd.py:23-49)
                  diff_b = wrap(theta[e_out] - theta[e_in])
                  corr_b = (diff_b - rho_b) * kb_b * 0.5
                  theta' = theta + omega / deg_e * (sum_in corr - sum_out corr)
              position pass (pbd.py:55-89), per edge (s, g):
                  r = min_image(p[g] - p[s]),  L = |r|,  tv = l0 (cos th', sin th')
                  F = k_st/2 * ((1 - k_sh) * r (L - l0) / L + k_sh * (r - tv))
                  p'[n] = p[n] + (sum_{s=n} F - sum_{g=n} F) / deg_n
    x_new = wrap(p_C),  v_new = (p_C - x) / dt                 simulate.py:83-98

reverse (adjoints xb, vb, thb of x_new, v_new, theta_new):
    pb = xb + vb / dt;  xb_direct = -vb / dt
    cycle c (last to first):
      Fb_e = pb[s] / deg_s - pb[g] / deg_g
      rb_e = J_e Fb_e,  J = k_st/2 ((1 - k_sh)(I - l0/L (I - rhat rhat^T)) + k_sh I)
      pb[g] += rb_e;  pb[s] -= rb_e
      thb'_e += -k_st/2 k_sh l0 (-sin th' Fb_x + cos th' Fb_y)
      corrb_b = omega (thb'[e_in] / deg_in - thb'[e_out] / deg_out)
      rhob_b -= corrb_b kb_b / 2;  diffb_b = corrb_b kb_b / 2
      thb[e_out] += diffb_b;  thb[e_in] -= diffb_b
    steric push p1 = p0 + f(p0) strength dt:  pb0 = pb1 + (df/dp0)^T pb1 strength dt
    xb_prev = pb0 + xb_direct;  vb_prev = pb0 * d * dt

The minimum image, the angle wrap and the final periodic wrap have derivative
1 almost everywhere; their jumps are not differentiated (as in the reference's
``jax.grad``, whose ``%`` derivative is 1).

Pinned by central finite differences of the float64 forward (``step_ref.step``)
in ``tests/test_rod_adjoint_oracle.py``: per step (directional derivatives
for x, v and theta, both solvers, steric on and off), the steric force's VJP
alone, and whole rollouts (rest-angle and rest-length gradients), including
the folding experiment's default physics and its objective over frames.
"""
from __future__ import annotations

from dataclasses import replace

import numpy as np

from . import losses as LS
from . import step_ref as R

OMEGA = 1.8  # pbd.py:43


def _degrees(s: R.State):
    src, tgt = s.pairs[:, 0], s.pairs[:, 1]
    n = s.position.shape[0]
    deg_n = np.zeros(n)
    np.add.at(deg_n, src, 1.0)
    np.add.at(deg_n, tgt, 1.0)
    deg_n = np.maximum(deg_n, 1.0)
    deg_e = np.asarray(R._edge_bending_degree(s.theta.shape[0], s.bending_pairs, np.float64))
    return deg_n, deg_e


def _check(p: R.Params):
    if p.enable_fluid or p.solver_type not in ("pbd", "cosserat"):
        raise NotImplementedError("the rod adjoint covers the rod-only step (PBD or stable "
                                  "Cosserat, with or without the steric term) without fluid")


def _sample_w(V, yy, xx):
    """Corners, weights and the x / y derivatives of the bilinear sample
    sample_bilinear_wrap(V, yy, xx) (forces.py:20-38 via map_coordinates)."""
    h, w = V.shape
    ly, lx = np.floor(yy), np.floor(xx)
    wy1, wx1 = yy - ly, xx - lx
    wy0, wx0 = 1.0 - wy1, 1.0 - wx1
    iy, ix = ly.astype(np.int64), lx.astype(np.int64)
    y0, y1 = np.mod(iy, h), np.mod(iy + 1, h)
    x0, x1 = np.mod(ix, w), np.mod(ix + 1, w)
    corners = [(y0, x0, wy0 * wx0), (y0, x1, wy0 * wx1), (y1, x0, wy1 * wx0), (y1, x1, wy1 * wx1)]
    dsx = wy0 * (V[y0, x1] - V[y0, x0]) + wy1 * (V[y1, x1] - V[y1, x0])
    dsy = wx0 * (V[y1, x0] - V[y0, x0]) + wx1 * (V[y1, x1] - V[y0, x1])
    return corners, dsx, dsy


def steric_vjp(pos0, grid_shape, p: R.Params, gbar):
    """(d f / d pos0)^T gbar for f = compute_field_steric_force_corrected(pos0)
    (forces.py:41-118): f = -grad_FD V(pos) - c sum_j K(r_ij) r_ij with
    V = G * D(pos), D the CIC deposit, G the Gaussian filter (a symmetric
    operator, so its adjoint is itself), c = sv / sigma^2 and
    d(K r)/dr = K (I - r r^T / sigma^2)."""
    sigma, nsk, sv = float(p.steric_sigma), int(p.steric_neighbor_skip), float(p.steric_scatter_value)
    pos = np.asarray(R.periodic_boundary(grid_shape, np.asarray(pos0, np.float64)), np.float64)
    g = np.asarray(gbar, np.float64)
    n = pos.shape[0]
    _, V = R.update_steric_potential(pos, grid_shape, sigma, sv)
    V = np.asarray(V, np.float64)
    out = np.zeros((n, 2))
    # window term: out_m += c sum_j J(r_mj) (g_j - g_m)
    c = sv / (sigma * sigma)
    for k in range(-nsk, nsk + 1):
        j = np.arange(n) + k
        valid = (j >= 0) & (j < n)
        js = np.clip(j, 0, n - 1)
        r = np.asarray(R.displacement(grid_shape, pos, pos[js]), np.float64)
        d2 = (r * r).sum(axis=1)
        K = np.exp(-d2 / (2 * sigma * sigma)) / (2 * np.pi * sigma * sigma)
        dg = g[js] - g
        rd = (r * dg).sum(axis=1)
        out += (c * valid * K)[:, None] * (dg - r * (rd / (sigma * sigma))[:, None])
    # field term f = -(gx, gy), gx = S(y, x+e) - S(y, x-e), gy = S(y+e, x) - S(y-e, x)
    e = 0.5
    x, y = pos[:, 0], pos[:, 1]
    samples = [(y, x + e, 0, 1.0), (y, x - e, 0, -1.0), (y + e, x, 1, 1.0), (y - e, x, 1, -1.0)]
    Vbar = np.zeros_like(V)
    jac = np.zeros((n, 2, 2))  # d(gx, gy)/d(x, y)
    for yy, xx, comp, sgn in samples:
        corners, dsx, dsy = _sample_w(V, yy, xx)
        jac[:, comp, 0] += sgn * dsx / (2 * e)
        jac[:, comp, 1] += sgn * dsy / (2 * e)
        for cy, cx, wgt in corners:
            np.add.at(Vbar, (cy, cx), -g[:, comp] * sgn * wgt / (2 * e))
    out[:, 0] += -(jac[:, 0, 0] * g[:, 0] + jac[:, 1, 0] * g[:, 1])
    out[:, 1] += -(jac[:, 0, 1] * g[:, 0] + jac[:, 1, 1] * g[:, 1])
    Dbar = np.asarray(R.gaussian_filter(Vbar, sigma), np.float64)
    h, w = grid_shape
    fx, fy = np.floor(x), np.floor(y)
    dx, dy = x - fx, y - fy
    x0 = np.mod(fx.astype(np.int64), w)
    y0 = np.mod(fy.astype(np.int64), h)
    x1, y1 = np.mod(x0 + 1, w), np.mod(y0 + 1, h)
    d00, d10, d01, d11 = Dbar[y0, x0], Dbar[y0, x1], Dbar[y1, x0], Dbar[y1, x1]
    out[:, 0] += sv * (-(1 - dy) * d00 + (1 - dy) * d10 - dy * d01 + dy * d11)
    out[:, 1] += sv * (-(1 - dx) * d00 - dx * d10 + (1 - dx) * d01 + dx * d11)
    return out


def step_tape(s: R.State, dt: float, p: R.Params):
    """Forward step that also returns the per-cycle positions (before each
    position pass) and angles (after each angle pass)."""
    _check(p)
    s = s.copy()
    prev = s.position
    s.velocity = s.velocity * p.damping
    s.position = prev + s.velocity * dt
    if p.enable_steric:
        _, s.steric, fst = R.compute_field_steric_force_corrected(
            s.position, s.grid_shape, p.steric_sigma, p.steric_neighbor_skip,
            p.steric_scatter_value)
        s.position = s.position + fst * p.steric_strength * dt
    cos = p.solver_type == "cosserat"
    if cos:
        s.debug_vector = s.position
    ps, ths = [], []
    for _ in range(p.cycles_per_step):
        ps.append(s.position.copy())
        if cos:
            R.stable_cosserat_constraint(s, p)
            ths.append(s.theta.copy())
        else:
            t = s.copy()
            deg_e = R._edge_bending_degree(t.theta.shape[0], t.bending_pairs, t.theta.dtype)
            delta = R._theta_bending_delta(t.theta, t.bending_pairs, t.bending_rest_angles,
                                           t.bending_stiffness)
            ths.append(t.theta + delta / deg_e * OMEGA)
            R.pbd_rod_constraint(s, p)
    s.velocity = (s.position - prev) / dt
    s.position = R.periodic_boundary(s.grid_shape, s.position)
    return s, ps, ths


def _cosserat_cycles_vjp(s, p, ps, ths, pb, thb, rhob, deg_e):
    """Reverse of the stable Cosserat cycles (stable_cosserat.py:13-119):
      th1 = theta + delta / deg_e;  th2 = th1 + 0.05 wrap(atan2(r) - th1)
      F_e = k/l (r / l - (cos th2, sin th2));  h_n = w + sum k / l^2
      x' = x + (w (y - x) + sum_src F - sum_tgt F) / h
    The rest length l enters F (dF/dl = -2 k r / l^3 + k / l^2 (cos, sin))
    and h (dh_n/dl = -2 k / l^3 at both ends, dx'/dh = -(x' - x) / h).
    Returns (pb, thb, yb, rhob, lb)."""
    src, tgt = s.pairs[:, 0], s.pairs[:, 1]
    bin_, bout = s.bending_pairs[:, 0], s.bending_pairs[:, 1]
    l = np.asarray(s.rest_lengths, np.float64)
    kb = np.asarray(s.bending_stiffness, np.float64)
    k = float(p.stiffness_stretch)
    w = 1.0 / (float(p.config_dt) ** 2)
    h = np.full(s.position.shape[0], w)
    np.add.at(h, src, k / l ** 2)
    np.add.at(h, tgt, k / l ** 2)
    yb = np.zeros_like(pb)
    lb = np.zeros_like(l)
    y = np.asarray(ps[0], np.float64)
    for c in reversed(range(p.cycles_per_step)):
        pos = np.asarray(ps[c], np.float64)
        th2 = np.asarray(ths[c], np.float64)
        Fb = pb[src] / h[src, None] - pb[tgt] / h[tgt, None]
        # the cycle's update (x' - x) = num / h, from its input positions
        rr = np.asarray(R.displacement(s.grid_shape, pos[tgt], pos[src]), np.float64)
        d2 = np.stack([np.cos(th2), np.sin(th2)], axis=1)
        F = (k / l)[:, None] * (rr / l[:, None] - d2)
        num = w * (y - pos)
        np.add.at(num, src, F)
        np.add.at(num, tgt, -F)
        hb = -(pb * num).sum(axis=1) / h ** 2
        lb = lb + (-2.0 * k / l ** 3) * ((rr * Fb).sum(axis=1) + hb[src] + hb[tgt]) \
            + (k / l ** 2) * (d2 * Fb).sum(axis=1)
        yb = yb + pb * (w / h)[:, None]
        pnew = pb * (1.0 - w / h)[:, None]
        rb = (k / l ** 2)[:, None] * Fb
        th2b = thb + (-(k / l)) * (-np.sin(th2) * Fb[:, 0] + np.cos(th2) * Fb[:, 1])
        th1b = 0.95 * th2b
        ab = 0.05 * th2b
        r = np.asarray(R.displacement(s.grid_shape, pos[tgt], pos[src]), np.float64)
        L2 = (r * r).sum(axis=1)
        rb = rb + (ab / L2)[:, None] * np.stack([-r[:, 1], r[:, 0]], axis=1)
        np.add.at(pnew, tgt, rb)
        np.add.at(pnew, src, -rb)
        corrb = th1b[bin_] / deg_e[bin_] - th1b[bout] / deg_e[bout]
        diffb = corrb * kb * 0.5
        rhob = rhob - diffb
        thb = th1b.copy()
        np.add.at(thb, bout, diffb)
        np.add.at(thb, bin_, -diffb)
        pb = pnew
    return pb, thb, yb, rhob, lb


def _prologue_vjp(s, dt, p: R.Params, pb):
    """Adjoint of the steric push (simulate.py:52-63) on the prediction
    pred0 = x + (v d) dt:  pred1 = pred0 + f(pred0) strength dt."""
    if not p.enable_steric:
        return pb
    pred0 = s.position + (s.velocity * p.damping) * dt
    return pb + steric_vjp(pred0, s.grid_shape, p, pb * float(p.steric_strength) * dt)


def step_vjp(s: R.State, dt: float, p: R.Params, xb, vb, thb):
    """Adjoint of one step at state ``s``: returns (xb_prev, vb_prev,
    thb_prev, rhob) given the adjoints of the step's outputs."""
    return step_vjp_params(s, dt, p, xb, vb, thb)[:4]


def step_vjp_params(s: R.State, dt: float, p: R.Params, xb, vb, thb):
    """``step_vjp`` plus the adjoint of the rest lengths: returns (xb_prev,
    vb_prev, thb_prev, rhob, lb).  PBD: dF/dl0 = -k_st/2 ((1 - k_sh) rhat +
    k_sh (cos th', sin th'))."""
    _, ps, ths = step_tape(s, dt, p)
    deg_n, deg_e = _degrees(s)
    src, tgt = s.pairs[:, 0], s.pairs[:, 1]
    bin_, bout = s.bending_pairs[:, 0], s.bending_pairs[:, 1]
    l0 = np.asarray(s.rest_lengths, np.float64)
    kb = np.asarray(s.bending_stiffness, np.float64)
    kst, ksh = float(p.stiffness_stretch), float(p.stiffness_shear)
    xb = np.asarray(xb, np.float64)
    vb = np.asarray(vb, np.float64)
    pb = xb + vb / dt
    xb_direct = -vb / dt
    thb = np.array(thb, np.float64, copy=True)
    rhob = np.zeros(s.bending_pairs.shape[0])
    if p.solver_type == "cosserat":
        pb, thb, yb, rhob, lb = _cosserat_cycles_vjp(s, p, ps, ths, pb, thb, rhob, deg_e)
        pb = pb + yb  # y is the prologue's prediction (simulate.py:71-72)
        pb = _prologue_vjp(s, dt, p, pb)
        return pb + xb_direct, pb * float(p.damping) * dt, thb, rhob, lb
    lb = np.zeros_like(l0)
    for c in reversed(range(p.cycles_per_step)):
        pos = np.asarray(ps[c], np.float64)
        th = np.asarray(ths[c], np.float64)
        # position pass
        Fb = pb[src] / deg_n[src, None] - pb[tgt] / deg_n[tgt, None]
        r = np.asarray(R.displacement(s.grid_shape, pos[tgt], pos[src]), np.float64)
        L = np.sqrt((r * r).sum(axis=1))
        Ls = np.maximum(L, 1e-8)
        rh = r / Ls[:, None]
        a = (1.0 - l0 / Ls)[:, None]
        # J Fb with J = k_st/2 ((1 - k_sh)(a I + l0/L rh rh^T) + k_sh I)
        proj = (rh * Fb).sum(axis=1)[:, None] * rh
        rb = 0.5 * kst * ((1.0 - ksh) * (a * Fb + (l0 / Ls)[:, None] * proj) + ksh * Fb)
        np.add.at(pb, tgt, rb)
        np.add.at(pb, src, -rb)
        lb = lb - 0.5 * kst * ((1.0 - ksh) * (rh * Fb).sum(axis=1)
                               + ksh * (np.cos(th) * Fb[:, 0] + np.sin(th) * Fb[:, 1]))
        thb = thb + (-0.5 * kst * ksh * l0) * (-np.sin(th) * Fb[:, 0] + np.cos(th) * Fb[:, 1])
        # angle pass
        corrb = OMEGA * (thb[bin_] / deg_e[bin_] - thb[bout] / deg_e[bout])
        diffb = corrb * kb * 0.5
        rhob = rhob - diffb
        np.add.at(thb, bout, diffb)
        np.add.at(thb, bin_, -diffb)
    pb = _prologue_vjp(s, dt, p, pb)
    return pb + xb_direct, pb * float(p.damping) * dt, thb, rhob, lb


def msd_loss(position, target, grid_shape):
    """Folding loss: mean over nodes of the squared minimum-image distance to
    the target curve (experiments/filament_folding/experiment.py:57-165 uses
    point-set distances of this kind); returns (loss, d loss / d position)."""
    d = np.asarray(R.displacement(grid_shape, np.asarray(position, np.float64),
                                  np.asarray(target, np.float64)), np.float64)
    n = d.shape[0]
    return float((d * d).sum() / n), 2.0 * d / n


def rollout_grad(s0: R.State, dt: float, p: R.Params, num_steps: int, target,
                 dtype=np.float64, loss_fn=None):
    """Loss after ``num_steps`` steps and its gradient with respect to the
    (time-constant) bending rest angles, by storing every step's state.
    ``dtype`` is the precision of the forward trajectory (and of each step's
    recomputed tape); the adjoint arithmetic is float64 either way.  A float32
    trajectory is the device's own (bit-identical steps), which separates the
    adjoint's accuracy from the divergence of float32 and float64 rollouts.
    ``loss_fn(position) -> (loss, d loss / d position)`` replaces the mean
    squared distance to ``target`` (e.g. the folding losses of
    ``oracle.losses``)."""
    states = [s0.astype(dtype)]
    for _ in range(num_steps):
        states.append(R.step(states[-1], dt, p))
    if loss_fn is None:
        loss, xb = msd_loss(states[-1].position, target, s0.grid_shape)
    else:
        loss, xb = loss_fn(np.asarray(states[-1].position, np.float64))
    vb = np.zeros_like(xb)
    thb = np.zeros(s0.theta.shape[0])
    rhob = np.zeros(s0.bending_pairs.shape[0])
    for t in reversed(range(num_steps)):
        xb, vb, thb, rb = step_vjp(states[t], dt, p, xb, vb, thb)
        rhob = rhob + rb
    return loss, rhob


def rollout_frames_grad(s0: R.State, dt: float, p: R.Params, num_steps: int, frame_losses,
                        dtype=np.float64):
    """Multi-frame objective: ``frame_losses`` maps a step count t (1 ..
    num_steps) to ``loss_fn(position after t steps) -> (loss, d loss / d
    position)``; returns (sum of the frame losses, its gradient with respect
    to the rest angles).  The reverse pass adds each frame's cotangent to the
    position adjoint when it reaches that frame."""
    def objective(pos):
        total, bars = 0.0, {}
        for t, fn in frame_losses.items():
            l, g = fn(pos[t])
            total += l
            bars[t] = g
        return total, bars
    total, rhob, _ = rollout_objective_grad(s0, dt, p, num_steps, list(frame_losses), objective,
                                            dtype)
    return total, rhob


def rollout_objective_grad(s0: R.State, dt: float, p: R.Params, num_steps: int, steps,
                           objective, dtype=np.float64, lengths=False):
    """An objective coupling several frames: ``objective({t: position after
    t steps for t in steps})`` returns (loss, {t: d loss / d that position}),
    as the folding loss's stability term compares frames with each other.
    Returns (loss, its gradient with respect to the rest angles, the frame
    positions); with ``lengths``, (loss, rest-angle gradient, rest-length
    gradient, frame positions)."""
    states = [s0.astype(dtype)]
    for _ in range(num_steps):
        states.append(R.step(states[-1], dt, p))
    pos = {int(t): np.asarray(states[int(t)].position, np.float64) for t in steps}
    total, bars = objective(pos)
    xb = np.zeros((s0.position.shape[0], 2))
    vb = np.zeros_like(xb)
    thb = np.zeros(s0.theta.shape[0])
    rhob = np.zeros(s0.bending_pairs.shape[0])
    lenb = np.zeros(s0.pairs.shape[0])
    for t in reversed(range(num_steps)):
        if t + 1 in bars:
            xb = xb + bars[t + 1]
        xb, vb, thb, rb, lb = step_vjp_params(states[t], dt, p, xb, vb, thb)
        rhob = rhob + rb
        lenb = lenb + lb
    if lengths:
        return total, rhob, lenb, pos
    return total, rhob, pos


def folding_raw_grad(s0: R.State, dt: float, p: R.Params, num_steps: int, raw,
                     curvature_scale: float, target, eval_steps, stab_steps, anchor_step,
                     smoothness_weight: float = 0.0, rest=None, dtype=np.float64,
                     raw_lengths=None, distance_range=(0.1, 4.0), lengths=None, **objective_kw):
    """The folding experiment's loss for one initial state
    (experiments/filament_folding/experiment.py:739-745, 803-916): rest
    angles tanh(raw) * curvature_scale (or ``rest``, the same values computed
    elsewhere, so a float32 rollout can follow another's rounding), a rollout
    of ``num_steps`` steps scored by ``losses.folding_objective`` at the given
    frames, plus the smoothness term.  Returns (total, d total / d raw, frame
    positions); with ``raw_lengths``, the rest lengths are lo + sigmoid(raw)
    (hi - lo) over ``distance_range`` (experiment.py:741-742; or ``lengths``,
    the same values computed elsewhere) and the result is (total, d / d raw,
    d / d raw_lengths, frame positions)."""
    raw = np.asarray(raw, np.float64)
    th = np.tanh(raw)
    r = th * curvature_scale if rest is None else np.asarray(rest, np.float64)
    s = replace(s0, bending_rest_angles=r)
    if raw_lengths is not None:
        lo, hi = distance_range
        sg = 1.0 / (1.0 + np.exp(-np.asarray(raw_lengths, np.float64)))
        ln = lo + sg * (hi - lo) if lengths is None else np.asarray(lengths, np.float64)
        s = replace(s, rest_lengths=ln)
    steps = sorted(set(eval_steps) | set(stab_steps) | ({anchor_step} - {None}))

    def objective(pos):
        total, _, _, bars = LS.folding_objective(pos, target, s0.grid_shape, eval_steps,
                                                 stab_steps, anchor_step, **objective_kw)
        return total, bars

    total, g_rest, g_len, pos = rollout_objective_grad(s, dt, p, num_steps, steps, objective,
                                                       dtype, lengths=True)
    if smoothness_weight > 0.0:
        l, g = LS.smoothness(r)
        total += smoothness_weight * l
        g_rest = g_rest + smoothness_weight * g
    g_raw = g_rest * curvature_scale * (1.0 - th * th)
    if raw_lengths is None:
        return total, g_raw, pos
    return total, g_raw, g_len * (hi - lo) * sg * (1.0 - sg), pos


def rollout_frames_loss(s0: R.State, dt: float, p: R.Params, num_steps: int, frame_losses,
                        rest=None):
    s = s0.astype(np.float64)
    if rest is not None:
        s = replace(s, bending_rest_angles=np.asarray(rest, np.float64))
    total = 0.0
    for t in range(1, num_steps + 1):
        s = R.step(s, dt, p)
        if t in frame_losses:
            total += frame_losses[t](np.asarray(s.position, np.float64))[0]
    return total


def rollout_loss(s0: R.State, dt: float, p: R.Params, num_steps: int, target, rest=None):
    s = s0.astype(np.float64)
    if rest is not None:
        s = replace(s, bending_rest_angles=np.asarray(rest, np.float64))
    for _ in range(num_steps):
        s = R.step(s, dt, p)
    return msd_loss(s.position, target, s0.grid_shape)[0]
