"""step / simulate (simulate.py:15-100) on the MI355X device path.

``step(nodes, edges, fields, dt, solver_config)`` keeps the reference's
signature and functional semantics (new output arrays; inputs untouched) and
runs the whole coupled step as one stream-ordered sequence of HIP kernels:
steric field -> rod solve -> immersed boundary -> fused LBM collide-stream.
"""
from __future__ import annotations

import ctypes
from dataclasses import replace

import torch

from . import _lib
from ._lib import McActuation, McState, check
from ._runtime import Plan, as_f32, edge_params, params_from_config, plan_for, ptr, stream_handle
from .structs import Edges, Fields, Nodes


def _device_of(nodes, fields):
    dev = fields.f_grid.device if fields.f_grid is not None else nodes.position.device
    if dev.type != "cuda":
        raise RuntimeError("microcosmos_amd runs on the GPU only: move the state to a cuda device")
    return dev


def step(nodes: Nodes, edges: Edges, fields: Fields, dt: float, solver_config):
    """One coupled step (simulate.py:37-100)."""
    dev = _device_of(nodes, fields)
    lib = _lib.load()
    params = params_from_config(solver_config)
    H, W = fields.grid_shape
    f_in = as_f32(fields.f_grid, dev)
    Hf, Wf = f_in.shape[1], f_in.shape[2]
    N = nodes.position.shape[0]
    plan = plan_for(N, edges.pairs, edges.bending_pairs, (H, W), (Hf, Wf), dev)
    pos = as_f32(nodes.position, dev)
    vel = as_f32(nodes.velocity, dev)
    dbg = as_f32(nodes.debug_vector, dev)
    th = as_f32(edges.theta, dev)
    st = as_f32(fields.steric, dev)
    fv = as_f32(fields.fluid_velocity, dev)
    o_pos, o_vel, o_dbg, o_th = (torch.empty_like(pos), torch.empty_like(vel),
                                 torch.empty_like(dbg), torch.empty_like(th))
    o_st = torch.empty_like(st) if params.enable_steric else st
    if params.enable_fluid:
        o_fv, o_f = torch.empty_like(fv), torch.empty_like(f_in)
    else:
        o_fv, o_f = fv, f_in
    sin = McState(ptr(pos), ptr(vel), ptr(dbg), ptr(th), ptr(st), ptr(fv), ptr(f_in))
    sout = McState(ptr(o_pos), ptr(o_vel), ptr(o_dbg), ptr(o_th), ptr(o_st), ptr(o_fv), ptr(o_f))
    ep, keep = edge_params(edges, dev)
    check(lib.mc_step(plan.handle, ctypes.byref(params), ctypes.byref(sin), ctypes.byref(sout),
                      ctypes.byref(ep), float(dt), stream_handle(dev)))
    new_nodes = Nodes(position=o_pos, velocity=o_vel, color_bend=nodes.color_bend,
                      debug_vector=o_dbg)
    new_edges = replace(edges, theta=o_th)
    new_fields = replace(fields, steric=o_st, fluid_velocity=o_fv, f_grid=o_f)
    return new_nodes, new_edges, new_fields


def simulate(nodes: Nodes, edges: Edges, fields: Fields, dt: float, num_steps: int,
             solver_config, *, record_every: int = 1):
    """simulate.py:15-35 -- returns (nodes_timeseries, fields_timeseries), each a
    pytree of arrays stacked over steps; f_grid is dropped from the history and
    the fluid velocity is recorded at the grid's resolution (linearly resized
    from the lattice when they differ).  The steps run as graph-replayed
    device rollouts (mc_rollout_record) with the frames written on the device;
    ``record_every`` > 1 keeps every k-th step's frame only."""
    num_steps = int(num_steps)
    k = int(record_every)
    if k < 1:
        raise ValueError("record_every must be at least 1")
    ro = Rollout(nodes, edges, fields, solver_config, dt)
    frames = ro.record(num_steps, stride=k)
    n_frames = num_steps // k
    cb = torch.as_tensor(nodes.color_bend)
    nts = Nodes(position=frames["position"], velocity=frames["velocity"],
                color_bend=cb.unsqueeze(0).expand(n_frames, *cb.shape).clone(),
                debug_vector=frames["debug_vector"])
    fts = replace(fields, steric=frames["steric"], fluid_velocity=frames["fluid_velocity"],
                  f_grid=None)
    return nts, fts


def tuning(plan_handle) -> dict:
    """A plan's launch choices (mc_plan_tuning)."""
    lib = _lib.load()
    wm, mode = ctypes.c_int32(), ctypes.c_int32()
    check(lib.mc_plan_tuning(plan_handle, ctypes.byref(wm), ctypes.byref(mode)))
    return {"wide_min": int(wm.value),
            "lbm_pass": ("own launch", "in next rod launch", "second stream")[int(mode.value)]}


class Rollout:
    """Device-resident, in-place rollout state for long runs (mc_rollout).

    Holds one copy of every state array on the device and advances it with
    graph-captured steps; optional in-kernel actuation of the rest angles.
    """

    def __init__(self, nodes: Nodes, edges: Edges, fields: Fields, solver_config, dt: float,
                 actuation=None):
        self.dev = _device_of(nodes, fields)
        self.params = params_from_config(solver_config)
        self.dt = float(dt)
        d = self.dev
        self.position = as_f32(nodes.position, d).clone()
        self.velocity = as_f32(nodes.velocity, d).clone()
        self.debug_vector = as_f32(nodes.debug_vector, d).clone()
        self.color_bend = nodes.color_bend
        self.theta = as_f32(edges.theta, d).clone()
        self.steric = as_f32(fields.steric, d).clone()
        self.fluid_velocity = as_f32(fields.fluid_velocity, d).clone()
        self.f_grid = as_f32(fields.f_grid, d).clone()
        self.edges = edges
        self.fields = fields
        self.grid_shape = tuple(fields.grid_shape)
        self.plan = plan_for(self.position.shape[0], edges.pairs, edges.bending_pairs,
                             self.grid_shape, tuple(self.f_grid.shape[1:]), d)
        self.ep, self._keep = edge_params(edges, d)
        self.t = 0
        self.act = None
        if actuation is not None:
            amp, omega, phase = (as_f32(a, d) for a in actuation)
            self._act_keep = (amp, omega, phase)
            self.act = (amp, omega, phase)

    def state(self) -> McState:
        return McState(ptr(self.position), ptr(self.velocity), ptr(self.debug_vector),
                       ptr(self.theta), ptr(self.steric), ptr(self.fluid_velocity),
                       ptr(self.f_grid))

    def run(self, num_steps: int, use_graph: bool = True, record_position=None,
            record_fluid_velocity=None):
        lib = _lib.load()
        st = self.state()
        act = None
        if self.act is not None:
            a, o, p = self.act
            act = McActuation(ptr(a), ptr(o), ptr(p), int(self.t))
        check(lib.mc_rollout(self.plan.handle, ctypes.byref(self.params), ctypes.byref(st),
                             ctypes.byref(self.ep), ctypes.byref(act) if act else None,
                             self.dt, int(num_steps), ptr(record_position),
                             ptr(record_fluid_velocity), int(bool(use_graph)),
                             stream_handle(self.dev)))
        self.t += int(num_steps)

    def run_actions(self, d_rest_length=None, d_bending_angle=None):
        """One graph-replayed step per action row (gym/base.py:163-176): before
        step k the rest lengths and rest angles are the edges' base values
        plus row k of ``d_rest_length`` (T, E) and ``d_bending_angle`` (T, B)
        (float32, either None for a zero action)."""
        from .batch import _action_rows
        rows, keep = _action_rows(d_rest_length, d_bending_angle, int(self.edges.pairs.shape[0]),
                                  int(self.edges.bending_pairs.shape[0]), self.dev)
        lib = _lib.load()
        st = self.state()
        check(lib.mc_rollout_actions(self.plan.handle, ctypes.byref(self.params),
                                     ctypes.byref(st), ctypes.byref(self.ep), ctypes.byref(rows),
                                     self.dt, int(rows.num_rows), None, stream_handle(self.dev)))
        self._action_keep = keep  # read by the enqueued steps
        self.t += int(rows.num_rows)

    def autotune(self, steps: int = 8) -> dict:
        """Times the plan's launch choices on a copy of the current state and
        keeps the fastest (mc_plan_autotune); returns them.  Results do not
        change: every choice computes the same steps bit for bit."""
        lib = _lib.load()
        st = self.state()
        check(lib.mc_plan_autotune(self.plan.handle, ctypes.byref(self.params), ctypes.byref(st),
                                   ctypes.byref(self.ep), self.dt, int(steps),
                                   stream_handle(self.dev)))
        return tuning(self.plan.handle)

    def set_tuning(self, wide_min: int = 768, lbm_pass: int = 1):
        """Sets the plan's launch choices by hand (mc_plan_set_tuning):
        lbm_pass 0 its own launch, 1 in the next rod launch, 2 on a second
        stream."""
        check(_lib.load().mc_plan_set_tuning(self.plan.handle, int(wide_min), int(lbm_pass)))
        return tuning(self.plan.handle)

    def record(self, num_steps: int, stride: int = 1, use_graph: bool = True):
        """num_steps steps with a frame of the node arrays, the steric field
        and the grid-resolution fluid velocity after every stride-th step,
        written on the device inside the step graphs (mc_rollout_record).
        Returns a dict of (frames, ...) device tensors: position, velocity,
        debug_vector (F, N, 2), steric (F, H, W), fluid_velocity
        (F, 2, H, W), F = num_steps // stride."""
        lib = _lib.load()
        d = self.dev
        F = int(num_steps) // int(stride)
        N = self.position.shape[0]
        H, W = self.grid_shape
        out = {
            "position": torch.empty((F, N, 2), dtype=torch.float32, device=d),
            "velocity": torch.empty((F, N, 2), dtype=torch.float32, device=d),
            "debug_vector": torch.empty((F, N, 2), dtype=torch.float32, device=d),
            "steric": torch.empty((F, H, W), dtype=torch.float32, device=d),
            "fluid_velocity": torch.empty((F, 2, H, W), dtype=torch.float32, device=d),
        }
        rec = _lib.McRecord(ptr(out["position"]), ptr(out["velocity"]),
                            ptr(out["debug_vector"]), ptr(out["steric"]),
                            ptr(out["fluid_velocity"]), int(stride))
        st = self.state()
        act = None
        if self.act is not None:
            a, o, p = self.act
            act = McActuation(ptr(a), ptr(o), ptr(p), int(self.t))
        check(lib.mc_rollout_record(self.plan.handle, ctypes.byref(self.params), ctypes.byref(st),
                                    ctypes.byref(self.ep), ctypes.byref(act) if act else None,
                                    self.dt, int(num_steps), ctypes.byref(rec),
                                    int(bool(use_graph)), stream_handle(d)))
        self.t += int(num_steps)
        return out

    def profile(self, num_steps: int):
        """num_steps steps (no graph, each LBM pass its own launch) with HIP
        events carried by each LBM launch: returns (average LBM kernel ms,
        average step ms)."""
        lib = _lib.load()
        st = self.state()
        act = None
        if self.act is not None:
            a, o, p = self.act
            act = McActuation(ptr(a), ptr(o), ptr(p), int(self.t))
        lbm = ctypes.c_double()
        stp = ctypes.c_double()
        check(lib.mc_profile_step(self.plan.handle, ctypes.byref(self.params), ctypes.byref(st),
                                  ctypes.byref(self.ep), ctypes.byref(act) if act else None,
                                  self.dt, int(num_steps), ctypes.byref(lbm), ctypes.byref(stp),
                                  stream_handle(self.dev)))
        self.t += int(num_steps)
        return lbm.value, stp.value

    def profile_fused(self, num_steps: int):
        """num_steps (even) steps with each LBM pass carried by the next
        step's rod launch, as in the rollout graphs, and HIP events carried by
        those launches: returns (average rod+LBM launch ms, average step ms)."""
        lib = _lib.load()
        st = self.state()
        act = None
        if self.act is not None:
            a, o, p = self.act
            act = McActuation(ptr(a), ptr(o), ptr(p), int(self.t))
        kern = ctypes.c_double()
        stp = ctypes.c_double()
        check(lib.mc_profile_fused_step(self.plan.handle, ctypes.byref(self.params),
                                        ctypes.byref(st), ctypes.byref(self.ep),
                                        ctypes.byref(act) if act else None, self.dt,
                                        int(num_steps), ctypes.byref(kern), ctypes.byref(stp),
                                        stream_handle(self.dev)))
        self.t += int(num_steps)
        return kern.value, stp.value

    def track(self, groups=None, start_step: int = 0):
        """Displacement tracking of swim_base.py:179-209 over this rollout's
        steps, on the device and inside the step graphs.  groups: node
        slices / offsets (default: all nodes).  The rollout switches to a
        private plan so that other callers of the same topology are not
        tracked.  Returns the DisplacementTracker (reset to the current
        positions)."""
        from .fitness import DisplacementTracker
        self.plan = Plan(self.grid_shape, tuple(self.f_grid.shape[1:]), self.position.shape[0],
                         self.plan.pairs, self.plan.bending_pairs)
        self.tracker = DisplacementTracker(self.plan, self.grid_shape, groups,
                                           start_step=start_step)
        self.tracker.init(self.position)
        return self.tracker

    def check(self):
        """Waits for the rollout's stream and raises MicrocosmosError if its
        steps did not complete as specified (mc_plan_status)."""
        self.plan.status(stream_handle(self.dev))

    def nodes(self) -> Nodes:
        return Nodes(position=self.position, velocity=self.velocity, color_bend=self.color_bend,
                     debug_vector=self.debug_vector)

    def edges_now(self) -> Edges:
        return replace(self.edges, theta=self.theta)

    def fields_now(self) -> Fields:
        return replace(self.fields, steric=self.steric, fluid_velocity=self.fluid_velocity,
                       f_grid=self.f_grid)
