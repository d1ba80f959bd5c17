"""Batched evaluation of many independent simulations (mc_batch_plan_create,
mc_batch_rollout).

The reference evaluates an ES population with ``jax.vmap`` over rollouts of
one shape (examples/01_es_locomotion.ipynb, ``rollout_fitness`` inside
``jax.vmap(fitness)``).  Here instances may differ in grid size, lattice size,
filament count and node count: ``Batch`` packs their states into concatenated
device buffers (instance order), builds one batched plan, and every kernel
launch of a step covers all instances.  Each instance computes exactly what
it computes alone.
"""
from __future__ import annotations

import ctypes
from dataclasses import replace

import numpy as np
import torch

from . import _lib
from ._lib import McActuation, McState, check
from ._runtime import KMAX, _host_topology, as_f32, params_from_config, ptr, stream_handle
from .structs import Edges, Fields, Nodes


class _nullctx:
    def __enter__(self):
        return None

    def __exit__(self, *a):
        return False


def _action_rows(d_rest_length, d_bending_angle, num_edges: int, num_hinges: int, dev):
    """The mc_actions of per-step action arrays (T, E) / (T, B), either None;
    returns it with the device tensors it points at."""
    arrs = []
    rows = None
    for a, width, name in ((d_rest_length, num_edges, "d_rest_length"),
                           (d_bending_angle, num_hinges, "d_bending_angle")):
        if a is None:
            arrs.append(None)
            continue
        t = as_f32(a, dev).contiguous()
        if t.dim() != 2 or t.shape[1] != width:
            raise ValueError(f"{name} must be (steps, {width}), got {tuple(t.shape)}")
        if rows is not None and t.shape[0] != rows:
            raise ValueError("d_rest_length and d_bending_angle have different step counts")
        rows = t.shape[0]
        arrs.append(t)
    if rows is None:
        raise ValueError("no action arrays given")
    m = _lib.McActions(ptr(arrs[0]), ptr(arrs[1]), int(rows))
    return m, arrs


class BatchPlan:
    """Owns one batched mc_plan handle."""

    def __init__(self, shapes, node_offsets, pairs, bending_pairs, kmax: int = KMAX):
        lib = _lib.load()
        self.shapes = np.ascontiguousarray(shapes, dtype=np.int32).reshape(-1, 4)
        self.node_offsets = np.ascontiguousarray(node_offsets, dtype=np.int32)
        self.pairs = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1, 2)
        self.bending_pairs = np.ascontiguousarray(bending_pairs, dtype=np.int32).reshape(-1, 2)
        h = ctypes.c_void_p()
        check(lib.mc_batch_plan_create(
            ctypes.byref(h), int(self.shapes.shape[0]), self.shapes.ctypes.data,
            self.node_offsets.ctypes.data,
            self.pairs.ctypes.data if self.pairs.size else None, self.pairs.shape[0],
            self.bending_pairs.ctypes.data if self.bending_pairs.size else None,
            self.bending_pairs.shape[0], int(kmax)))
        self.handle = h
        self._lib = lib

    @property
    def num_nodes(self):
        return int(self.node_offsets[-1])

    def workspace_bytes(self) -> int:
        b = ctypes.c_size_t()
        check(self._lib.mc_plan_workspace_bytes(self.handle, ctypes.byref(b)))
        return int(b.value)

    def status(self, stream: int):
        """Waits for ``stream`` and raises MicrocosmosError if the plan's work
        did not complete as specified (mc_plan_status)."""
        check(self._lib.mc_plan_status(self.handle, stream))

    def release(self, stream: int):
        """Frees the plan after the work enqueued on ``stream`` (a hipStream_t
        handle) has finished, without synchronising other streams."""
        if getattr(self, "handle", None) and self.handle.value:
            check(self._lib.mc_plan_destroy_after(self.handle, stream))
            self.handle = ctypes.c_void_p()

    def __del__(self):
        if getattr(self, "handle", None) and self.handle.value:
            self._lib.mc_plan_destroy(self.handle)
            self.handle = ctypes.c_void_p()


class Batch:
    """Independent instances stepped together.

    instances: list of (Nodes, Edges, Fields) with each instance's own node,
    edge and hinge numbering (as the reference builds them).  actuation: None
    or a list (one per instance) of (amp, omega, phase) per-hinge arrays for
    rest_b(t) = amp sin(omega t + phase), applied in-kernel.
    """

    def __init__(self, instances, solver_config, dt: float, actuation=None, device=None,
                 stream=None):
        if not instances:
            raise ValueError("empty batch")
        dev = torch.device(device) if device is not None else instances[0][0].position.device
        self.dev = dev
        self.stream = stream
        with torch.cuda.stream(stream) if stream is not None else _nullctx():
            self._build(instances, solver_config, dt, actuation, dev)

    def _build(self, instances, solver_config, dt, actuation, dev):
        self.params = params_from_config(solver_config)
        self.dt = float(dt)
        self.n = len(instances)
        shapes, noff, eoff, boff = [], [0], [0], [0]
        pairs, bps = [], []
        for nodes, edges, fields in instances:
            h, w = fields.grid_shape
            hf, wf = fields.f_grid.shape[1:]
            shapes.append((h, w, hf, wf))
            pa = _host_topology(edges.pairs)[0].reshape(-1, 2)
            bp = _host_topology(edges.bending_pairs)[0].reshape(-1, 2)
            pairs.append(pa + noff[-1])
            bps.append(bp + eoff[-1])
            noff.append(noff[-1] + nodes.position.shape[0])
            eoff.append(eoff[-1] + pa.shape[0])
            boff.append(boff[-1] + bp.shape[0])
        self.shapes = np.asarray(shapes, np.int64)
        self.node_off = np.asarray(noff, np.int64)
        self.edge_off = np.asarray(eoff, np.int64)
        self.hinge_off = np.asarray(boff, np.int64)
        self.plan = BatchPlan(shapes, noff, np.concatenate(pairs) if pairs else np.zeros((0, 2)),
                              np.concatenate(bps) if bps else np.zeros((0, 2)))

        def cat(ts):
            return torch.cat([as_f32(t, dev).reshape(-1) for t in ts]).contiguous()

        self.position = cat([i[0].position for i in instances])
        self.velocity = cat([i[0].velocity for i in instances])
        self.debug_vector = cat([i[0].debug_vector for i in instances])
        self.theta = cat([i[1].theta for i in instances])
        self.rest_lengths = cat([i[1].rest_lengths for i in instances])
        self.rest_angles = cat([i[1].bending_rest_angles for i in instances])
        self.stiffness = cat([i[1].bending_stiffness for i in instances])
        self.steric = cat([i[2].steric for i in instances])
        self.fluid_velocity = cat([i[2].fluid_velocity for i in instances])
        self.f_grid = cat([i[2].f_grid for i in instances])
        self._templates = instances
        self.act = None
        if actuation is not None:
            self.act = tuple(cat([a[k] for a in actuation]) for k in range(3))
        self.t = 0

    def _stream(self) -> int:
        return int(self.stream.cuda_stream) if self.stream is not None else stream_handle(self.dev)

    def _state(self):
        return McState(ptr(self.position), ptr(self.velocity), ptr(self.debug_vector),
                       ptr(self.theta), ptr(self.steric), ptr(self.fluid_velocity),
                       ptr(self.f_grid))

    def run(self, num_steps: int, track: bool = True):
        """num_steps graph-replayed steps of every instance.  Returns the
        (n, 2) cumulative centre-of-mass displacement of each instance over
        these steps (swim_base.py:179-209) when track, else None."""
        lib = _lib.load()
        st = self._state()
        ep = _lib.McEdgeParams(ptr(self.rest_lengths), ptr(self.rest_angles), ptr(self.stiffness))
        act = None
        if self.act is not None:
            a, o, p = self.act
            act = McActuation(ptr(a), ptr(o), ptr(p), int(self.t))
        disp = torch.empty((self.n, 2), dtype=torch.float32, device=self.dev) if track else None
        check(lib.mc_batch_rollout(self.plan.handle, ctypes.byref(self.params), ctypes.byref(st),
                                   ctypes.byref(ep), ctypes.byref(act) if act else None, self.dt,
                                   int(num_steps), ptr(disp), self._stream()))
        self.t += int(num_steps)
        return disp

    def autotune(self, steps: int = 8) -> dict:
        """mc_plan_autotune on the batch's plan (a copy of its state is
        stepped); returns the chosen launch knobs."""
        from .simulate import tuning
        lib = _lib.load()
        st = self._state()
        ep = _lib.McEdgeParams(ptr(self.rest_lengths), ptr(self.rest_angles), ptr(self.stiffness))
        check(lib.mc_plan_autotune(self.plan.handle, ctypes.byref(self.params), ctypes.byref(st),
                                   ctypes.byref(ep), self.dt, int(steps), self._stream()))
        return tuning(self.plan.handle)

    def run_actions(self, d_rest_length=None, d_bending_angle=None, track: bool = True):
        """One graph-replayed step per action row (gym/base.py:163-176): before
        step k every instance's rest lengths and rest angles are its base
        values plus row k of ``d_rest_length`` (T, total edges) and
        ``d_bending_angle`` (T, total hinges), each float32 on the device or
        None (a zero action), in the batch's concatenated numbering
        (``edge_off`` / ``hinge_off``).  Returns the (n, 2) displacement as
        ``run`` does."""
        rows, keep = _action_rows(d_rest_length, d_bending_angle, int(self.edge_off[-1]),
                                  int(self.hinge_off[-1]), self.dev)
        lib = _lib.load()
        st = self._state()
        ep = _lib.McEdgeParams(ptr(self.rest_lengths), ptr(self.rest_angles), ptr(self.stiffness))
        disp = torch.empty((self.n, 2), dtype=torch.float32, device=self.dev) if track else None
        check(lib.mc_rollout_actions(self.plan.handle, ctypes.byref(self.params),
                                     ctypes.byref(st), ctypes.byref(ep), ctypes.byref(rows),
                                     self.dt, int(rows.num_rows), ptr(disp), self._stream()))
        self._action_keep = keep  # read by the enqueued steps
        self.t += int(rows.num_rows)
        return disp

    def advance(self, num_steps: int, instance_t0, displacement: torch.Tensor):
        """num_steps steps of every instance, instance i at its own clock
        (actuation time instance_t0[i] + steps this batch has run), adding
        each instance's centre-of-mass displacement over these steps to
        ``displacement`` (n, 2) in place (mc_batch_advance).  Stream-ordered:
        returns without waiting."""
        lib = _lib.load()
        st = self._state()
        ep = _lib.McEdgeParams(ptr(self.rest_lengths), ptr(self.rest_angles), ptr(self.stiffness))
        act = None
        t0 = np.ascontiguousarray(np.asarray(instance_t0, dtype=np.int64).reshape(-1))
        if t0.shape[0] != self.n:
            raise ValueError(f"instance_t0 has {t0.shape[0]} entries for {self.n} instances")
        if displacement.shape != (self.n, 2) or displacement.dtype != torch.float32:
            raise ValueError("displacement must be a float32 (n, 2) device tensor")
        if self.act is not None:
            a, o, p = self.act
            act = McActuation(ptr(a), ptr(o), ptr(p), int(self.t))
        check(lib.mc_batch_advance(self.plan.handle, ctypes.byref(self.params), ctypes.byref(st),
                                   ctypes.byref(ep), ctypes.byref(act) if act else None,
                                   t0.ctypes.data, self.dt, int(num_steps), ptr(displacement),
                                   self._stream()))
        self.t += int(num_steps)

    def check(self):
        """Waits for this batch's stream and raises MicrocosmosError if its
        steps did not complete as specified (mc_plan_status)."""
        self.plan.status(self._stream())

    def close(self):
        """Releases the plan once this batch's stream has finished its work."""
        if self.plan is not None:
            self.plan.release(self._stream())
            self.plan = None

    def instance(self, i: int):
        """(Nodes, Edges, Fields) views of instance i's current state."""
        nodes, edges, fields = self._templates[i]
        n0, n1 = int(self.node_off[i]), int(self.node_off[i + 1])
        e0, e1 = int(self.edge_off[i]), int(self.edge_off[i + 1])
        h, w, hf, wf = (int(v) for v in self.shapes[i])
        g0 = int(sum(int(s[0]) * int(s[1]) for s in self.shapes[:i]))
        f0 = int(sum(int(s[2]) * int(s[3]) for s in self.shapes[:i]))
        nd = Nodes(position=self.position[2 * n0:2 * n1].view(-1, 2),
                   velocity=self.velocity[2 * n0:2 * n1].view(-1, 2),
                   color_bend=nodes.color_bend,
                   debug_vector=self.debug_vector[2 * n0:2 * n1].view(-1, 2))
        ed = replace(edges, theta=self.theta[e0:e1])
        fl = replace(fields, steric=self.steric[g0:g0 + h * w].view(h, w),
                     fluid_velocity=self.fluid_velocity[2 * f0:2 * f0 + 2 * hf * wf].view(2, hf, wf),
                     f_grid=self.f_grid[9 * f0:9 * f0 + 9 * hf * wf].view(9, hf, wf))
        return nd, ed, fl
