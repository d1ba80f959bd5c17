"""Continuous-batching request server for rollouts.

The reference's callers run one jitted rollout per request (an ES candidate
or an optimisation rollout: gym/base.py:163-196 ``Environment.step`` in a
``lax.scan``; examples/01_es_locomotion.ipynb ``rollout_fitness``), and can
batch only requests of one shape, with ``jax.vmap``.  ``Server`` takes
requests of mixed grid sizes, filament counts and rollout lengths as they
arrive and keeps every admitted request's lattice and rod state resident on
the device until it completes:

* **Admission by HBM footprint.**  A request waits until its estimated
  device footprint (state, plan workspace and the transient copy made when a
  batch is re-formed) fits in the budget next to the admitted requests.
* **Lanes.**  Each lane is a HIP stream running one batch at a time: the
  lane's admitted requests packed into one batched plan, so that every
  kernel launch of a step covers all of them.  The lane advances its batch
  in chunks (``mc_batch_advance``, graph-replayed steps).  With two lanes or
  more, the last lane takes the crowded requests -- those whose filaments
  cover their lattice densely enough to chain into clusters of more than 768
  nodes, solved by the multi-workgroup immersed-boundary launch -- and the
  others the light ones, so that a light request's steps never wait on a
  crowded one's; a lane with nothing of its own class waiting takes the
  other class.
* **Continuous batching.**  At each chunk boundary the lane retires the
  requests that have completed (their final state is copied out), admits
  waiting requests that fit, and re-forms its batch from the survivors'
  current states plus the newcomers' initial states.  Instances keep their
  own clock (actuation time, displacement sum), so a request computes
  bit-for-bit what it computes alone, whatever it was batched with and
  wherever its chunks fell.
* **Overlap.**  While one lane re-forms its batch on the host, the other
  lanes keep the device busy, and the lanes' kernels run concurrently.
* **Gradient requests** (the folding gradient of the reference's
  differentiable experiments, experiments/filament_folding/experiment.py
  :985-1018) run on a lane of their own, one at a time: a checkpointed
  forward rollout and the hand-written adjoint (``grad.RodAdjoint``).

Latency of a request = completion (observed by the server) - arrival.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field, replace
from typing import List, Optional

import numpy as np
import torch

from .batch import Batch
from .structs import Edges, Fields, Nodes


def footprint_bytes(grid_shape, fluid_shape, num_nodes: int, num_edges: int,
                    num_hinges: int, kmax: int = 6) -> int:
    """Estimated device bytes of one request inside a batch: the state, the
    batched plan's workspace (ping-pong lattice, immersed-boundary and
    deposit scratch, FFT spectrum) and the copy made when its batch is
    re-formed."""
    h, w = grid_shape
    hf, wf = fluid_shape
    lat = hf * wf
    grid = h * w
    taps = kmax * kmax
    state = lat * (36 + 8) + grid * 4 + num_nodes * 24 + num_edges * 8 + num_hinges * 12
    plan = lat * (36 + 4 + 4 + 8 + 1) + grid * (8 + 4) + num_nodes * (taps * 20 + 128)
    return int(2 * state + plan)


@dataclass
class Request:
    """One rollout request: an initial state, a number of steps and the
    per-hinge actuation rest_b(t) = amp_b sin(omega_b t + phase_b)."""
    nodes: Nodes
    edges: Edges
    fields: Fields
    steps: int
    actuation: Optional[tuple] = None
    arrival: float = 0.0
    rid: int = -1
    # gradient request: d loss / d bending rest angles of the folding loss
    # (mean squared minimum-image distance of the final positions to target)
    gradient: bool = False
    target: Optional[torch.Tensor] = None
    loss: Optional[torch.Tensor] = None
    rest_angle_grad: Optional[torch.Tensor] = None
    # filled in by the server
    footprint: int = 0
    crowded: bool = False
    admitted: Optional[float] = None
    finished: Optional[float] = None
    result: Optional[tuple] = None          # (Nodes, Edges, Fields) after `steps` steps
    displacement: Optional[torch.Tensor] = None  # (2,) cumulative COM displacement

    @property
    def latency(self) -> Optional[float]:
        return None if self.finished is None else self.finished - self.arrival


@dataclass
class _Slot:
    req: Request
    done: int = 0            # steps completed
    joined_at: int = 0       # steps completed when it joined the lane's batch


@dataclass
class _GradLane:
    stream: torch.cuda.Stream
    reqs: List[Request] = field(default_factory=list)
    adj: Optional[object] = None
    event: Optional[torch.cuda.Event] = None


@dataclass
class _Lane:
    stream: torch.cuda.Stream
    crowded: Optional[bool] = None   # the class of requests it takes (None: any)
    slots: List[_Slot] = field(default_factory=list)
    batch: Optional[Batch] = None
    disp: Optional[torch.Tensor] = None
    event: Optional[torch.cuda.Event] = None
    in_flight: int = 0       # steps of the chunk on the device
    chunks: int = 0
    reforms: int = 0


class Server:
    """Continuous-batching rollout server on one device.

    solver_config / dt: shared by every request (one physics configuration
    per server, as one jitted step function in the reference).
    hbm_budget: device bytes the admitted requests may occupy (default:
    60 % of the device's free memory at construction).
    lanes: concurrent batches (HIP streams).  chunk: the most steps a lane
    runs between scheduling points.  max_batch: most requests per lane.
    """

    def __init__(self, solver_config, dt: float, hbm_budget: Optional[int] = None,
                 lanes: int = 2, chunk: int = 64, max_batch: int = 64, device=None,
                 gradient_config=None, checkpoint_every: int = 50):
        self.dev = torch.device(device) if device is not None else torch.device(
            "cuda", torch.cuda.current_device())
        if self.dev.type != "cuda":
            raise RuntimeError("the server runs on a cuda device")
        self.config = solver_config
        self.dt = float(dt)
        if hbm_budget is None:
            free, _ = torch.cuda.mem_get_info(self.dev)
            hbm_budget = int(0.6 * free)
        self.budget = int(hbm_budget)
        self.chunk = int(chunk)
        self.max_batch = int(max_batch)
        if lanes < 1 or self.chunk < 1 or self.max_batch < 1:
            raise ValueError("lanes, chunk and max_batch must be positive")
        self.lanes = [_Lane(stream=torch.cuda.Stream(device=self.dev),
                            crowded=None if lanes < 2 else (i == lanes - 1))
                      for i in range(lanes)]
        if gradient_config is None:
            from .solver.config import PBD_SCHEME_NO_FLUID
            gradient_config = replace(PBD_SCHEME_NO_FLUID, enable_steric=False)
        self.gradient_config = gradient_config
        self.checkpoint_every = int(checkpoint_every)
        self.glane = _GradLane(stream=torch.cuda.Stream(device=self.dev))
        self.gwaiting: List[Request] = []
        self.waiting: List[Request] = []
        self.completed: List[Request] = []
        self.used = 0
        self.peak_used = 0
        self._next_id = 0
        self.clock = time.perf_counter

    # ---- submission -------------------------------------------------------
    def submit(self, req: Request) -> Request:
        if req.steps < 1:
            raise ValueError("a request needs at least one step")
        n = int(req.nodes.position.shape[0])
        if n == 0:
            raise ValueError("a request needs at least one node")
        fshape = tuple(req.fields.f_grid.shape[1:])
        req.footprint = footprint_bytes(req.fields.grid_shape, fshape, n,
                                        int(req.edges.pairs.shape[0]),
                                        int(req.edges.bending_pairs.shape[0]))
        # crowded: K^2 stencil cells per node cover at least half the lattice
        k = int(getattr(self.config, "ibm_kernel_size", 4))
        req.crowded = n > 768 and n * k * k >= 0.5 * fshape[0] * fshape[1]
        if req.footprint > self.budget:
            raise ValueError(f"request needs ~{req.footprint} B, more than the budget {self.budget} B")
        if req.gradient:
            # checkpoints and one recomputed segment of the rollout on top
            per_state = 4 * (4 * n + int(req.edges.pairs.shape[0]))
            req.footprint += per_state * (req.steps // self.checkpoint_every + 2 +
                                          self.checkpoint_every)
            if req.target is None:
                raise ValueError("a gradient request needs a target curve")
        if req.rid < 0:
            req.rid = self._next_id
        self._next_id += 1
        (self.gwaiting if req.gradient else self.waiting).append(req)
        return req

    @property
    def active(self) -> int:
        return sum(len(l.slots) for l in self.lanes) + len(self.glane.reqs)

    def idle(self) -> bool:
        return not self.waiting and not self.gwaiting and self.active == 0

    # ---- scheduling -------------------------------------------------------
    def poll(self) -> List[Request]:
        """One scheduling pass: harvest lanes whose chunk finished, retire
        completed requests, admit waiting ones, launch the next chunks.
        Returns the requests completed in this pass."""
        done: List[Request] = self._poll_gradient()
        free = [l for l in self.lanes if l.event is None or l.event.query()]
        if not free:
            return done
        for lane in free:
            if lane.event is not None:
                # the chunk has finished: a barrier timeout of its wide IB
                # solve is reported here, at the chunk boundary
                if lane.batch is not None:
                    lane.batch.check()
                lane.event = None
                for sl in lane.slots:
                    sl.done += lane.in_flight
                lane.in_flight = 0
            done.extend(self._reform(lane, self._share(lane, free)))
            self._launch(lane)
        return done

    def _poll_gradient(self) -> List[Request]:
        """The gradient lane: retire the finished batch, start the next one --
        the waiting gradient requests at the head of the queue with the same
        step count (up to max_batch, within the budget), one batched adjoint
        (each request's gradient is the one it gets alone)."""
        g = self.glane
        out: List[Request] = []
        if g.reqs:
            if not g.event.query():
                return out
            now = self.clock()
            for r in g.reqs:
                r.finished = now
                self.used -= r.footprint
                self.completed.append(r)
                out.append(r)
            g.adj.batch.plan.release(int(g.stream.cuda_stream))
            g.reqs, g.adj, g.event = [], None, None
        take: List[Request] = []
        while (self.gwaiting and len(take) < self.max_batch
               and self.gwaiting[0].steps == (take[0].steps if take else self.gwaiting[0].steps)
               and self.used + self.gwaiting[0].footprint <= self.budget):
            r = self.gwaiting.pop(0)
            self.used += r.footprint
            self.peak_used = max(self.peak_used, self.used)
            r.admitted = self.clock()
            take.append(r)
        if not take:
            return out
        from .grad import BatchRodAdjoint, min_image
        with torch.cuda.stream(g.stream):
            g.stream.wait_stream(torch.cuda.current_stream(self.dev))
            adj = BatchRodAdjoint([(r.nodes, r.edges, r.fields) for r in take],
                                  self.gradient_config, self.dt)
            xs = adj.forward(take[0].steps, self.checkpoint_every)
            bars = []
            for r, x in zip(take, xs):
                d = min_image(tuple(r.fields.grid_shape), x, r.target)
                n = x.shape[0]
                r.loss = (d * d).sum() / n
                bars.append(2.0 * d / n)
            for r, gr in zip(take, adj.backward(bars)):
                r.rest_angle_grad = gr
            ev = torch.cuda.Event()
            ev.record(g.stream)
        g.reqs, g.adj, g.event = take, adj, ev
        return out

    def _takes(self, lane: _Lane, r: Request) -> bool:
        """Whether ``lane`` admits ``r``: its own class, or the other class
        when none of its own is waiting and it is idle."""
        if lane.crowded is None or r.crowded == lane.crowded:
            return True
        own = any(w.crowded == lane.crowded for w in self.waiting)
        return not own and not lane.slots

    def _share(self, lane: _Lane, free) -> int:
        """Waiting requests this lane may take at this boundary: its class's
        waiting count split among the free lanes of that class."""
        mine = [w for w in self.waiting if self._takes(lane, w)]
        peers = [l for l in free if l.crowded == lane.crowded] or [lane]
        return max(1, -(-len(mine) // len(peers)))

    def _admit(self, lane: _Lane, share: int) -> List[Request]:
        out = []
        i = 0
        while (i < len(self.waiting) and len(out) < share
               and len(lane.slots) + len(out) < self.max_batch):
            r = self.waiting[i]
            if not self._takes(lane, r):
                i += 1
                continue
            if self.used + r.footprint > self.budget:
                break  # first come first served within the class
            self.waiting.pop(i)
            self.used += r.footprint
            self.peak_used = max(self.peak_used, self.used)
            r.admitted = self.clock()
            out.append(r)
        return out

    def _reform(self, lane: _Lane, share: int) -> List[Request]:
        finished = [i for i, sl in enumerate(lane.slots) if sl.done >= sl.req.steps]
        newcomers = self._admit(lane, share)
        if not finished and not newcomers:
            return []
        s = lane.stream
        old = lane.batch
        keep = [i for i in range(len(lane.slots)) if i not in set(finished)]
        out = []
        with torch.cuda.stream(s):
            for i in finished:
                sl = lane.slots[i]
                n, e, f = old.instance(i)
                sl.req.result = (
                    Nodes(position=n.position.clone(), velocity=n.velocity.clone(),
                          color_bend=n.color_bend, debug_vector=n.debug_vector.clone()),
                    replace(e, theta=e.theta.clone()),
                    replace(f, steric=f.steric.clone(), fluid_velocity=f.fluid_velocity.clone(),
                            f_grid=f.f_grid.clone()))
                sl.req.displacement = lane.disp[i].clone()
            insts, acts, disps, slots = [], [], [], []
            for i in keep:
                insts.append(old.instance(i))
                acts.append(lane.slots[i].req.actuation)
                disps.append(lane.disp[i:i + 1])
                slots.append(lane.slots[i])
            if newcomers:
                # initial states were written on the submitting stream
                s.wait_stream(torch.cuda.current_stream(self.dev))
            for r in newcomers:
                for t in (r.nodes.position, r.nodes.velocity, r.nodes.debug_vector,
                          r.edges.theta, r.fields.steric, r.fields.fluid_velocity,
                          r.fields.f_grid):
                    if isinstance(t, torch.Tensor) and t.is_cuda:
                        t.record_stream(s)
                insts.append((r.nodes, r.edges, r.fields))
                acts.append(r.actuation)
                disps.append(torch.zeros((1, 2), dtype=torch.float32, device=self.dev))
                slots.append(_Slot(req=r))
            if insts:
                if any(a is None for a in acts) and not all(a is None for a in acts):
                    raise ValueError("requests of one server must all have or all lack actuation")
                new = Batch(insts, self.config, self.dt,
                            actuation=None if acts[0] is None else acts, device=self.dev, stream=s)
                disp = torch.cat(disps).contiguous()
            else:
                new, disp = None, None
        if old is not None:
            old.close()  # waits for this lane's copies only
        for sl in slots:
            sl.joined_at = sl.done
        now = self.clock()
        for i in finished:
            r = lane.slots[i].req
            r.finished = now
            self.used -= r.footprint
            out.append(r)
        lane.slots, lane.batch, lane.disp = slots, new, disp
        lane.reforms += 1
        self.completed.extend(out)
        return out

    def _launch(self, lane: _Lane):
        if lane.batch is None or not lane.slots or lane.event is not None:
            return
        left = min(sl.req.steps - sl.done for sl in lane.slots)
        k = min(self.chunk, left)
        t0 = [sl.joined_at for sl in lane.slots]
        lane.batch.advance(k, t0, lane.disp)
        ev = torch.cuda.Event()
        ev.record(lane.stream)
        lane.event = ev
        lane.in_flight = k
        lane.chunks += 1

    # ---- driving ----------------------------------------------------------
    def run(self, requests: List[Request], realtime: bool = True) -> List[Request]:
        """Serves ``requests`` (sorted by arrival, seconds from the start of
        the run).  realtime: release each request at its arrival time on the
        wall clock; otherwise submit all at once.  Returns them completed."""
        pending = sorted(requests, key=lambda r: r.arrival)
        t_start = self.clock()
        base = {id(r): r.arrival for r in pending}
        for r in pending:
            r.arrival = t_start + (base[id(r)] if realtime else 0.0)
        i = 0
        while i < len(pending) or not self.idle():
            now = self.clock()
            while i < len(pending) and pending[i].arrival <= now:
                self.submit(pending[i])
                i += 1
            got = self.poll()
            if not got and self.active and all(l.event is not None for l in self.lanes if l.slots):
                time.sleep(20e-6)
        for r in pending:
            r.arrival -= t_start
            r.admitted -= t_start
            r.finished -= t_start
        return pending


def latency_report(requests: List[Request]) -> dict:
    """p50 / p99 latency, mean queue wait and throughput of a served trace."""
    lat = np.array([r.latency for r in requests], dtype=np.float64)
    wait = np.array([r.admitted - r.arrival for r in requests], dtype=np.float64)
    first = min(r.arrival for r in requests)
    last = max(r.finished for r in requests)
    return {"requests": len(requests),
            "p50_latency_s": float(np.percentile(lat, 50)),
            "p99_latency_s": float(np.percentile(lat, 99)),
            "mean_queue_wait_s": float(wait.mean()),
            "makespan_s": float(last - first),
            "rollouts_per_s": float(len(requests) / (last - first))}


def fifo_report(arrivals, service_times) -> dict:
    """Latency of the same trace served one request at a time in arrival
    order (a single FIFO server) from measured per-request service times."""
    order = np.argsort(arrivals, kind="stable")
    t = -np.inf
    lat = np.empty(len(order))
    fin = np.empty(len(order))
    for k in order:
        start = max(t, arrivals[k])
        t = start + service_times[k]
        fin[k] = t
        lat[k] = t - arrivals[k]
    first = float(np.min(arrivals))
    return {"requests": int(len(order)),
            "p50_latency_s": float(np.percentile(lat, 50)),
            "p99_latency_s": float(np.percentile(lat, 99)),
            "makespan_s": float(fin.max() - first),
            "rollouts_per_s": float(len(order) / (fin.max() - first))}
