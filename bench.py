"""Coupled fluid-filament step on one MI355X: MLUPS of the 1024^2 plain case.

A step is one full coupled step (steric field by a direct Gaussian filter, 8
PBD cycles on 256 x 64 filaments, 5 immersed-boundary iterations with the
K = 4 stencil, fused D2Q9 TRT + Guo collide-stream) over the device-resident
1024^2 lattice, replayed from a HIP graph.  Inputs are resident in HBM before the timed region.

Prints one JSON line.  Also reports the roofline fraction of the launch that
carries the lattice traffic in the timed graphs (k_rod_lbm: the filament solve
with the previous step's LBM pass; algorithmic bytes / HIP-event duration),
the standalone LBM launch's and the whole step's, the CPU restatement's rate on a
bounded sample of the same workload, and the batched ES evaluation rate: a
seeded population of ragged instances (grid sides 128/192/256, 1-16
filaments of 16-64 nodes) rolled out together by mc_batch_rollout, against
the same rollouts run one instance at a time on the plain GPU path.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time
from dataclasses import replace
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "coupled-step MLUPS (% HBM peak); batched evals/s + p99 latency vs plain"
HBM_PEAK_GBS = 8000.0
LBM_BYTES_PER_CELL = 80  # read f 36 + write f 36 + write u* 8
NODE_BYTES = 100  # per node and step: positions, velocities, debug_vector, theta, edge/hinge params


def cpu_baseline(wl, steps, runs=1):
    """The numpy FP32 restatement timed on this host: `runs` runs of `steps`
    steps, each after one warm-up step; returns (median MLUPS, list of run
    seconds)."""
    import numpy as np
    from oracle import step_ref as R
    from tests.helpers import oracle_params, oracle_state
    p = oracle_params(wl)
    secs = []
    for _ in range(runs):
        s = oracle_state(wl)
        s = R.step(replace(s, bending_rest_angles=wl.rest_angles_at(0)), wl.dt, p)
        t0 = time.perf_counter()
        for t in range(1, steps + 1):
            s = R.step(replace(s, bending_rest_angles=wl.rest_angles_at(t)), wl.dt, p)
        secs.append(time.perf_counter() - t0)
    h, w = wl.grid_shape
    return h * w * steps / float(np.median(secs)) / 1e6, secs


def batched_es(n_inst, steps, plain_sample, dev):
    """Evaluations per second of a batched population vs the plain loop."""
    import torch
    from microcosmos_amd import PBD_SCHEME, Rollout
    from microcosmos_amd import workloads as W
    from microcosmos_amd.batch import Batch
    from microcosmos_amd.workloads import device_state
    pop = W.es_population(n_inst, seed=2, steps=steps)
    cfg = replace(PBD_SCHEME, **pop[0].config_overrides)

    def act(w):
        return tuple(torch.as_tensor(a, device=dev) for a in (w.act_amp, w.act_omega, w.act_phase))

    b = Batch([device_state(w, dev) for w in pop], cfg, pop[0].dt,
              actuation=[act(w) for w in pop])
    b.autotune(steps=4)
    b.run(16)
    torch.cuda.synchronize()
    print(f"[bench] batched warmup done ({n_inst} instances)", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    disp = b.run(steps)
    torch.cuda.synchronize()
    tb = time.perf_counter() - t0
    finite = bool(torch.isfinite(disp).all().item())
    cells = sum(w.grid_shape[0] * w.grid_shape[1] for w in pop)
    # the plain per-instance loop on a sample of the same instances, stratified
    # by grid side (the first instances of each side); every sample rollout
    # gets its plan and graphs, and runs once untimed, before the clock starts
    sides = sorted({w.grid_shape[0] for w in pop})
    per = max(1, plain_sample // len(sides))
    sample = [w for s in sides for w in [x for x in pop if x.grid_shape[0] == s][:per]]
    ros = []
    for w in sample:
        n, e, f = device_state(w, dev)
        ro = Rollout(n, e, f, cfg, w.dt, actuation=(w.act_amp, w.act_omega, w.act_phase))
        tr = ro.track()
        ro.run(steps)
        tr.init(ro.position)
        ros.append(ro)
    torch.cuda.synchronize()
    side_time = {s: [] for s in sides}
    for w, ro in zip(sample, ros):
        t0 = time.perf_counter()
        ro.run(steps)
        torch.cuda.synchronize()
        side_time[w.grid_shape[0]].append(time.perf_counter() - t0)
    tp = sum(sum(v) for v in side_time.values())
    # the population's plain-loop time: each side's mean rollout time times
    # the number of instances of that side
    count = {s: sum(1 for w in pop if w.grid_shape[0] == s) for s in sides}
    t_pop = sum(count[s] * sum(side_time[s]) / len(side_time[s]) for s in sides)
    print(f"[bench] batched {tb:.3f} s, plain sample {tp:.3f} s", file=sys.stderr, flush=True)
    ev_b = n_inst / tb
    ev_p = n_inst / t_pop
    return {"workload": f"es-population-{n_inst}", "instances": n_inst, "rollout_steps": steps,
            "lattice_cells": cells, "evals_per_s": round(ev_b, 2),
            "instance_steps_per_s": round(n_inst * steps / tb, 1),
            "mlups": round(cells * steps / tb / 1e6, 1),
            "plain_loop": {"sample_instances": len(sample),
                           "sample": f"{per} instances of each side {sides}, each timed over "
                                     f"{steps} graph-replayed steps after an untimed run of the "
                                     f"same length; population time = sum over sides of "
                                     f"count x mean",
                           "mean_rollout_s": {str(s): round(sum(v) / len(v), 5)
                                              for s, v in side_time.items()},
                           "evals_per_s": round(ev_p, 2)},
            "vs_plain_loop": round(ev_b / ev_p, 2), "finite": finite,
            "p99_latency": None}


def serving(n_req, dev, lanes, chunk, max_batch, plain=True):
    """Served trace (Poisson arrivals of mixed rollout requests) through the
    continuous-batching server vs the same requests one at a time on the
    plain GPU path (per request: plan, graph-replayed rollout, displacement)."""
    import numpy as np
    import torch
    from microcosmos_amd import PBD_SCHEME, Rollout
    from microcosmos_amd import workloads as W
    from microcosmos_amd.serve import Request, Server, fifo_report, latency_report
    from microcosmos_amd.workloads import device_state
    trace = W.serving_trace(n_req, seed=4)
    first = next(it for it in trace if not it.gradient)  # the rollout requests' settings
    cfg = replace(PBD_SCHEME, **first.workload.config_overrides)
    dt = first.workload.dt

    def requests(items):
        out = []
        for it in items:
            wl = it.workload
            n, e, f = device_state(wl, dev)
            e = replace(e, pairs=wl.pairs, bending_pairs=wl.bending_pairs)
            if it.gradient:
                out.append(Request(nodes=n, edges=e, fields=f, steps=it.steps, arrival=it.arrival,
                                   gradient=True, target=torch.as_tensor(it.target, device=dev)))
                continue
            act = tuple(torch.as_tensor(a, device=dev)
                        for a in (wl.act_amp, wl.act_omega, wl.act_phase))
            out.append(Request(nodes=n, edges=e, fields=f, steps=it.steps, actuation=act,
                               arrival=it.arrival))
        return out

    warm = W.serving_trace(3, seed=99, max_filaments=4, max_steps=300)
    Server(cfg, dt, lanes=lanes, chunk=chunk, max_batch=max_batch).run(requests(warm))
    torch.cuda.synchronize()
    reqs = requests(trace)
    torch.cuda.synchronize()
    print(f"[bench] serving {n_req} requests ...", file=sys.stderr, flush=True)
    srv = Server(cfg, dt, lanes=lanes, chunk=chunk, max_batch=max_batch)
    done = srv.run(reqs, realtime=True)
    torch.cuda.synchronize()
    rep = latency_report(done)
    finite = all(bool(torch.isfinite(r.rest_angle_grad if r.gradient else r.result[0].position)
                      .all().item()) for r in done)
    cells = sum(it.workload.grid_shape[0] * it.workload.grid_shape[1] * it.steps for it in trace
                if not it.gradient)
    out = {"workload": f"serving-trace-{n_req}", "arrival_rate_per_s": 200.0,
           "grid_sides": [128, 256, 512, 1024], "filaments": [1, 256], "nodes_per_filament": [16, 64],
           "rollout_steps": [200, 2000], "lanes": lanes, "chunk_steps": chunk,
           "max_batch": max_batch, "lattice_cell_updates": cells,
           "gradient_requests": sum(it.gradient for it in trace),
           "gradient": "folding gradient (256^2, 8 x 48 nodes, 500 steps, checkpoints every 50) "
                       "through the rod adjoint kernels",
           "p50_latency_s": round(rep["p50_latency_s"], 4),
           "p99_latency_s": round(rep["p99_latency_s"], 4),
           "mean_queue_wait_s": round(rep["mean_queue_wait_s"], 4),
           "rollouts_per_s": round(rep["rollouts_per_s"], 3),
           "mlups": round(cells / rep["makespan_s"] / 1e6, 1),
           "batch_reforms": sum(l.reforms for l in srv.lanes),
           "peak_admitted_bytes": srv.peak_used, "finite": finite}
    print(f"[bench] served: {rep}", file=sys.stderr, flush=True)
    if plain:
        from microcosmos_amd import PBD_SCHEME_NO_FLUID
        from microcosmos_amd.grad import rollout_value_and_grad
        gcfg = replace(PBD_SCHEME_NO_FLUID, enable_steric=False)
        service = []
        for it in trace:
            wl = it.workload
            n, e, f = device_state(wl, dev)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            if it.gradient:
                rollout_value_and_grad(n, e, f, gcfg, dt, it.steps,
                                       torch.as_tensor(it.target, device=dev), 50)
            else:
                ro = Rollout(n, e, f, cfg, dt, actuation=(wl.act_amp, wl.act_omega, wl.act_phase))
                ro.track()
                ro.run(it.steps)
            torch.cuda.synchronize()
            service.append(time.perf_counter() - t0)
        fifo = fifo_report(np.array([it.arrival for it in trace]), np.array(service))
        print(f"[bench] one by one: {fifo}", file=sys.stderr, flush=True)
        out["plain_one_by_one"] = {k: round(v, 4) if isinstance(v, float) else v
                                   for k, v in fifo.items()}
        out["vs_plain_rollouts_per_s"] = round(rep["rollouts_per_s"] / fifo["rollouts_per_s"], 2)
        out["vs_plain_p99"] = round(fifo["p99_latency_s"] / rep["p99_latency_s"], 2)
    return out


def profiled_traffic(h, w, kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc
    summary (FETCH_SIZE + WRITE_SIZE, each scaled by its plain-copy
    calibration) of this workload, or None when the summary is absent or for
    another lattice."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        "coupled_step_1024x1024_256x64.md")
    if (h, w) != (1024, 1024) or not os.path.exists(path):
        return None
    for line in open(path):
        m = re.match(r"`%s`: ([0-9.]+) MB per launch" % re.escape(kernel), line)
        if m:
            return int(round(float(m.group(1)) * 1e6))
    return None


def gradient_leg(dev, reps=5, scheme="pbd"):
    """The folding gradient: loss of the final positions after a rollout on
    the folding workload (256^2, 8 filaments x 48 nodes) and its gradient with
    respect to the 368 hinge rest angles (384 per-node curvatures), forward
    and adjoint kernels on the device.  scheme "pbd" / "cosserat": the
    rod-only step, 500 steps at dt 1, checkpoints every 50; "folding": the
    folding experiment's default physics (stable Cosserat with the steric
    term, dt 0.02, experiments/conf/experiment/filament_folding.yaml:53-90),
    150 steps, every step kept."""
    import torch
    from microcosmos_amd import COSSERAT_SCHEME_NO_FLUID, PBD_SCHEME_NO_FLUID
    from microcosmos_amd import workloads as W
    from microcosmos_amd.grad import RodAdjoint, min_image
    from microcosmos_amd.workloads import device_state
    if scheme == "folding":
        steps, every, dt = W.FOLDING_STEPS, 1, W.FOLDING_DT
        wl, tgt = W.folding(steps=steps)
        cfg = replace(COSSERAT_SCHEME_NO_FLUID, dt=dt, **W.FOLDING_COSSERAT)
    else:
        steps, every = 500, 50
        wl, tgt = W.folding()
        dt = wl.dt
        base = COSSERAT_SCHEME_NO_FLUID if scheme == "cosserat" else PBD_SCHEME_NO_FLUID
        cfg = replace(base, enable_steric=False)
    target = torch.as_tensor(tgt, device=dev)

    def one():
        n, e, f = device_state(wl, dev)
        ad = RodAdjoint(n, e, f, cfg, dt)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        x = ad.forward(steps, every)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        d = min_image(ad.grid_shape, x, target)
        g = ad.backward(2.0 * d / x.shape[0])[0]
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        return t1 - t0, t2 - t1, float(torch.isfinite(g).all().item())

    one()
    fw, bw = [], []
    for _ in range(reps):
        a, b, ok = one()
        fw.append(a)
        bw.append(b)
    import numpy as np
    f_ms, b_ms = float(np.median(fw)) * 1e3, float(np.median(bw)) * 1e3
    return {"workload": wl.name, "scheme": scheme, "cycles": cfg.cycles_per_step, "steps": steps,
            "dt": dt, "steric": bool(cfg.enable_steric), "checkpoint_every": every,
            "parameters": 384,
            "hinges": int(wl.bending_pairs.shape[0]), "forward_ms": round(f_ms, 3),
            "backward_ms": round(b_ms, 3), "gradients_per_s": round(1e3 / (f_ms + b_ms), 2),
            "finite": bool(ok)}


def objective_leg(dev, batch=1, reps=3):
    """The folding experiment's whole loss and its gradient
    (experiments/filament_folding/experiment.py:803-923) with its default
    settings (filament_folding.yaml): the default physics, 150 + 30 steps, DCD
    centred without PCA rotation over 4 evaluation frames, the stability term over 4 frames against the main and
    anchor frames, the smooth max, and gradients with respect to the raw rest
    angles (368) and raw rest lengths (376), both starting at zero as in the
    reference; the target cloud is the folding workload's target curve (384
    points).  ``batch`` initial states (seeds 3..) share the parameters and
    go through one batched rollout and adjoint (optimization.batch_size = 4).
    One evaluation = loss + gradient of the whole batch."""
    import numpy as np
    import torch
    from microcosmos_amd import COSSERAT_SCHEME_NO_FLUID
    from microcosmos_amd import folding as F
    from microcosmos_amd import workloads as W
    from microcosmos_amd.workloads import device_state
    dt = W.FOLDING_DT
    cfg = replace(COSSERAT_SCHEME_NO_FLUID, dt=dt, **W.FOLDING_COSSERAT)
    wls = [W.folding(seed=3 + k, steps=W.FOLDING_STEPS) for k in range(batch)]
    cloud = torch.as_tensor(wls[0][1], device=dev)
    lc = F.FoldingLossConfig()
    nb = int(wls[0][0].bending_pairs.shape[0])
    ne = int(wls[0][0].pairs.shape[0])
    raw = torch.zeros(nb, dtype=torch.float32, device=dev)
    raw_l = torch.zeros(ne, dtype=torch.float32, device=dev)

    def one():
        inst = [device_state(wl, dev) for wl, _ in wls]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        if batch == 1:
            out = F.folding_value_and_grad(*inst[0], cfg, dt, W.FOLDING_STEPS, 30, raw, cloud, lc,
                                           raw_lengths=raw_l)
        else:
            out = F.batch_folding_value_and_grad(inst, cfg, dt, W.FOLDING_STEPS, 30, raw, cloud,
                                                 lc, raw_lengths=raw_l)
        torch.cuda.synchronize()
        return time.perf_counter() - t0, float(out[0].item()), \
            bool(torch.isfinite(out[3]).all().item() and torch.isfinite(out[4]).all().item())

    one()
    ts = []
    for _ in range(reps):
        t, loss, ok = one()
        ts.append(t)
    ms = float(np.median(ts)) * 1e3
    return {"workload": wls[0][0].name, "batch": batch, "steps": W.FOLDING_STEPS + 30,
            "eval_frames": lc.num_eval_frames, "stability_frames": lc.num_stability_frames,
            "loss_mode": lc.loss_mode, "pca_rotation": lc.use_pca_rotation,
            "parameters": nb + ne, "ms_per_evaluation": round(ms, 3),
            "evaluations_per_s": round(1e3 / ms, 2),
            "instance_gradients_per_s": round(batch * 1e3 / ms, 2), "loss": round(loss, 4),
            "finite": ok}


def gradient_batch_leg(dev, batch=4, reps=5):
    """The folding experiment's batched objective (experiment.py:915-923
    vmaps the loss over optimization.batch_size = 4,
    experiments/conf/experiment/filament_folding.yaml:67): ``batch`` folding
    instances (seeds 3..) with the experiment's default physics, 150 steps,
    one batched forward and one batched adjoint; gradients/s counts instance
    gradients."""
    import numpy as np
    import torch
    from microcosmos_amd import COSSERAT_SCHEME_NO_FLUID
    from microcosmos_amd import workloads as W
    from microcosmos_amd.grad import BatchRodAdjoint, min_image
    from microcosmos_amd.workloads import device_state
    steps, dt = W.FOLDING_STEPS, W.FOLDING_DT
    cfg = replace(COSSERAT_SCHEME_NO_FLUID, dt=dt, **W.FOLDING_COSSERAT)
    wls = [W.folding(seed=3 + k, steps=steps) for k in range(batch)]
    tg = [torch.as_tensor(t, device=dev) for _, t in wls]

    def one():
        ad = BatchRodAdjoint([device_state(wl, dev) for wl, _ in wls], cfg, dt)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        xs = ad.forward(steps, 1)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        bars = []
        for (wl, _), x, t in zip(wls, xs, tg):
            d = min_image(wl.grid_shape, x, t)
            bars.append(2.0 * d / x.shape[0])
        gs = ad.backward(bars)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        return t1 - t0, t2 - t1, all(bool(torch.isfinite(g).all().item()) for g in gs)

    one()
    fw, bw = [], []
    ok = True
    for _ in range(reps):
        a, b, o = one()
        fw.append(a)
        bw.append(b)
        ok = ok and o
    f_ms, b_ms = float(np.median(fw)) * 1e3, float(np.median(bw)) * 1e3
    return {"workload": f"folding-256x256-8x48 x {batch}", "scheme": "folding", "batch": batch,
            "steps": steps, "dt": dt, "forward_ms": round(f_ms, 3), "backward_ms": round(b_ms, 3),
            "gradients_per_s": round(batch * 1e3 / (f_ms + b_ms), 2), "finite": bool(ok)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--cpu-steps", type=int, default=20)
    ap.add_argument("--cpu-c0-steps", type=int, default=1000,
                    help="steps of the 128^2 plain case on the CPU restatement (0: skip)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--autotune", action="store_true",
                    help="time the plan's launch choices on the initial state first "
                         "(mc_plan_autotune); the defaults otherwise")
    ap.add_argument("--profile-steps", type=int, default=100)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch-instances", type=int, default=512)
    ap.add_argument("--batch-steps", type=int, default=500)
    ap.add_argument("--plain-sample", type=int, default=12)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=48)
    ap.add_argument("--serve-lanes", type=int, default=2)
    ap.add_argument("--serve-chunk", type=int, default=64)
    ap.add_argument("--serve-max-batch", type=int, default=32)
    ap.add_argument("--serve-no-plain", action="store_true")
    ap.add_argument("--only-serve", action="store_true",
                    help="run only the serving leg and print its JSON")
    ap.add_argument("--no-grad", action="store_true")
    ap.add_argument("--only-grad", choices=["pbd", "cosserat", "folding", "folding-batch4",
                                            "objective", "objective-batch4"],
                    help="run only the gradient leg of this scheme and print its JSON")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 9 x 1024 x 1024 float32 device copy (37.7 MB read + "
                         "37.7 MB written per launch) to calibrate HBM byte counters")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        src = torch.rand((9, 1024, 1024), device="cuda")
        dst = torch.empty_like(src)
        for _ in range(20):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibration": "copy", "bytes_read_per_launch": src.numel() * 4,
                          "bytes_written_per_launch": src.numel() * 4, "launches": 20}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not "
                                   "part of it (instances would shard by whole lattice)",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import torch
    from microcosmos_amd import PBD_SCHEME, Rollout
    from microcosmos_amd import workloads as W
    from microcosmos_amd.workloads import device_state

    dev = torch.device("cuda", 0)
    if args.only_grad:
        if args.only_grad.startswith("objective"):
            leg = objective_leg(dev, 4 if args.only_grad.endswith("batch4") else 1)
        elif args.only_grad == "folding-batch4":
            leg = gradient_batch_leg(dev, 4)
        else:
            leg = gradient_leg(dev, scheme=args.only_grad)
        print(json.dumps(leg), flush=True)
        return
    if args.only_serve:
        print(json.dumps(serving(args.serve_requests, dev, args.serve_lanes, args.serve_chunk,
                                 args.serve_max_batch, plain=not args.serve_no_plain)), flush=True)
        return
    wl = W.plain_1024()
    nodes, edges, fields = device_state(wl, dev)
    cfg = replace(PBD_SCHEME, **wl.config_overrides)
    ro = Rollout(nodes, edges, fields, cfg, wl.dt,
                 actuation=(wl.act_amp, wl.act_omega, wl.act_phase))
    use_graph = not args.no_graph
    if args.autotune:
        # the plan's launch knobs timed on a copy of the initial state
        # (results do not depend on them)
        ro.autotune(steps=32)
    from microcosmos_amd.simulate import tuning as plan_tuning
    tuning = plan_tuning(ro.plan.handle)
    ro.run(args.warmup, use_graph=use_graph)
    # one untimed run of the timed length instantiates the graphs the timed
    # run replays (its bulk graph and its remainder graph)
    ro.run(args.steps, use_graph=use_graph)
    torch.cuda.synchronize()
    ro.check()
    print(f"[bench] warmup {args.warmup} (+{args.steps} graph instantiation) done",
          file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    ro.run(args.steps, use_graph=use_graph)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ro.check()
    print(f"[bench] timed {args.steps} steps in {elapsed:.3f} s", file=sys.stderr, flush=True)
    finite = bool(torch.isfinite(ro.f_grid).all().item() and torch.isfinite(ro.position).all().item())
    h, w = wl.grid_shape
    ms = elapsed / args.steps * 1e3
    mlups = h * w * args.steps / elapsed / 1e6
    # roofline of the launch that carries the lattice traffic in the timed
    # graphs: the rod solve with the previous step's LBM pass (k_rod_lbm),
    # timed by HIP events carried by those launches over a run of fused steps
    nprof = max(2, args.profile_steps + (args.profile_steps & 1))
    fused_ms, fused_step_ms = ro.profile_fused(nprof)
    # the standalone LBM launch (the unfused step), for comparison
    lbm_ms, step_ms_ev = ro.profile(nprof)
    torch.cuda.synchronize()
    lbm_bytes = LBM_BYTES_PER_CELL * h * w
    fused_bytes = lbm_bytes + NODE_BYTES * wl.num_nodes
    achieved = fused_bytes / (fused_ms * 1e-3) / 1e9
    lbm_gbs = lbm_bytes / (lbm_ms * 1e-3) / 1e9
    # the whole step against the same roof: (80 B per cell + 100 B per node)
    # per step over the timed wall clock
    step_bytes = LBM_BYTES_PER_CELL * h * w + NODE_BYTES * wl.num_nodes
    step_gbs = step_bytes * args.steps / elapsed / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": profiled_traffic(h, w, "k_rod_lbm"),
                "traffic_source": "committed profile, not this run: "
                                  "profiles/coupled_step_1024x1024_256x64.md (rocprofv3 --pmc "
                                  "FETCH_SIZE + WRITE_SIZE, plain-copy calibrated)",
                "kernel": "k_rod_lbm", "avg_kernel_ms": round(fused_ms, 5),
                "bytes_per_launch": fused_bytes,
                "basis": "80 B/cell (LBM pass) + 100 B/node (filament solve) per launch / "
                         "HIP-event duration of the fused launch, steps 2..n of "
                         f"{nprof} fused steps",
                "fused_step_ms_events": round(fused_step_ms, 5),
                "k_lbm": {"avg_kernel_ms": round(lbm_ms, 5), "bytes_per_launch": lbm_bytes,
                          "achieved": round(lbm_gbs, 1),
                          "frac": round(lbm_gbs / HBM_PEAK_GBS, 4),
                          "traffic": profiled_traffic(h, w, "k_lbm"),
                          "basis": "the standalone LBM launch of the unfused step"},
                "step": {"bytes_per_step": step_bytes, "achieved": round(step_gbs, 1),
                         "frac": round(step_gbs / HBM_PEAK_GBS, 4),
                         "basis": "(80 B/cell + 100 B/node) x steps / timed wall clock"}}
    batched = None
    if not args.no_batch:
        batched = batched_es(args.batch_instances, args.batch_steps, args.plain_sample, dev)
    grad = None
    if not args.no_grad:
        grad = gradient_leg(dev)
        grad["stable_cosserat"] = gradient_leg(dev, scheme="cosserat")
        grad["folding_physics"] = gradient_leg(dev, scheme="folding")
        grad["folding_physics_batch4"] = gradient_batch_leg(dev, 4)
        grad["folding_objective"] = objective_leg(dev, 1)
        grad["folding_objective_batch4"] = objective_leg(dev, 4)
        print(f"[bench] gradient {grad}", file=sys.stderr, flush=True)
    served = None
    if not args.no_serve:
        served = serving(args.serve_requests, dev, args.serve_lanes, args.serve_chunk,
                         args.serve_max_batch, plain=not args.serve_no_plain)
        if batched is not None:
            batched["p99_latency"] = served["p99_latency_s"]
    cpu = None
    if not args.no_cpu:
        print("[bench] cpu baseline ...", file=sys.stderr, flush=True)
        v, secs = cpu_baseline(wl, args.cpu_steps, runs=3)
        cpu = {"value": round(v, 3), "unit": "MLUPS", "cores": 1, "kind": "port",
               "sample": f"median of 3 runs of {args.cpu_steps} steps of the same 1024^2 256x64 "
                         f"workload, each after one warm-up step, numpy FP32 restatement "
                         f"(runs {', '.join(f'{t:.2f}' for t in secs)} s)",
               "host_cpus": len(os.sched_getaffinity(0)),
               "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}
        if args.cpu_c0_steps > 0:
            wl0 = W.plain_128()
            v0, s0 = cpu_baseline(wl0, args.cpu_c0_steps, runs=1)
            cpu["plain_128"] = {"value": round(v0, 3), "unit": "MLUPS", "steps": args.cpu_c0_steps,
                                "seconds": round(s0[0], 2),
                                "coupled_steps_per_s": round(args.cpu_c0_steps / s0[0], 2),
                                "sample": "the 128^2 32-node plain case, numpy FP32 restatement"}
    out = {"metric": METRIC, "value": round(mlups, 2), "unit": "MLUPS", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 5),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic",
           "config": {"workload": wl.name, "grid": [h, w], "filaments": 256,
                      "nodes_per_filament": 64, "pbd_cycles": 8, "ibm_kernel_size": 4,
                      "ibm_iterations": 5, "steric_sigma": 2.0, "tau_plus": 0.6,
                      "lambda_trt": 0.1875, "hip_graph": use_graph, "finite": finite,
                      "tuning": tuning},
           "roofline": roofline, "cpu_baseline": cpu, "batched": batched, "serving": served,
           "gradient": grad,
           "coupled_steps_per_s": round(args.steps / elapsed, 2),
           "step_ms_events": round(step_ms_ev, 5)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
