"""Parity of the HIP path with the CPU restatement, stage by stage and for the
whole coupled step, on seeded inputs.  Every call goes through the C ABI.

Tolerances:
  * integer work and in-order folds: bit-exact (CIC density, kernel weights);
  * single operator calls whose only rounding difference is the FFT or a
    double-evaluated transcendental: relative L2 <= 1e-6;
  * the coupled step: relative L2 <= 1e-5 after 100 steps (the FP32 tolerance
    the path is specified to), stated per result together with the FP32-vs-FP64
    oracle drift of the same run.
"""
from __future__ import annotations

import sys
from dataclasses import replace

import numpy as np
import pytest

from oracle import step_ref as R
from microcosmos_amd import workloads as W

from .helpers import (device_config, device_state, max_ulp, oracle_params, oracle_state,
                      pos_rel_l2, rel_l2, run_oracle)

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda", 0)


def _np(t):
    return t.detach().to("cpu").numpy()


@pytest.mark.parametrize("shape,sigma", [((48, 48), 2.0), ((64, 96), 1.5), ((10, 10), 1.0),
                                         ((128, 128), 2.0), ((50, 50), 2.0),
                                         ((1024, 1024), 2.0), ((256, 128), 5.0),
                                         ((256, 256), 1.5), ((34, 51), 2.0), ((262, 131), 2.0)])
def test_gaussian_filter(dev, shape, sigma):
    """Direct filter (short kernels; also sides with prime factors above 13,
    which have no FFT plan) or FFT (sigma 1.5 on 256^2), each against its
    restatement and against numpy's float64 FFT filter."""
    _say(f"gaussian_filter shape={shape} sigma={sigma}")
    from microcosmos_amd.utils import gaussian_filter
    rng = np.random.default_rng(3)
    d = rng.random(shape).astype(np.float32)
    got = _np(gaussian_filter(torch.as_tensor(d, device=dev), sigma))
    exp64 = R.gaussian_filter_pocketfft(d.astype(np.float64), sigma)
    exp32 = R.gaussian_filter(d, sigma)
    # same algorithm (direct filter or Stockham network), same roundings
    np.testing.assert_array_equal(got, exp32)
    # FP32 error against the float64 DFT, stated
    e = rel_l2(got, exp64)
    _say(f"  fp32 vs fp64 rel-L2 {e:.2e}")
    assert e <= 1e-6


def test_fused_filter_equals_two_pass(dev, monkeypatch):
    """The one-launch 2-D filter (kernels of at most 33 taps; at 1024^2 with
    the CIC deposit inside it) and the row deposit with the row and column
    launches give the same steric field and steps, bit for bit."""
    _say("fused 2-D steric filter vs two passes, 6 steps on 64x80 and 1024x1024")
    from microcosmos_amd import Rollout
    from microcosmos_amd._runtime import Plan
    for wl in (W.small(grid=(64, 80), n_fil=4, n_nodes=24, seed=29), W.plain_1024()):
        cfg = device_config(wl)
        runs = []
        for two_pass in (False, True):
            if two_pass:
                monkeypatch.setenv("MC_FILTER_TWO_PASS", "1")
            nodes, edges, fields = device_state(wl, dev)
            ro = Rollout(nodes, edges, fields, cfg, wl.dt,
                         actuation=(wl.act_amp, wl.act_omega, wl.act_phase))
            ro.plan = Plan(ro.grid_shape, tuple(ro.f_grid.shape[1:]), ro.position.shape[0],
                           ro.plan.pairs, ro.plan.bending_pairs)
            ro.run(6)
            torch.cuda.synchronize()
            runs.append((_np(ro.steric), _np(ro.position), _np(ro.f_grid)))
            monkeypatch.delenv("MC_FILTER_TWO_PASS", raising=False)
        for a, b in zip(runs[0], runs[1]):
            np.testing.assert_array_equal(a, b)
        assert np.abs(runs[0][0]).max() > 0.0


def test_gaussian_filter_needs_fft_side(dev):
    """A kernel too long for the direct filter on a side without an FFT plan
    (262 = 2 x 131, sigma 1.5) is refused with a clear error."""
    _say("gaussian_filter 262x262 sigma 1.5: refused")
    from microcosmos_amd.utils import gaussian_filter
    assert R.gauss_taps(262, 1.5) is None
    d = torch.zeros((262, 262), device=dev)
    with pytest.raises(Exception, match="prime factors"):
        gaussian_filter(d, 1.5)
    # the same plan shape still filters with a short kernel
    d[3, 5] = 1.0
    f = _np(gaussian_filter(d, 2.0))
    np.testing.assert_array_equal(f, R.gaussian_filter(_np(d), 2.0))


def test_gaussian_filter_reference_kat(dev):
    """tests/test_utils.py:59-68: delta at (5,5) on 10x10 keeps its peak and symmetry."""
    _say("gaussian_filter KAT 10x10")
    from microcosmos_amd.utils import gaussian_filter
    d = torch.zeros((10, 10), device=dev)
    d[5, 5] = 1.0
    f = _np(gaussian_filter(d, 1.0))
    assert int(np.argmax(f)) == 55
    assert np.allclose(f[5, 4], f[5, 6]) and np.allclose(f[4, 5], f[6, 5])


@pytest.mark.parametrize("case", ["small48", "plain128"])
def test_steric_field(dev, case):
    _say(f"steric field {case}")
    from microcosmos_amd.forces import compute_field_steric_force_corrected, deposit_density
    wl = W.small() if case == "small48" else W.plain_128()
    nodes, edges, fields = device_state(wl, dev)
    dens = _np(deposit_density(nodes.position, wl.grid_shape, 1.0))
    exp_dens, exp_st, exp_f = R.compute_field_steric_force_corrected(
        wl.position, wl.grid_shape, 2.0, 5, 1.0)
    # CIC deposit: integer cells + in-order fold -> bit-exact
    assert np.array_equal(dens, exp_dens), max_ulp(dens, exp_dens)
    f2, force = compute_field_steric_force_corrected(nodes, fields, 2.0, 5, 1.0)
    _, st64, f64 = R.compute_field_steric_force_corrected(
        wl.position.astype(np.float64), wl.grid_shape, 2.0, 5, 1.0)
    assert rel_l2(_np(f2.steric), exp_st) <= 1e-7
    assert rel_l2(_np(force), exp_f) <= 1e-6, max_ulp(_np(force), exp_f)
    # the steric force is a difference of nearly equal samples: its FP32 error
    # against float64 is stated and bounded by the restatement's own
    e_gpu, e_o32 = rel_l2(_np(force), f64), rel_l2(exp_f, f64)
    _say(f"  force rel-L2 vs fp64: device {e_gpu:.2e}, fp32 restatement {e_o32:.2e}")
    assert e_gpu <= 1.5 * e_o32 + 1e-7
    assert rel_l2(_np(f2.steric), st64) <= 1e-6


@pytest.mark.parametrize("shape", [(48, 48), (48, 80)])
def test_cic_deposit_dense_rows(dev, shape):
    """Rows holding more nodes than the row pass keeps in LDS (here 300 and
    120 nodes against 64) take the global overflow path; a stack of nodes on
    one cell checks the fold order of equal keys.  Bit-exact against the
    restatement."""
    _say(f"cic deposit dense rows {shape}")
    from microcosmos_amd.forces import deposit_density
    H, W = shape
    rng = np.random.default_rng(21)
    a = np.stack([rng.uniform(0, W, 300), rng.uniform(10.0, 11.0, 300)], 1)
    b = np.stack([rng.uniform(0, W, 120), rng.uniform(30.2, 30.9, 120)], 1)
    c = np.tile([[7.25, 40.5]], (40, 1)) + rng.normal(0, 1e-3, (40, 2))
    d = rng.uniform(-W, 2 * W, (60, 2))  # wrapped on the device
    pos = np.concatenate([a, b, c, d]).astype(np.float32)
    rng.shuffle(pos)
    dens = _np(deposit_density(torch.as_tensor(pos, device=dev), shape, 0.7))
    exp_dens = R.compute_field_steric_force_corrected(pos, shape, 2.0, 5, 0.7)[0]
    assert np.array_equal(dens, exp_dens), max_ulp(dens, exp_dens)


@pytest.mark.parametrize("shape", [(128, 160), (256, 256)])
def test_fused_deposit_dense_windows(dev, shape):
    """The deposit inside the 2-D filter launch (grids wider and taller than
    a tile's window, at most one node per 32 cells): a patch holding more
    CIC entries than a window keeps in LDS (600 nodes in 30 x 30 cells) takes
    the arena path; a stack of nodes on one cell checks the fold order; nodes
    across the periodic edges and on cell corners.  Density bit-exact against
    the restatement, the filtered field too (the same taps, the same order)."""
    _say(f"fused cic deposit dense windows {shape}")
    from microcosmos_amd.forces import deposit_density, update_steric_potential
    from microcosmos_amd.structs import Fields, Nodes
    H, W = shape
    rng = np.random.default_rng(23)
    a = rng.uniform(40.0, 70.0, (600, 2))
    b = np.tile([[11.5, 90.25]], (16, 1)) + rng.normal(0, 1e-3, (16, 2))
    c = np.array([[0.0, 0.0], [W - 1e-3, H - 1e-3], [W * 0.5, 0.0], [3.0, 7.0]])
    pos = np.concatenate([a, b, c]).astype(np.float32)
    rng.shuffle(pos)
    assert pos.shape[0] * 32 <= H * W
    dens = _np(deposit_density(torch.as_tensor(pos, device=dev), shape, 0.7))
    exp_dens, exp_st, _ = R.compute_field_steric_force_corrected(pos, shape, 2.0, 5, 0.7)
    assert np.array_equal(dens, exp_dens), max_ulp(dens, exp_dens)
    n = pos.shape[0]
    z = torch.zeros((n, 2), device=dev)
    nodes = Nodes(position=torch.as_tensor(pos, device=dev), velocity=z,
                  color_bend=torch.zeros(n, dtype=torch.int32, device=dev), debug_vector=z)
    fields = Fields(grid_shape=shape, steric=torch.zeros(shape, device=dev),
                    fluid_velocity=torch.zeros((2, H, W), device=dev), f_grid=None)
    st = _np(update_steric_potential(nodes, fields, sigma=2.0, scatter_value=0.7).steric)
    np.testing.assert_array_equal(st, exp_st)


@pytest.mark.parametrize("solver", ["pbd", "cosserat"])
def test_rod_solve(dev, solver):
    _say(f"rod solve {solver}")
    from microcosmos_amd._runtime import rod_iterations
    wl = W.small(n_fil=3, n_nodes=20, seed=11)
    nodes, edges, fields = device_state(wl, dev)
    rng = np.random.default_rng(5)
    y = wl.position + rng.normal(0, 0.2, wl.position.shape).astype(np.float32)
    rest = (rng.normal(0, 0.3, wl.bending_rest_angles.shape)).astype(np.float32)
    base = R.COSSERAT_SCHEME if solver == "cosserat" else R.PBD_SCHEME
    p = replace(base, cycles_per_step=6)
    s = oracle_state(wl)
    s = replace(s, debug_vector=y, bending_rest_angles=rest)
    fn = R.stable_cosserat_constraint if solver == "cosserat" else R.pbd_rod_constraint
    for _ in range(6):
        fn(s, p)
    from microcosmos_amd import COSSERAT_SCHEME, PBD_SCHEME
    cfg = COSSERAT_SCHEME if solver == "cosserat" else PBD_SCHEME
    nodes = replace(nodes, debug_vector=torch.as_tensor(y, device=dev))
    edges = replace(edges, bending_rest_angles=torch.as_tensor(rest, device=dev))
    n2, e2 = rod_iterations((nodes, edges), cfg, wl.grid_shape, 6, solver)
    gp, gt = _np(n2.position), _np(e2.theta)
    assert max_ulp(gp, s.position) <= 2, max_ulp(gp, s.position)
    assert max_ulp(gt, s.theta) <= 2, max_ulp(gt, s.theta)


def _ib_case(dev, wl, K, seed=21):
    from microcosmos_amd.solver.fluid import immersed_boundary_interaction
    rng = np.random.default_rng(seed)
    vel = rng.normal(0, 0.05, wl.position.shape).astype(np.float32)
    wl.velocity = vel
    p = oracle_params(wl, ibm_kernel_size=K)
    v_exp, u_exp, f_exp = R.immersed_boundary_interaction(wl.position, vel, wl.grid_shape,
                                                          wl.f_grid, 1.0, p)
    nodes, edges, fields = device_state(wl, dev)
    cfg = device_config(wl, ibm_kernel_size=K)
    n2, fl2, f2 = immersed_boundary_interaction(nodes, edges, fields, fields.f_grid, 1.0, cfg)
    gv, gu, gf = _np(n2.velocity), _np(fl2.fluid_velocity), _np(f2)
    assert max_ulp(gf, f_exp) <= 4, max_ulp(gf, f_exp)
    assert max_ulp(gu, u_exp) <= 4, max_ulp(gu, u_exp)
    assert max_ulp(gv, v_exp) <= 4, max_ulp(gv, v_exp)


@pytest.mark.parametrize("K", [1, 5, 6])
def test_ib_lbm_kernel_sizes(dev, K):
    _say(f"IB+LBM single call K={K}")
    _ib_case(dev, W.small(grid=(40, 44), n_fil=3, n_nodes=20, seed=17), K)


def test_ib_lbm_crowded_cluster(dev):
    # 480 nodes on 40^2: the stencils chain every filament into one cluster
    # whose working set exceeds the LDS (but below the wide-solve size), so
    # the per-cluster solve runs from global memory
    _say("IB+LBM single call, one 480-node cluster")
    _ib_case(dev, W.small(grid=(40, 40), n_fil=12, n_nodes=40, seed=3), 4)


@pytest.mark.parametrize("K", [4, 6])
def test_ib_lbm_wide_cluster(dev, K):
    # 1920 nodes on 64^2 chain into one cluster of more than 768 nodes: it
    # is solved by the wide workgroups with grid barriers
    _say(f"IB+LBM single call, one 1920-node cluster (wide solve), K={K}")
    _ib_case(dev, W.small(grid=(64, 64), n_fil=48, n_nodes=40, seed=3), K)


def _branch(wl, extra):
    """The workload with extra edges (pairs of node ids) appended: nodes of
    degree 3 and more, whose synthetic-point tangents sum several edges in
    scatter order; filaments joined by an edge become one component."""
    from dataclasses import replace as _r
    extra = np.asarray(extra, np.int32).reshape(-1, 2)
    k = len(extra)
    return _r(wl, pairs=np.concatenate([wl.pairs, extra]),
              theta=np.concatenate([wl.theta, np.zeros(k, wl.theta.dtype)]),
              rest_lengths=np.concatenate([wl.rest_lengths,
                                           np.ones(k, wl.rest_lengths.dtype)]))


def _ib_syn_case(dev, wl, K, width, seed=21):
    """One IB + LBM call with synthetic points against the restatement: the
    points (tangent sums, normal, offsets), the IB over 2N points and the
    averaged drag, bit for bit."""
    from microcosmos_amd.solver.fluid import immersed_boundary_interaction
    rng = np.random.default_rng(seed)
    vel = rng.normal(0, 0.05, wl.position.shape).astype(np.float32)
    wl.velocity = vel
    p = oracle_params(wl, ibm_kernel_size=K, synthetic_node_width=width)
    v_exp, u_exp, f_exp = R.immersed_boundary_interaction(wl.position, vel, wl.grid_shape,
                                                          wl.f_grid, 1.0, p, pairs=wl.pairs)
    nodes, edges, fields = device_state(wl, dev)
    cfg = device_config(wl, ibm_kernel_size=K, synthetic_node_width=width)
    n2, fl2, f2 = immersed_boundary_interaction(nodes, edges, fields, fields.f_grid, 1.0, cfg)
    np.testing.assert_array_equal(_np(f2), f_exp)
    np.testing.assert_array_equal(_np(fl2.fluid_velocity), u_exp)
    np.testing.assert_array_equal(_np(n2.velocity), v_exp)


@pytest.mark.parametrize("K,width", [(4, 0.5), (3, 1.25), (6, 0.3)])
def test_ib_lbm_synthetic_points(dev, K, width):
    _say(f"IB+LBM single call, synthetic points K={K} width={width}")
    _ib_syn_case(dev, W.small(grid=(48, 48), n_fil=3, n_nodes=24, seed=13), K, width)


def test_ib_lbm_synthetic_branched(dev):
    # nodes of degree 3 and 4, two filaments joined, singleton components
    _say("IB+LBM single call, synthetic points on a branched graph")
    wl = W.small(grid=(40, 44), seed=5, node_counts=[1, 3, 1, 12, 2, 1, 14])
    wl = _branch(wl, [[9, 4], [9, 13], [20, 9], [6, 22]])
    _ib_syn_case(dev, wl, 4, 0.7)


def test_ib_lbm_synthetic_wide_cluster(dev):
    # 1920 nodes = 3840 points on 64^2: one cluster above the wide-solve size
    _say("IB+LBM single call, synthetic points, one 3840-point cluster (wide solve)")
    _ib_syn_case(dev, W.small(grid=(64, 64), n_fil=48, n_nodes=40, seed=3), 4, 0.5)


def test_step_parity_synthetic_points(dev):
    """Twelve coupled steps with synthetic IB points (a 16-step graph's
    remainder graph, and step-by-step): bit-identical to the restatement."""
    _say("12 coupled steps with synthetic points, 48^2")
    wl = W.small(grid=(48, 48), n_fil=4, n_nodes=20, seed=9)
    p = oracle_params(wl, synthetic_node_width=0.6)
    o = run_oracle(wl, 12, p)
    from microcosmos_amd import Rollout
    for graph in (True, False):
        nodes, edges, fields = device_state(wl, dev)
        ro = Rollout(nodes, edges, fields, device_config(wl, synthetic_node_width=0.6), wl.dt,
                     actuation=(wl.act_amp, wl.act_omega, wl.act_phase))
        ro.run(12, use_graph=graph)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(_np(ro.position), o.position)
        np.testing.assert_array_equal(_np(ro.velocity), o.velocity)
        np.testing.assert_array_equal(_np(ro.f_grid), o.f_grid)
        np.testing.assert_array_equal(_np(ro.fluid_velocity), o.fluid_velocity)


def test_step_parity_wide_cluster(dev):
    """Ten coupled steps of a crowded 64^2 case whose filaments form one wide
    cluster: bit-identical to the float32 restatement."""
    _say("10 coupled steps, 64^2 with 48 x 40 nodes (wide cluster)")
    wl = W.small(grid=(64, 64), n_fil=48, n_nodes=40, seed=3)
    p = oracle_params(wl)
    o = run_oracle(wl, 10, p)
    from microcosmos_amd import Rollout
    nodes, edges, fields = device_state(wl, dev)
    ro = Rollout(nodes, edges, fields, device_config(wl), wl.dt,
                 actuation=(wl.act_amp, wl.act_omega, wl.act_phase))
    ro.run(10, use_graph=True)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(_np(ro.position), o.position)
    np.testing.assert_array_equal(_np(ro.f_grid), o.f_grid)
    np.testing.assert_array_equal(_np(ro.fluid_velocity), o.fluid_velocity)
    np.testing.assert_array_equal(_np(ro.velocity), o.velocity)


def test_ib_lbm_singletons(dev):
    # filaments of 1 and 2 nodes (components without edges or hinges)
    _say("IB+LBM single call, singleton components")
    wl = W.small(grid=(32, 36), seed=5, node_counts=[1, 3, 1, 12, 2, 1])
    _ib_case(dev, wl, 4)


@pytest.mark.parametrize("K", [4, 2, 3])
def test_ib_lbm_single_call(dev, K):
    _say(f"IB+LBM single call K={K}")
    from microcosmos_amd.solver.fluid import immersed_boundary_interaction
    wl = W.small(grid=(48, 48), n_fil=3, n_nodes=24, seed=13)
    rng = np.random.default_rng(21)
    vel = rng.normal(0, 0.05, wl.position.shape).astype(np.float32)
    wl.velocity = vel
    p = oracle_params(wl, ibm_kernel_size=K)
    tr = R.IBTrace()
    v_exp, u_exp, f_exp = R.immersed_boundary_interaction(wl.position, vel, wl.grid_shape,
                                                          wl.f_grid, 1.0, p, trace=tr)
    nodes, edges, fields = device_state(wl, dev)
    cfg = device_config(wl, ibm_kernel_size=K)
    n2, fl2, f2 = immersed_boundary_interaction(nodes, edges, fields, fields.f_grid, 1.0, cfg)
    gv, gu, gf = _np(n2.velocity), _np(fl2.fluid_velocity), _np(f2)
    assert max_ulp(gf, f_exp) <= 4, max_ulp(gf, f_exp)
    assert max_ulp(gu, u_exp) <= 4, max_ulp(gu, u_exp)
    assert max_ulp(gv, v_exp) <= 4, max_ulp(gv, v_exp)


def _compare_states(tag, g, o, o64):
    gpos = _np(g[0].position)
    gu = _np(g[2].fluid_velocity)
    gf = _np(g[2].f_grid)
    gth = _np(g[1].theta)
    r = {
        "pos": pos_rel_l2(o.grid_shape, gpos, o.position),
        "u": rel_l2(gu, o.fluid_velocity),
        "f": rel_l2(gf, o.f_grid),
        "theta": rel_l2(gth, o.theta),
        "steric": rel_l2(_np(g[2].steric), o.steric),
        "rho": rel_l2(gf.astype(np.float64).sum(axis=0), o.f_grid.astype(np.float64).sum(axis=0)),
    }
    r64 = {"pos": pos_rel_l2(o.grid_shape, gpos, o64.position),
           "u": rel_l2(gu, o64.fluid_velocity), "f": rel_l2(gf, o64.f_grid)} if o64 else None
    d64 = {"pos": pos_rel_l2(o.grid_shape, o.position, o64.position),
           "u": rel_l2(o.fluid_velocity, o64.fluid_velocity),
           "f": rel_l2(o.f_grid, o64.f_grid)} if o64 else None
    _say(f"{tag}: gpu-vs-fp32 {r}  gpu-vs-fp64 {r64}  fp32-vs-fp64 {d64}")
    return r


@pytest.mark.parametrize("steps", [1, 10])
def test_step_parity_small(dev, steps):
    _say(f"step parity small steps={steps}")
    from microcosmos_amd import step
    wl = W.small()
    p = oracle_params(wl)
    o = run_oracle(wl, steps, p)
    o64 = run_oracle(wl, steps, p, dtype=np.float64)
    nodes, edges, fields = device_state(wl, dev)
    cfg = device_config(wl)
    for t in range(steps):
        edges = replace(edges, bending_rest_angles=torch.as_tensor(wl.rest_angles_at(t), device=dev))
        nodes, edges, fields = step(nodes, edges, fields, wl.dt, cfg)
    r = _compare_states(f"small {steps} steps", (nodes, edges, fields), o, o64)
    for k in ("pos", "u", "f", "theta"):
        assert r[k] <= 1e-5, (k, r)


@pytest.mark.parametrize("bc,st", [("cosserat", "pbd"), ("pbd", "cosserat"), ("cosserat", "cosserat")])
def test_step_follows_bending_constraint(dev, bc, st):
    """The kernel follows bending_constraint; solver_type only sets the
    inertial target (a Cosserat constraint under solver_type "pbd" pulls toward
    y = 0), five steps with the fluid against the restatement."""
    _say(f"step dispatch bending_constraint={bc} solver_type={st}")
    from microcosmos_amd import step, pbd_rod_constraint, stable_cosserat_constraint
    wl = W.small(seed=11)
    base = R.COSSERAT_SCHEME if bc == "cosserat" else R.PBD_SCHEME
    p = replace(oracle_params(wl, base, cycles_per_step=4), solver_type=st, bending_constraint=bc)
    o = run_oracle(wl, 5, p)
    from microcosmos_amd import COSSERAT_SCHEME, PBD_SCHEME
    cfg = device_config(wl, COSSERAT_SCHEME if bc == "cosserat" else PBD_SCHEME, cycles_per_step=4)
    cfg = replace(cfg, solver_type=st, bending_constraint=(
        stable_cosserat_constraint if bc == "cosserat" else pbd_rod_constraint))
    nodes, edges, fields = device_state(wl, dev)
    for t in range(5):
        edges = replace(edges, bending_rest_angles=torch.as_tensor(wl.rest_angles_at(t), device=dev))
        nodes, edges, fields = step(nodes, edges, fields, wl.dt, cfg)
    r = _compare_states(f"dispatch {bc}/{st}", (nodes, edges, fields), o, None)
    for k in ("pos", "u", "f", "theta"):
        assert r[k] <= 1e-5, (k, r)
    dbg = _np(nodes.debug_vector)
    if st == "cosserat":
        assert rel_l2(dbg, o.debug_vector) <= 1e-5
    else:
        assert np.all(dbg == 0)


def test_rollout_permuted_node_ids(dev):
    """Filaments whose node ids are not contiguous ranges (the small case with
    its node numbering shuffled) take the listed-id paths of every kernel;
    12 graph-replayed steps (one 8-step graph and 4 single steps) against
    the restatement on the same numbering."""
    _say("rollout with shuffled node ids, 12 steps")
    from microcosmos_amd import Rollout
    wl = W.small()
    rng = np.random.default_rng(8)
    order = rng.permutation(wl.num_nodes)  # new node k is old node order[k]
    new_id = np.argsort(order).astype(np.int32)
    wl = replace(wl, position=wl.position[order].copy(), velocity=wl.velocity[order].copy(),
                 pairs=new_id[wl.pairs],
                 color_bend=None if wl.color_bend is None else wl.color_bend[order].copy())
    p = oracle_params(wl)
    o = run_oracle(wl, 12, p)
    nodes, edges, fields = device_state(wl, dev)
    ro = Rollout(nodes, edges, fields, device_config(wl), wl.dt,
                 actuation=(wl.act_amp, wl.act_omega, wl.act_phase))
    ro.run(12)
    torch.cuda.synchronize()
    r = _compare_states("shuffled ids 12 steps", (ro.nodes(), ro.edges_now(), ro.fields_now()), o,
                        None)
    for k in ("pos", "u", "f", "theta", "steric"):
        assert r[k] <= 1e-5, (k, r)


@pytest.mark.parametrize("use_graph", [True, False])
def test_rollout_parity_100(dev, use_graph):
    _say(f"rollout parity 128^2 100 steps graph={use_graph}")
    from microcosmos_amd import Rollout
    wl = W.plain_128()
    p = oracle_params(wl)
    o = run_oracle(wl, 100, p)
    o64 = run_oracle(wl, 100, p, dtype=np.float64)
    nodes, edges, fields = device_state(wl, dev)
    cfg = device_config(wl)
    ro = Rollout(nodes, edges, fields, cfg, wl.dt,
                 actuation=(wl.act_amp, wl.act_omega, wl.act_phase))
    ro.run(100, use_graph=use_graph)
    torch.cuda.synchronize()
    r = _compare_states("plain128 100 steps", (ro.nodes(), ro.edges_now(), ro.fields_now()), o, o64)
    for k in ("pos", "u", "f", "theta"):
        assert r[k] <= 1e-5, (k, r)


def test_rollout_parity_1000(dev):
    """1000 steps of the 128^2 plain case: density, velocity, positions within
    relative L2 1e-3 of the float32 restatement (and 1e-5 at step 100 on the
    way, from the same run)."""
    _say("rollout parity 128^2 1000 steps")
    from microcosmos_amd import Rollout
    wl = W.plain_128()
    p = oracle_params(wl)
    nodes, edges, fields = device_state(wl, dev)
    cfg = device_config(wl)
    ro = Rollout(nodes, edges, fields, cfg, wl.dt,
                 actuation=(wl.act_amp, wl.act_omega, wl.act_phase))
    s = oracle_state(wl)
    s64 = oracle_state(wl, np.float64)
    done = 0
    for stop, tol in ((100, 1e-5), (1000, 1e-3)):
        ro.run(stop - done)
        torch.cuda.synchronize()
        for t in range(done, stop):
            s = R.step(replace(s, bending_rest_angles=wl.rest_angles_at(t)), wl.dt, p)
            s64 = R.step(replace(s64, bending_rest_angles=wl.rest_angles_at(t).astype(np.float64)),
                         wl.dt, p)
        done = stop
        r = _compare_states(f"plain128 {stop} steps", (ro.nodes(), ro.edges_now(), ro.fields_now()),
                            s, s64)
        for k in ("pos", "u", "rho", "f"):
            assert r[k] <= tol, (stop, k, r)


def test_plain_1024_two_steps(dev):
    """The full 1024^2, 256 x 64 workload against the float32 restatement over
    two coupled steps (graph-replayed rollout)."""
    _say("plain 1024^2 two steps vs oracle")
    from microcosmos_amd import Rollout
    wl = W.plain_1024()
    p = oracle_params(wl)
    o = run_oracle(wl, 2, p)
    nodes, edges, fields = device_state(wl, dev)
    ro = Rollout(nodes, edges, fields, device_config(wl), wl.dt,
                 actuation=(wl.act_amp, wl.act_omega, wl.act_phase))
    ro.run(2)
    torch.cuda.synchronize()
    r = _compare_states("plain1024 2 steps", (ro.nodes(), ro.edges_now(), ro.fields_now()), o, None)
    for k in ("pos", "u", "rho", "f", "theta", "steric"):
        assert r[k] <= 1e-5, (k, r)


def test_plain_1024_deterministic(dev):
    """Two 50-step rollouts of the 1024^2 workload from the same state give
    bit-identical lattices, velocities and node states: no result depends on
    the order in which workgroups or atomics run."""
    _say("plain 1024^2 determinism 2 x 50 steps")
    from microcosmos_amd import Rollout
    wl = W.plain_1024()
    outs = []
    for _ in range(2):
        nodes, edges, fields = device_state(wl, dev)
        ro = Rollout(nodes, edges, fields, device_config(wl), wl.dt,
                     actuation=(wl.act_amp, wl.act_omega, wl.act_phase))
        ro.run(50)
        torch.cuda.synchronize()
        fl = ro.fields_now()
        outs.append([_np(x).copy() for x in (ro.nodes().position, ro.nodes().velocity,
                                             ro.edges_now().theta, fl.f_grid, fl.fluid_velocity,
                                             fl.steric)])
        del ro
    for a, b in zip(*outs):
        assert np.isfinite(a).all()
        assert np.array_equal(a.view(np.int32), b.view(np.int32))
