"""Host topology builders (graph.py:156-401): the reference's integer KATs
(tests/test_graphs.py) on microcosmos_amd.topology, bit-exact against the
oracle's restatement, and the seeded workloads' shapes."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import step_ref as R
from microcosmos_amd import topology as T
from microcosmos_amd import workloads as W


def test_loop_edges():
    e, d = T.make_edges(10, "loop")
    assert e.shape == (10, 2) and e.dtype == np.int32
    for i in range(10):
        assert e[i, 0] == i and e[i, 1] == (i + 1) % 10
    assert np.all(d == 0)


def test_line_edges_and_degrees():
    e, _ = T.make_edges(10, "line")
    assert e.shape[0] == 9
    assert np.sum(e[:, 0] == 0) == 1 and np.sum(e[:, 1] == 0) == 0
    assert np.sum(e[:, 1] == 9) == 1 and np.sum(e[:, 0] == 9) == 0
    for i in range(1, 9):
        assert np.sum(e[:, 0] == i) == 1 and np.sum(e[:, 1] == i) == 1


def test_tree_edges_and_junction():
    e, d = T.make_edges(50, "tree")
    assert e.shape[0] == 49
    junction = int(50 * 0.6) // 2
    assert np.sum(e[:, 0] == junction) + np.sum(e[:, 1] == junction) >= 3
    assert np.sum(np.abs(d) > 0.1) >= 2
    bp, rest = T.compute_bending_pairs(e, d, 50)
    assert sum(int(e[b[0], 1]) == junction for b in bp) == 3
    assert np.sum(np.abs(rest) > 0.1) >= 2


def test_multiple_instances():
    e, _ = T.make_edges(20, "loop", num_instances=2)
    assert e.shape[0] == 20
    for s, t in e:
        assert (s < 10) == (t < 10)
    e, _ = T.make_edges(20, "line", num_instances=2)
    assert e.shape[0] == 18


def test_bending_pair_counts():
    for g, n, nb in (("line", 10, 8), ("loop", 10, 10)):
        e, d = T.make_edges(n, g)
        bp, rest = T.compute_bending_pairs(e, d, n)
        assert bp.shape == (nb, 2)
        assert np.all(rest == 0)


@pytest.mark.parametrize("g,n,k", [("loop", 17, 3), ("line", 23, 2), ("tree", 40, 1), ("tree", 41, 3)])
def test_matches_oracle(g, n, k):
    e, d = T.make_edges(n, g, k)
    e2, d2 = R.make_edges(n, g, k)
    assert np.array_equal(e, e2) and np.array_equal(d, d2)
    bp, r = T.compute_bending_pairs(e, d, n)
    bp2, r2 = R.compute_bending_pairs(e2, d2, n)
    assert np.array_equal(bp, bp2) and np.array_equal(r, r2)


def test_line_instances_equals_make_edges():
    p, bp, rest = T.line_instances([16] * 4)
    e, d = T.make_edges(64, "line", 4)
    bp2, r2 = T.compute_bending_pairs(e, d, 64)
    assert np.array_equal(p, e) and np.array_equal(bp, bp2) and np.array_equal(rest, r2)


def test_workload_shapes():
    wl = W.plain_128()
    assert wl.grid_shape == (128, 128) and wl.position.shape == (32, 2)
    assert wl.pairs.shape == (31, 2) and wl.bending_pairs.shape == (30, 2)
    assert wl.f_grid.shape == (9, 128, 128) and wl.f_grid.dtype == np.float32
    wl = W.plain_1024()
    assert wl.position.shape == (16384, 2) and wl.pairs.shape == (16128, 2)
    assert wl.bending_pairs.shape == (15872, 2)
    assert np.all((wl.position >= 0) & (wl.position < 1024))
    # feq(1, u0) has unit density to float32 rounding and reproduces u0
    rho = wl.f_grid.astype(np.float64).sum(0)
    assert np.abs(rho - 1).max() < 1e-5


def test_workload_is_seeded():
    a, b = W.small(seed=3), W.small(seed=3)
    assert np.array_equal(a.position, b.position) and np.array_equal(a.f_grid, b.f_grid)
    assert not np.array_equal(a.position, W.small(seed=4).position)
