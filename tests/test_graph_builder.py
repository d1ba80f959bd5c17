"""Host topology builders (microcosmos_amd.graph_builder): the reference's own
GraphBuilder tests (tests/test_graph_builder.py), the JSON graph format
against networkx's undirected graph built the reference's way
(graph_builder.py:185-192), and the initial positions' closed forms
(graph.py:263-345)."""
from __future__ import annotations

import json
import math

import numpy as np
import pytest

import networkx as nx

from microcosmos_amd.graph_builder import GraphBuilder, initialize_positions, parse_graph_json
from microcosmos_amd.topology import compute_bending_pairs


def test_simple_loop():
    b = GraphBuilder()
    loop = b.add_loop(10)
    e, d = b.build()
    assert e.shape == (10, 2) and d.shape == (10,) and e.dtype == np.int32 and d.dtype == np.float32
    assert loop == list(range(10))
    assert np.all(d == 0.0)
    assert set(e[:, 0].tolist()) == set(range(10)) == set(e[:, 1].tolist())
    assert e.tolist() == [[i, (i + 1) % 10] for i in range(10)]


def test_simple_line():
    b = GraphBuilder()
    b.add_line(15)
    e, d = b.build()
    assert e.shape == (14, 2) and np.all(d == 0.0)
    assert np.sum(e[:, 0] == 0) == 1 and np.sum(e[:, 1] == 0) == 0
    assert np.sum(e[:, 1] == 14) == 1 and np.sum(e[:, 0] == 14) == 0


def test_loop_with_three_branches():
    b = GraphBuilder()
    b.add_loop(50)
    for at in (1, 10, 20):
        b.add_branch(from_node=at, length=40)
    e, d = b.build()
    assert e.shape == (50 + 3 * 39 + 3, 2)
    for at in (1, 10, 20):
        assert np.sum(e[:, 0] == at) + np.sum(e[:, 1] == at) >= 3


def test_tree_topology_and_edge_order():
    b = GraphBuilder()
    trunk = b.add_line(40)
    bp = trunk[20]
    b1 = b.add_branch(from_node=bp, length=20)
    b2 = b.add_branch(from_node=bp, length=20)
    e, d = b.build()
    assert e.shape == (79, 2)
    assert np.sum(e[:, 0] == bp) + np.sum(e[:, 1] == bp) == 4
    assert np.sum(np.abs(d) > 0.1) == 40
    # edges grouped by source in node order; the branch node's out-edges in
    # attachment order: trunk, first branch (+pi/2), second branch (-pi/2)
    src = e[:, 0]
    assert np.all(np.diff(src) >= 0)
    out = e[src == bp]
    assert out.tolist() == [[bp, bp + 1], [bp, b1[0]], [bp, b2[0]]]
    assert d[src == bp].tolist() == [0.0, np.float32(math.pi / 2), np.float32(-math.pi / 2)]
    # the branch hinge at bp: (trunk in-edge, each out-edge) with rest angle
    # dep[out] - dep[in]
    pairs, rest = compute_bending_pairs(e, d, 80)
    ein = int(np.nonzero((e[:, 0] == bp - 1) & (e[:, 1] == bp))[0][0])
    at_bp = [(int(p[1]), float(r)) for p, r in zip(pairs, rest) if p[0] == ein]
    assert [r for _, r in at_bp] == [0.0, float(np.float32(math.pi / 2)),
                                    float(np.float32(-math.pi / 2))]


def test_complex_branching():
    b = GraphBuilder()
    trunk = b.add_line(30)
    prim = b.add_branch(from_node=trunk[15], length=20)
    b.add_branch(from_node=prim[10], length=10)
    b.add_branch(from_node=prim[10], length=10)
    e, _ = b.build()
    assert e.shape == (69, 2)
    sp = prim[10]
    assert np.sum(e[:, 0] == sp) + np.sum(e[:, 1] == sp) == 4


def test_branch_departure_angles():
    b = GraphBuilder()
    trunk = b.add_line(20)
    b.add_branch(from_node=trunk[10], length=10)
    b.add_branch(from_node=trunk[10], length=10)
    e, d = b.build()
    j = d[e[:, 0] == trunk[10]]
    nz = j[np.abs(j) > 0.1]
    assert len(nz) == 2
    assert np.isclose(abs(nz[0]), np.pi / 2, atol=0.01) and np.isclose(abs(nz[1]), np.pi / 2, atol=0.01)
    assert np.sign(nz[0]) != np.sign(nz[1])


def test_explicit_departure_angle_and_multiple_loops():
    b = GraphBuilder()
    l1, l2, l3 = b.add_loop(10), b.add_loop(15), b.add_loop(8)
    assert (l1, l2, l3) == (list(range(10)), list(range(10, 25)), list(range(25, 33)))
    br = b.add_branch(from_node=3, length=4, departure_angle=0.7)
    e, d = b.build()
    assert e.shape == (33 + 4, 2)
    for s, t in e.tolist():
        for lo, hi in ((0, 10), (10, 25), (25, 33)):
            if lo <= s < hi and s != 3:
                assert lo <= t < hi
    assert np.all(d[np.isin(e[:, 1], br)] == np.float32(0.7))


def test_empty_graph():
    e, d = GraphBuilder().build()
    assert e.shape == (0, 2) and d.shape == (0,)


def test_non_contiguous_nodes_error():
    b = GraphBuilder()
    b.G.add_edge(0, 1, departure_angle=0.0)
    b.G.add_edge(1, 5, departure_angle=0.0)
    b.next_id = 6
    with pytest.raises(ValueError, match="contiguous integers"):
        b.build()


def _nx_parse(data):
    """The reference's parse (graph_builder.py:185-192) on networkx."""
    g = nx.Graph()
    for edge in data.get("edges", []):
        g.add_edge(int(edge[0]), int(edge[1]))
    for n in data.get("nodes", []):
        g.add_node(int(n))
    return list(g.nodes()), [list(e) for e in g.edges()]


@pytest.mark.parametrize("seed", range(6))
def test_json_parse_matches_networkx_order(tmp_path, seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(3, 40))
    ids = rng.permutation(n + 5)[:n]
    edges = []
    for _ in range(int(rng.integers(1, 3 * n))):
        a, c = rng.choice(ids, 2, replace=bool(rng.random() < 0.1))
        edges.append([int(a), int(c)])
    if seed % 2:
        edges += [list(reversed(edges[0])), edges[1 % len(edges)]]  # repeated bonds
    nodes = [int(x) for x in rng.permutation(ids)] + [int(n + 7)]  # one isolated node
    data = {"nodes": nodes, "edges": edges}
    want_nodes, want_edges = _nx_parse(data)
    p = tmp_path / "g.json"
    p.write_text(json.dumps(data))
    for src in (data, str(p), json.dumps(data)):
        got_nodes, got_edges = parse_graph_json(src)
        assert got_nodes == want_nodes
        assert got_edges.dtype == np.int32 and got_edges.tolist() == want_edges


def test_json_parse_of_a_ring_feeds_the_bending_pairs():
    data = {"nodes": list(range(6)), "edges": [[i, (i + 1) % 6] for i in range(6)]}
    nodes, e = parse_graph_json(data)
    assert nodes == list(range(6))
    # the undirected ring reports (0,1) ... (4,5) then (0,5) from node 0
    assert e.tolist() == [[0, 1], [0, 5], [1, 2], [2, 3], [3, 4], [4, 5]]


def test_positions_line_closed_form():
    pos = initialize_positions(10, "line", (64, 96))
    m = 64.0
    assert pos.dtype == np.float32 and pos.shape == (10, 2)
    # shape / 2 = (32, 48): x holds shape[0] / 2, y runs along shape[1] / 2
    assert np.allclose(pos[:, 0], 32.0, atol=1e-5)
    assert np.allclose(pos[:, 1], np.linspace(48 - 0.4 * m, 48 + 0.4 * m, 10), atol=1e-4)


def test_positions_loop_and_tree_shapes():
    shape = (128, 128)
    pos = initialize_positions(64, "loop", shape, generator=3)
    r = np.hypot(pos[:, 0] - 64.0, pos[:, 1] - 64.0)
    assert abs(r.mean() - 32.0) < 1.0 and r.std() < 3.0
    pos2 = initialize_positions(64, "loop", shape, generator=3)
    assert np.array_equal(pos, pos2)
    t = initialize_positions(50, "tree", shape, generator=1)
    assert t.shape == (50, 2)
    trunk = int(50 * 0.6)
    assert np.all(np.abs(t[:trunk, 0] - 64.0) < 8.0)
    assert np.all(t[trunk:trunk + 10, 0] > 60.0) and np.all(t[trunk + 10:, 0] < 68.0)
    multi = initialize_positions(90, "line", shape, generator=5, num_instances=4)
    assert multi.shape == (90, 2)
    starts = [0, 23, 46, 68, 90]
    for i in range(4):
        seg = multi[starts[i]:starts[i + 1]]
        cx = seg[:, 0].mean()
        assert 0.3 * 128 - 1e-3 <= cx < 0.7 * 128
