"""Host topology builders: the directed rod-graph builder, the JSON graph
format and the initial node positions.

``GraphBuilder`` follows graph_builder.py:30-157: one directed edge per bond
(forward along a chain, parent -> child at a branch), each with a departure
angle; ``build`` lists the edges grouped by source node in the order nodes
were first added, each node's successors in the order they were attached
(the edge order of the reference's networkx DiGraph), which fixes the edge
numbering that the bending pairs and every ordered scatter of the rod solvers
depend on.

``parse_graph_json`` is the JSON format of graph_builder.py:160-194
({"nodes": [...], "edges": [[u, v], ...]}): the undirected graph it builds
(edges first, then the listed nodes; an edge given twice, in either
direction, is one bond), as its node order and its edge list in the order an
undirected networkx graph reports them.  Turning that graph into a directed
Cosserat graph and laying it out (graph.py:13-154, 548-659 and
cactus_graph_layout.py) are host preprocessing outside this library.

``initialize_positions`` follows graph.py:263-345.  The deterministic parts
(line placement, the single-instance centre, the circle and the tree's
trunk and branches) are the reference's; the centres of several instances
and the noise are drawn from a torch generator, so they match the
reference's distributions but not its jax.random draws.
"""
from __future__ import annotations

import json
import math
from pathlib import Path

import numpy as np

import networkx as nx


class GraphBuilder:
    """graph_builder.py:30-157 -- a directed rod graph assembled from loops,
    lines and branches; ``G`` is the networkx DiGraph, ``next_id`` the next
    free node id."""

    def __init__(self):
        self.G = nx.DiGraph()
        self.next_id = 0

    def _take(self, length: int) -> list[int]:
        ids = list(range(self.next_id, self.next_id + length))
        self.next_id += length
        return ids

    def add_loop(self, length: int) -> list[int]:
        """graph_builder.py:45-61 -- a closed ring: i -> i+1, last -> first,
        departure angle 0."""
        ids = self._take(length)
        for k, a in enumerate(ids):
            self.G.add_edge(a, ids[(k + 1) % length], departure_angle=0.0)
        return ids

    def add_line(self, length: int) -> list[int]:
        """graph_builder.py:63-79 -- an open chain i -> i+1, departure angle 0."""
        ids = self._take(length)
        for a, b in zip(ids[:-1], ids[1:]):
            self.G.add_edge(a, b, departure_angle=0.0)
        return ids

    def add_branch(self, from_node: int, length: int,
                   departure_angle: float | None = None) -> list[int]:
        """graph_builder.py:81-128 -- a chain of ``length`` new nodes joined
        to ``from_node`` by a junction edge.  Without an explicit angle the
        first branch off a node departs at +pi/2 and any later one at -pi/2
        (a branch is an out-edge whose angle exceeds 0.1 in magnitude); every
        edge of the branch carries the junction's angle, so the branch's own
        hinges rest straight."""
        ids = self._take(length)
        for a, b in zip(ids[:-1], ids[1:]):
            self.G.add_edge(a, b, departure_angle=0.0)
        if departure_angle is None:
            branches = sum(1 for s in self.G.successors(from_node)
                           if abs(self.G[from_node][s].get("departure_angle", 0.0)) > 0.1)
            departure_angle = math.pi / 2.0 if branches == 0 else -math.pi / 2.0
        self.G.add_edge(from_node, ids[0], departure_angle=departure_angle)
        for a, b in zip(ids[:-1], ids[1:]):
            self.G[a][b]["departure_angle"] = departure_angle
        return ids

    def build(self):
        """graph_builder.py:130-157 -- (E, 2) int32 edges and (E,) float32
        departure angles in the DiGraph's edge order; raises ValueError unless
        the nodes are exactly 0 .. n-1."""
        n = self.G.number_of_nodes()
        if set(self.G.nodes()) != set(range(n)):
            raise ValueError(f"Graph nodes must be contiguous integers from 0 to {n - 1}. "
                             f"Got nodes: {sorted(self.G.nodes())}")
        pairs, dep = [], []
        for u, v, data in self.G.edges(data=True):
            pairs.append((u, v))
            dep.append(data.get("departure_angle", 0.0))
        if not pairs:
            return np.zeros((0, 2), np.int32), np.zeros(0, np.float32)
        return np.asarray(pairs, np.int32), np.asarray(dep, np.float32)


def parse_graph_json(source):
    """graph_builder.py:185-192 -- the undirected graph of a JSON graph file
    (a path, a JSON string or an already-decoded dict).  Returns (nodes,
    edges): the node ids in the graph's order (first appearance in the edge
    list, then the listed nodes not yet seen) and the (E, 2) int32 bonds, one
    per unordered pair, each reported from the endpoint that came first in
    node order, grouped by that endpoint."""
    if isinstance(source, dict):
        data = source
    else:
        text = str(source)
        data = json.loads(text) if text.lstrip().startswith("{") else json.loads(Path(text).read_text())
    order: dict[int, None] = {}
    adj: dict[int, dict[int, None]] = {}

    def node(a: int):
        if a not in order:
            order[a] = None
            adj[a] = {}

    for e in data.get("edges", []):
        a, b = int(e[0]), int(e[1])
        node(a)
        node(b)
        adj[a][b] = None
        adj[b][a] = None
    for a in data.get("nodes", []):
        node(int(a))
    rank = {a: i for i, a in enumerate(order)}
    pairs = []
    for a in order:
        for b in adj[a]:
            if rank[b] > rank[a] or b == a:
                pairs.append((a, b))
    edges = np.asarray(pairs, np.int32) if pairs else np.zeros((0, 2), np.int32)
    return list(order), edges


def initialize_positions(num_nodes: int, graph_type: str, shape, generator=None,
                         num_instances: int = 1) -> np.ndarray:
    """graph.py:263-345 -- (num_nodes, 2) float32 initial positions of
    ``num_instances`` filaments of make_edges' sizes: a vertical line of
    length 0.8 min(shape) ('line'), a trunk of 0.8 min(shape) with two
    horizontal branches of 0.3 min(shape) from its middle node plus noise of
    0.02 min(shape) ('tree'), or a circle of radius 0.25 min(shape) plus noise
    of 5 % of the radius (anything else); one instance sits at the domain's
    centre, several at uniform centres in [0.3, 0.7) x shape.  ``generator``
    is a torch.Generator (or a seed) for the centres and the noise."""
    import torch

    if isinstance(generator, int) or generator is None:
        g = torch.Generator().manual_seed(0 if generator is None else int(generator))
    else:
        g = generator
    shape = np.asarray(shape, np.float32)
    per, rem = divmod(num_nodes, num_instances)
    if num_instances == 1:
        centers = (shape / 2.0).reshape(1, 2)
    else:
        u = torch.rand((num_instances, 2), generator=g, dtype=torch.float32).numpy()
        lo, hi = np.float32(0.3) * shape, np.float32(0.7) * shape
        centers = (lo + u * (hi - lo)).astype(np.float32)
    m = float(np.min(shape))
    out = []
    for i in range(num_instances):
        size = per + (1 if i < rem else 0)
        c = centers[i].astype(np.float32)
        if graph_type == "tree":
            trunk = int(size * 0.6)
            rest = size - trunk
            b1 = rest // 2
            b2 = rest - b1
            half = np.array([0.0, 0.4 * m], np.float32)
            tpos = np.linspace(c - half, c + half, trunk, dtype=np.float32)
            bp = tpos[trunk // 2]
            s1 = np.linspace(1, b1, b1, dtype=np.float32) / np.float32(b1 + 1)
            s2 = np.linspace(1, b2, b2, dtype=np.float32) / np.float32(b2 + 1)
            length = np.float32(0.3 * m)
            p1 = bp + s1[:, None] * np.array([1.0, 0.0], np.float32) * length
            p2 = bp + s2[:, None] * np.array([-1.0, 0.0], np.float32) * length
            pos = np.concatenate([tpos, p1, p2], axis=0).astype(np.float32)
            noise = torch.randn(pos.shape, generator=g, dtype=torch.float32).numpy()
            pos = pos + noise * np.float32(0.02 * m)
        elif graph_type == "line":
            half = np.array([0.0, m * 0.8 / 2.0], np.float32)
            pos = np.linspace(c - half, c + half, size, dtype=np.float32)
        else:
            radius = np.float32(m * 0.25)
            ang = np.linspace(0.0, 2.0 * np.pi, size, endpoint=False).astype(np.float32)
            circ = np.stack([radius * np.cos(ang), radius * np.sin(ang)], axis=1)
            noise = torch.randn(circ.shape, generator=g, dtype=torch.float32).numpy()
            pos = c + circ + noise * np.float32(radius * 0.05)
        out.append(pos.astype(np.float32))
    return np.concatenate(out, axis=0).astype(np.float32)
