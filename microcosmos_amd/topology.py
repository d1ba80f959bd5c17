"""Host-side topology: edge lists, bending pairs and the device rod layout.

Edge lists and bending pairs follow graph.py:156-401 exactly (integer work,
bit-exact).  ``RodLayout`` is the MI355X-specific preprocessing: it groups
nodes / edges / hinges by connected component (one workgroup per component on
the device) and turns every ``.at[].add`` scatter of the rod solvers into an
ordered gather list (CSR), so the device folds each node's and each edge's
contributions in exactly the reference's update order without atomics:

* node i   <- +F[e] for edges with src == i (edge order), then -F[e] for edges
             with tgt == i (edge order)                 (pbd.py:85-86,
             stable_cosserat.py:111-114)
* edge e   <- +corr[b] for hinges with e_in == e (hinge order), then -corr[b]
             for hinges with e_out == e (hinge order)   (pbd.py:45-46,
             stable_cosserat.py:158-160)
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


def make_edges(num_nodes: int, graph: str = "loop", num_instances: int = 1):
    """graph.py:156-192 -- concatenated per-instance directed edge list and
    per-edge departure angles, as (E,2) int32 and (E,) float32 arrays."""
    per = num_nodes // num_instances
    rem = num_nodes % num_instances
    if per < 1:
        raise ValueError(f"num_nodes ({num_nodes}) must be at least {num_instances} "
                         f"to create {num_instances} instances.")
    all_e, all_d = [], []
    start = 0
    for i in range(num_instances):
        size = per + (1 if i < rem else 0)
        e, d = _single_edges(size, graph, start)
        all_e.append(e)
        all_d.append(d)
        start += size
    return (np.concatenate(all_e, axis=0).astype(np.int32),
            np.concatenate(all_d, axis=0).astype(np.float32))


def _single_edges(n: int, graph: str, offset: int):
    """graph.py:195-259 -- loop (N edges), line (N-1) or trunk + two branches."""
    if graph == "loop":
        s = np.arange(n) + offset
        return np.stack([s, np.roll(s, -1)], axis=1), np.zeros(n, np.float32)
    if graph == "line":
        s = np.arange(n - 1) + offset
        return np.stack([s, s + 1], axis=1), np.zeros(n - 1, np.float32)
    if graph == "tree":
        trunk = int(n * 0.6)
        rest = n - trunk
        b1 = rest // 2
        b2 = rest - b1
        t0 = offset
        b1s = offset + trunk
        b2s = b1s + b1
        junction = t0 + trunk // 2
        el, dl = [], []
        for i in range(trunk - 1):
            el.append((t0 + i, t0 + i + 1)); dl.append(0.0)
        el.append((junction, b1s)); dl.append(np.pi / 2.0)
        for i in range(b1 - 1):
            el.append((b1s + i, b1s + i + 1)); dl.append(np.pi / 2.0)
        el.append((junction, b2s)); dl.append(-np.pi / 2.0)
        for i in range(b2 - 1):
            el.append((b2s + i, b2s + i + 1)); dl.append(-np.pi / 2.0)
        return np.array(el, np.int32), np.array(dl, np.float32)
    raise ValueError(f"Unknown graph type: {graph}")


def compute_bending_pairs(edge_pairs, departure_angles, num_nodes: int):
    """graph.py:347-401 -- for each node in ascending order, every (incoming,
    outgoing) edge pair in edge order; rest angle dep[e_out] - dep[e_in]."""
    ep = np.asarray(edge_pairs)
    dep = np.asarray(departure_angles, np.float32)
    incoming = [[] for _ in range(num_nodes)]
    outgoing = [[] for _ in range(num_nodes)]
    for e in range(ep.shape[0]):
        outgoing[int(ep[e, 0])].append(e)
        incoming[int(ep[e, 1])].append(e)
    bp, rest = [], []
    for v in range(num_nodes):
        for ei in incoming[v]:
            for eo in outgoing[v]:
                bp.append((ei, eo))
                rest.append(float(dep[eo] - dep[ei]))
    if not bp:
        return np.zeros((0, 2), np.int32), np.zeros(0, np.float32)
    return np.array(bp, np.int32), np.array(rest, np.float32)


def line_instances(counts):
    """Concatenated line filaments with the given node counts: (pairs, bending
    pairs, rest angles).  Equals make_edges(..., "line") + compute_bending_pairs
    when all counts are equal."""
    pairs, bps = [], []
    n0 = 0
    e0 = 0
    for n in counts:
        if n >= 2:
            s = np.arange(n - 1) + n0
            pairs.append(np.stack([s, s + 1], axis=1))
            b = np.arange(n - 2) + e0
            if n >= 3:
                bps.append(np.stack([b, b + 1], axis=1))
            e0 += n - 1
        n0 += n
    p = np.concatenate(pairs).astype(np.int32) if pairs else np.zeros((0, 2), np.int32)
    b = np.concatenate(bps).astype(np.int32) if bps else np.zeros((0, 2), np.int32)
    return p, b, np.zeros(b.shape[0], np.float32)


# --------------------------------------------------------------------------
# Device layout
# --------------------------------------------------------------------------
@dataclass
class RodLayout:
    """Component-grouped, gather-ordered topology for the device rod kernel.

    All arrays are int32.  ``*_off`` arrays have length (count + 1).
      comp_node_off / comp_nodes     global node ids of each component, ascending
      comp_edge_off / comp_edges     global edge ids of each component, ascending
      comp_hinge_off / comp_hinges   global hinge ids of each component, ascending
      edge_src / edge_tgt            component-local node index of each listed edge
      hinge_in / hinge_out           component-local edge index of each listed hinge
      node_inc_off / node_inc        per listed node: (local edge << 1) | negate
      edge_inc_off / edge_inc        per listed edge: (local hinge << 1) | negate
    """
    num_nodes: int
    num_edges: int
    num_hinges: int
    comp_node_off: np.ndarray
    comp_nodes: np.ndarray
    comp_edge_off: np.ndarray
    comp_edges: np.ndarray
    comp_hinge_off: np.ndarray
    comp_hinges: np.ndarray
    edge_src: np.ndarray
    edge_tgt: np.ndarray
    hinge_in: np.ndarray
    hinge_out: np.ndarray
    node_inc_off: np.ndarray
    node_inc: np.ndarray
    edge_inc_off: np.ndarray
    edge_inc: np.ndarray

    @property
    def num_components(self):
        return self.comp_node_off.shape[0] - 1

    def max_component(self):
        dn = np.diff(self.comp_node_off)
        de = np.diff(self.comp_edge_off)
        db = np.diff(self.comp_hinge_off)
        ni = np.diff(self.node_inc_off)
        ei = np.diff(self.edge_inc_off)
        # per-component incidence totals
        c_ni = np.add.reduceat(ni, self.comp_node_off[:-1]) if ni.size else np.zeros_like(dn)
        c_ni = np.where(dn > 0, c_ni, 0)
        c_ei = (np.add.reduceat(ei, self.comp_edge_off[:-1]) if ei.size else np.zeros_like(de))
        c_ei = np.where(de > 0, c_ei, 0)
        return (int(dn.max(initial=0)), int(de.max(initial=0)), int(db.max(initial=0)),
                int(c_ni.max(initial=0)), int(c_ei.max(initial=0)))


def _components(num_nodes, pairs):
    parent = np.arange(num_nodes)

    def find(a):
        root = a
        while parent[root] != root:
            root = parent[root]
        while parent[a] != root:
            parent[a], a = root, parent[a]
        return root

    for s, t in pairs:
        rs, rt = find(int(s)), find(int(t))
        if rs != rt:
            if rs < rt:
                parent[rt] = rs
            else:
                parent[rs] = rt
    return np.array([find(i) for i in range(num_nodes)])


def build_rod_layout(num_nodes: int, pairs, bending_pairs) -> RodLayout:
    """Group by connected component and build the ordered gather lists."""
    pairs = np.asarray(pairs, np.int64).reshape(-1, 2)
    bp = np.asarray(bending_pairs, np.int64).reshape(-1, 2)
    E = pairs.shape[0]
    B = bp.shape[0]
    if E and (pairs.min() < 0 or pairs.max() >= num_nodes):
        raise ValueError("edge endpoint out of range")
    if B and (bp.min() < 0 or bp.max() >= E):
        raise ValueError("bending pair edge index out of range")
    root = _components(num_nodes, pairs)
    # component label = rank of its smallest node id (components ordered by first node)
    uniq, comp_of_node = np.unique(root, return_inverse=True)
    C = uniq.shape[0]
    comp_of_edge = comp_of_node[pairs[:, 0]] if E else np.zeros(0, np.int64)
    comp_of_hinge = comp_of_edge[bp[:, 0]] if B else np.zeros(0, np.int64)
    if B and np.any(comp_of_edge[bp[:, 1]] != comp_of_hinge):
        # a hinge joins two edges of different components: merge is not possible
        # by construction (hinges are built from shared nodes); treat as error
        raise ValueError("bending pair spans two components")
    node_order = np.lexsort((np.arange(num_nodes), comp_of_node))
    edge_order = np.lexsort((np.arange(E), comp_of_edge)) if E else np.zeros(0, np.int64)
    hinge_order = np.lexsort((np.arange(B), comp_of_hinge)) if B else np.zeros(0, np.int64)
    cn = np.bincount(comp_of_node, minlength=C)
    ce = np.bincount(comp_of_edge, minlength=C) if E else np.zeros(C, np.int64)
    cb = np.bincount(comp_of_hinge, minlength=C) if B else np.zeros(C, np.int64)
    node_off = np.concatenate([[0], np.cumsum(cn)])
    edge_off = np.concatenate([[0], np.cumsum(ce)])
    hinge_off = np.concatenate([[0], np.cumsum(cb)])
    # local indices
    local_node = np.empty(num_nodes, np.int64)
    local_node[node_order] = np.arange(num_nodes) - node_off[comp_of_node[node_order]]
    local_edge = np.empty(E, np.int64)
    if E:
        local_edge[edge_order] = np.arange(E) - edge_off[comp_of_edge[edge_order]]
    local_hinge = np.empty(B, np.int64)
    if B:
        local_hinge[hinge_order] = np.arange(B) - hinge_off[comp_of_hinge[hinge_order]]
    edge_src = local_node[pairs[edge_order, 0]] if E else np.zeros(0, np.int64)
    edge_tgt = local_node[pairs[edge_order, 1]] if E else np.zeros(0, np.int64)
    hinge_in = local_edge[bp[hinge_order, 0]] if B else np.zeros(0, np.int64)
    hinge_out = local_edge[bp[hinge_order, 1]] if B else np.zeros(0, np.int64)
    # node incidences: src adds (edge order) then tgt adds (edge order)
    inc_lists = [[] for _ in range(num_nodes)]
    for e in range(E):
        inc_lists[int(pairs[e, 0])].append((int(local_edge[e]) << 1) | 0)
    for e in range(E):
        inc_lists[int(pairs[e, 1])].append((int(local_edge[e]) << 1) | 1)
    node_inc_off = [0]
    node_inc = []
    for v in node_order:
        node_inc.extend(inc_lists[int(v)])
        node_inc_off.append(len(node_inc))
    # edge incidences: e_in adds (hinge order) then e_out adds (hinge order)
    einc = [[] for _ in range(E)]
    for b in range(B):
        einc[int(bp[b, 0])].append((int(local_hinge[b]) << 1) | 0)
    for b in range(B):
        einc[int(bp[b, 1])].append((int(local_hinge[b]) << 1) | 1)
    edge_inc_off = [0]
    edge_inc = []
    for e in edge_order:
        edge_inc.extend(einc[int(e)])
        edge_inc_off.append(len(edge_inc))
    i32 = lambda a: np.ascontiguousarray(np.asarray(a, np.int64), dtype=np.int32)
    return RodLayout(
        num_nodes=num_nodes, num_edges=E, num_hinges=B,
        comp_node_off=i32(node_off), comp_nodes=i32(node_order),
        comp_edge_off=i32(edge_off), comp_edges=i32(edge_order),
        comp_hinge_off=i32(hinge_off), comp_hinges=i32(hinge_order),
        edge_src=i32(edge_src), edge_tgt=i32(edge_tgt),
        hinge_in=i32(hinge_in), hinge_out=i32(hinge_out),
        node_inc_off=i32(node_inc_off), node_inc=i32(node_inc),
        edge_inc_off=i32(edge_inc_off), edge_inc=i32(edge_inc))
