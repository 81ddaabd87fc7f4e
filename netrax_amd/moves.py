"""Seeded topology moves producing candidate networks of one search step.

The move types follow src/moves/MoveType.hpp:9-17 (arc insertion, arc
removal, rSPR); the implementations here are small stand-alone versions that
only need to produce valid rooted binary networks for batch scoring -- the
reference's move engine (perform/undo with index remapping) is out of scope.
Each candidate is re-read through the Newick reader so its node and edge
indices follow NetRAX's layout.
"""
from __future__ import annotations

from typing import Dict, List, Optional, Tuple

import numpy as np

from .network import BRLEN_MIN, Network, parse_newick


class _Mut:
    """Mutable DAG: children lists, edge lengths and reticulation-edge probs."""

    def __init__(self, net: Network):
        self.kids: Dict[int, List[int]] = {v: [] for v in range(net.n_nodes)}
        self.pars: Dict[int, List[int]] = {v: [] for v in range(net.n_nodes)}
        self.blen: Dict[Tuple[int, int], float] = {}
        self.prob: Dict[Tuple[int, int], float] = {}
        for e in range(net.n_edges):
            u, v = int(net.edge_parent[e]), int(net.edge_child[e])
            self.kids[u].append(v)
            self.pars[v].append(u)
            self.blen[(u, v)] = float(net.brlen[e])
        for r in range(net.n_retic):
            e1, e2 = int(net.retic_first_edge[r]), int(net.retic_second_edge[r])
            p = float(net.retic_prob[r])
            self.prob[(int(net.edge_parent[e1]), int(net.edge_child[e1]))] = p
            self.prob[(int(net.edge_parent[e2]), int(net.edge_child[e2]))] = 1.0 - p
        self.labels = {v: net.tip_labels[v] for v in range(net.n_tips)}
        self.root = net.root
        self.next_id = net.n_nodes

    def new_node(self) -> int:
        v = self.next_id
        self.next_id += 1
        self.kids[v] = []
        self.pars[v] = []
        return v

    def edges(self):
        return [(u, v) for u in self.kids for v in self.kids[u]]

    def reaches(self, a: int, b: int) -> bool:
        stack, seen = [a], set()
        while stack:
            x = stack.pop()
            if x == b:
                return True
            if x in seen:
                continue
            seen.add(x)
            stack.extend(self.kids[x])
        return False

    def add_edge(self, u, v, length, prob=None):
        self.kids[u].append(v)
        self.pars[v].append(u)
        self.blen[(u, v)] = max(length, BRLEN_MIN)
        if prob is not None:
            self.prob[(u, v)] = prob

    def remove_edge(self, u, v):
        self.kids[u].remove(v)
        self.pars[v].remove(u)
        length = self.blen.pop((u, v))
        prob = self.prob.pop((u, v), None)
        return length, prob

    def split(self, u, v, frac) -> int:
        """Inserts a node on edge (u, v); the reticulation prob stays on the
        edge entering v."""
        length, prob = self.remove_edge(u, v)
        a = self.new_node()
        self.add_edge(u, a, length * frac)
        self.add_edge(a, v, length * (1.0 - frac), prob)
        return a

    def suppress(self, x) -> bool:
        """Removes a node with one parent and one child, joining the edges."""
        if len(self.pars[x]) != 1 or len(self.kids[x]) != 1:
            return False
        (w,), (z,) = self.pars[x], self.kids[x]
        if z in self.kids[w]:
            return False  # would create a parallel edge
        l1, _ = self.remove_edge(w, x)
        l2, prob = self.remove_edge(x, z)
        self.add_edge(w, z, l1 + l2, prob)
        del self.kids[x], self.pars[x]
        return True

    def to_newick(self) -> str:
        names, emitted = {}, set()

        def rec(v, via, length):
            if len(self.pars[v]) == 2:
                if v not in names:
                    names[v] = f"#H{len(names)}"
                prob = self.prob[(via, v)]
                suffix = f"{names[v]}:{length!r}::{prob!r}"
                if v in emitted:
                    return suffix
                emitted.add(v)
                (c,) = self.kids[v]
                return "(" + rec(c, v, self.blen[(v, c)]) + ")" + suffix
            if not self.kids[v]:
                return f"{self.labels[v]}:{length!r}"
            inner = ",".join(rec(c, v, self.blen[(v, c)]) for c in self.kids[v])
            return "(" + inner + ")" + ("" if via is None else f":{length!r}")

        return rec(self.root, None, 0.0) + ";"


def arc_insertion(net: Network, rng: np.random.Generator, mean_brlen: float = 0.1) -> Optional[Network]:
    m = _Mut(net)
    edges = m.edges()
    for _ in range(200):
        i, j = rng.choice(len(edges), 2, replace=False)
        (u1, v1), (u2, v2) = edges[i], edges[j]
        if u1 == u2 or m.reaches(v1, u2) or m.reaches(v2, u1):
            continue
        a = m.split(u1, v1, rng.uniform())
        b = m.split(u2, v2, rng.uniform())
        p = float(rng.uniform(0.2, 0.8))
        m.prob[(u2 if (u2, b) in m.blen else m.pars[b][0], b)] = p
        m.add_edge(a, b, float(rng.exponential(mean_brlen)), 1.0 - p)
        return parse_newick(m.to_newick())
    return None


def arc_removal(net: Network, rng: np.random.Generator) -> Optional[Network]:
    if net.n_retic == 0:
        return None
    for _ in range(50):
        m = _Mut(net)
        b = int(net.retic_node[rng.integers(net.n_retic)])
        y = m.pars[b][int(rng.integers(2))]
        if len(m.kids[y]) != 2 or len(m.pars[y]) > 1:
            continue  # keep the move simple: the removed arc leaves a tree node
        m.remove_edge(y, b)
        m.prob.pop((m.pars[b][0], b), None)
        if not m.suppress(b):
            continue
        if y == m.root:
            (z,) = m.kids[y]
            if len(m.kids[z]) != 2 or len(m.pars[z]) != 1:
                continue
            m.remove_edge(y, z)
            del m.kids[y], m.pars[y]
            m.root = z
        elif not m.suppress(y):
            continue
        return parse_newick(m.to_newick())
    return None


def rspr_tail(net: Network, rng: np.random.Generator) -> Optional[Network]:
    """Prunes the subtree below a tree node and regrafts it on another edge."""
    for _ in range(200):
        m = _Mut(net)
        cand = [u for u in m.kids if len(m.kids[u]) == 2 and len(m.pars[u]) == 1]
        u = cand[int(rng.integers(len(cand)))]
        v = m.kids[u][int(rng.integers(2))]
        z = [c for c in m.kids[u] if c != v][0]
        (w,) = m.pars[u]
        if z in m.kids[w]:
            continue
        lv = m.blen[(u, v)]
        # detach u with v below it, reconnect w -> z
        l1, _ = m.remove_edge(w, u)
        l2, pz = m.remove_edge(u, z)
        m.add_edge(w, z, l1 + l2, pz)
        targets = [(x, y) for (x, y) in m.edges() if x != u and y != u and not m.reaches(v, x)]
        targets = [t for t in targets if t != (w, z)]
        if not targets:
            continue
        x, y = targets[int(rng.integers(len(targets)))]
        if v in m.kids[x]:
            continue
        length, prob = m.remove_edge(x, y)
        f = float(rng.uniform())
        m.add_edge(x, u, length * f)
        m.add_edge(u, y, length * (1 - f), prob)
        m.blen[(u, v)] = lv
        try:
            return parse_newick(m.to_newick())
        except (ValueError, AssertionError, KeyError):
            continue
    return None


def random_move(net: Network, rng: np.random.Generator, max_retic: int = 5) -> Network:
    """One seeded move: arc insertion, arc removal or rSPR."""
    for _ in range(20):
        kind = int(rng.integers(3))
        if kind == 0 and net.n_retic < max_retic:
            cand = arc_insertion(net, rng)
        elif kind == 1 and net.n_retic > 0:
            cand = arc_removal(net, rng)
        else:
            cand = rspr_tail(net, rng)
        if cand is not None:
            return cand
    return net.copy()
