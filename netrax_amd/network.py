"""Rooted phylogenetic network model with NetRAX's node/edge index layout.

The index layout follows the reference's Newick conversion
(src/io/NetworkIO.cpp:114-185 and src/io/RootedNetworkParser.cpp:59-376):

* tips get clv_index = tip order of appearance in the Newick string;
* non-root inner tree nodes follow in ``collectNodes`` order
  (NetworkIO.cpp:40-58: stack DFS that visits the last child first), the
  root is the last inner tree node;
* reticulation nodes come last, in the same ``collectNodes`` order, and that
  order is also their reticulation index;
* the edge into a tip or a non-root tree node v has pmatrix_index = v;
  reticulation i owns edges ``n_tips + n_inner - 1 + 2*i`` (first parent)
  and ``+1`` (second parent).  The first parent is the parent at whose
  position the reticulation first appears in the string.

Parser defaults follow RootedNetworkParser.cpp: missing lengths read as 0,
a missing probability reads as 0.5, probabilities are clamped to
[brprob_min, brprob_max] (318-360), a top-level trifurcation is split with a
new inner node (278-296), and every length is clamped from below to
brlen_min (362-376).
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Optional

import numpy as np

# raxml-ng constants used by NetRAX options (NetraxOptions.hpp:100-103).
BRLEN_MIN = 1.0e-6
BRLEN_MAX = 100.0
BRPROB_MIN = 1.0e-6
BRPROB_MAX = 1.0 - 1.0e-6


@dataclasses.dataclass
class Network:
    """Flat arrays describing a rooted binary network.

    Edge e goes from ``edge_parent[e]`` to ``edge_child[e]``.  Reticulation
    r is node ``retic_node[r]`` with parent edges ``retic_first_edge[r]`` and
    ``retic_second_edge[r]``; ``retic_prob[r]`` is the first-parent
    probability (AnnotatedNetwork.cpp:89-100).
    """

    n_tips: int
    n_inner_tree: int  # includes the root
    n_retic: int
    root: int
    edge_parent: np.ndarray
    edge_child: np.ndarray
    brlen: np.ndarray
    retic_node: np.ndarray
    retic_first_edge: np.ndarray
    retic_second_edge: np.ndarray
    retic_prob: np.ndarray
    tip_labels: List[str]

    @property
    def n_nodes(self) -> int:
        return self.n_tips + self.n_inner_tree + self.n_retic

    @property
    def n_edges(self) -> int:
        return len(self.edge_parent)

    def children(self, v: int) -> List[int]:
        return [int(self.edge_child[e]) for e in np.nonzero(self.edge_parent == v)[0]]

    def child_edges(self, v: int) -> List[int]:
        return [int(e) for e in np.nonzero(self.edge_parent == v)[0]]

    def copy(self) -> "Network":
        return dataclasses.replace(
            self,
            edge_parent=self.edge_parent.copy(),
            edge_child=self.edge_child.copy(),
            brlen=self.brlen.copy(),
            retic_node=self.retic_node.copy(),
            retic_first_edge=self.retic_first_edge.copy(),
            retic_second_edge=self.retic_second_edge.copy(),
            retic_prob=self.retic_prob.copy(),
            tip_labels=list(self.tip_labels),
        )

    def validate(self) -> None:
        n, E = self.n_nodes, self.n_edges
        assert E == n - 1 + self.n_retic, (E, n, self.n_retic)
        indeg = np.bincount(self.edge_child, minlength=n)
        outdeg = np.bincount(self.edge_parent, minlength=n)
        assert indeg[self.root] == 0
        for v in range(n):
            if v < self.n_tips:
                assert outdeg[v] == 0 and indeg[v] == 1, v
            elif v == self.root:
                assert outdeg[v] in (1, 2), v
            elif v >= self.n_tips + self.n_inner_tree:
                assert indeg[v] == 2 and outdeg[v] == 1, v
            else:
                assert indeg[v] == 1 and outdeg[v] == 2, v
        for r in range(self.n_retic):
            assert self.edge_child[self.retic_first_edge[r]] == self.retic_node[r]
            assert self.edge_child[self.retic_second_edge[r]] == self.retic_node[r]

    def to_newick(self) -> str:
        """Extended Newick with ``:len::prob`` on reticulation edges."""
        seen = set()
        retic_of_node = {int(v): r for r, v in enumerate(self.retic_node)}

        def rec(v: int, via_edge: int) -> str:
            if v in retic_of_node:
                r = retic_of_node[v]
                name = f"#H{r}"
                prob = self.retic_prob[r] if via_edge == self.retic_first_edge[r] else 1.0 - self.retic_prob[r]
                suffix = f"{name}:{float(self.brlen[via_edge])!r}::{float(prob)!r}"
                if v in seen:
                    return suffix
                seen.add(v)
                (ce,) = self.child_edges(v)
                return "(" + rec(int(self.edge_child[ce]), ce) + ")" + suffix
            length = "" if via_edge < 0 else f":{float(self.brlen[via_edge])!r}"
            if v < self.n_tips:
                return self.tip_labels[v] + length
            parts = [rec(int(self.edge_child[ce]), ce) for ce in self.child_edges(v)]
            return "(" + ",".join(parts) + ")" + length

        return rec(self.root, -1) + ";"


# --------------------------------------------------------------------------
# Extended Newick reader (RootedNetworkParser.cpp)
# --------------------------------------------------------------------------


class _PNode:
    __slots__ = ("label", "children", "is_retic", "retic_name", "length", "first_parent",
                 "second_parent", "first_len", "second_len", "first_prob", "second_prob",
                 "tip_index", "clv_index", "retic_index")

    def __init__(self):
        self.label = ""
        self.children: List["_PNode"] = []
        self.is_retic = False
        self.retic_name = ""
        self.length = 0.0
        self.first_parent = None
        self.second_parent = None
        self.first_len = 0.0
        self.second_len = 0.0
        self.first_prob = 0.0
        self.second_prob = 0.0
        self.tip_index = -1
        self.clv_index = -1
        self.retic_index = -1


def _tolerant_float(s: str) -> float:
    s = s.strip()
    if not s:
        return 0.0
    return float(s)


def _read_len_support_prob(s: str):
    # RootedNetworkParser.cpp:59-83: defaults {0, 0, 0.5}
    res = [0.0, 0.0, 0.5]
    parts = s.split(":")
    if len(parts) > 4:
        raise ValueError(f"too many ':' in {s!r}")
    for i, p in enumerate(parts[1:]):
        res[i] = _tolerant_float(p)
    return res


def _split_top(s: str) -> List[str]:
    out, depth, last = [], 0, 0
    for i, ch in enumerate(s):
        if ch == "(":
            depth += 1
        elif ch == ")":
            depth -= 1
        elif ch == "," and depth == 0:
            out.append(s[last:i])
            last = i + 1
    out.append(s[last:])
    return out


def parse_newick(newick: str, brlen_min: float = BRLEN_MIN, brprob_min: float = BRPROB_MIN,
                 brprob_max: float = BRPROB_MAX) -> Network:
    newick = "".join(newick.split())
    semi = newick.find(";")
    if semi < 0:
        raise ValueError("No semicolon found! The string is: " + newick)
    body = newick[:semi]
    nodes: List[_PNode] = []
    retics = {}
    counters = {"tips": 0, "retic": 0}

    def build(label_str: str, parent: Optional[_PNode]) -> _PNode:
        if "#" in label_str:
            hpos = label_str.find("#")
            cpos = label_str.find(":")
            name = label_str[hpos + 1:] if cpos < 0 else label_str[hpos + 1:cpos]
            lsp = _read_len_support_prob(label_str)
            if name in retics:
                node = retics[name]
                node.second_parent = parent
                node.second_len, node.second_prob = lsp[0], lsp[2]
                return node
            node = _PNode()
            node.is_retic = True
            node.retic_name = name
            node.label = label_str[:hpos]
            node.first_parent = parent
            node.first_len, node.first_prob = lsp[0], lsp[2]
            counters["retic"] += 1
            retics[name] = node
            nodes.append(node)
            return node
        node = _PNode()
        cpos = label_str.find(":")
        node.label = label_str if cpos < 0 else label_str[:cpos]
        node.length = _read_len_support_prob(label_str)[0]
        nodes.append(node)
        return node

    def read_subtree(parent: Optional[_PNode], s: str) -> _PNode:
        lp, rp = s.find("("), s.rfind(")")
        if lp >= 0 and rp >= 0:
            node = build(s[rp + 1:], parent)
            for cs in _split_top(s[lp + 1:rp]):
                child = read_subtree(node, cs)
                node.children.append(child)
            return node
        if lp < 0 and rp < 0:
            node = build(s, parent)
            if not node.is_retic:
                node.tip_index = counters["tips"]
                counters["tips"] += 1
            return node
        raise ValueError("unbalanced ()'s")

    root = read_subtree(None, body)

    # probability post-processing (RootedNetworkParser.cpp:318-360)
    for nd in nodes:
        if nd.is_retic:
            if nd.second_parent is None:
                raise ValueError(f"reticulation {nd.retic_name} has a single parent")
            if nd.first_prob == 0 and nd.second_prob == 0:
                nd.first_prob = nd.second_prob = 0.5
            else:
                nd.first_prob = min(max(nd.first_prob, brprob_min), brprob_max)
                nd.second_prob = min(max(nd.second_prob, brprob_min), brprob_max)
                if abs(1.0 - (nd.first_prob + nd.second_prob)) >= 1e-3:
                    raise ValueError("invalid sum of reticulation probs in input network")

    # enforceToplevelBifurcation (RootedNetworkParser.cpp:278-296)
    if len(root.children) == 3:
        nn = _PNode()
        nn.children = [root.children[1], root.children[2]]
        root.children = [root.children[0], nn]
        nodes.append(nn)
    # A single-child root (an extra pair of parentheses around the whole
    # network, as in the celine*.nw fixtures) is kept as a unary root.
    if len(root.children) not in (1, 2):
        raise ValueError("root must have one or two children")

    # clamp lengths (RootedNetworkParser.cpp:362-376)
    for nd in nodes:
        nd.length = max(nd.length, brlen_min)
        if nd.is_retic:
            nd.first_len = max(nd.first_len, brlen_min)
            nd.second_len = max(nd.second_len, brlen_min)

    # collectNodes (NetworkIO.cpp:40-58)
    order, stack, visited = [], [root], set()
    while stack:
        a = stack.pop()
        if id(a) in visited:
            continue
        visited.add(id(a))
        order.append(a)
        for c in a.children:
            if id(c) not in visited:
                stack.append(c)
    tips = [a for a in order if not a.children]
    inner = [a for a in order if a.children and not a.is_retic and a is not root] + [root]
    rets = [a for a in order if a.children and a.is_retic]
    for a in tips:
        if a.tip_index < 0:
            raise ValueError("reticulation without children")
    n_tips, n_inner, n_ret = len(tips), len(inner), len(rets)
    if sorted(a.tip_index for a in tips) != list(range(n_tips)):
        raise ValueError("tip numbering mismatch")
    for a in tips:
        a.clv_index = a.tip_index
    for i, a in enumerate(inner):
        a.clv_index = n_tips + i
    for i, a in enumerate(rets):
        a.clv_index = n_tips + n_inner + i
        a.retic_index = i

    E = n_tips + n_inner - 1 + 2 * n_ret
    ep = np.full(E, -1, np.int32)
    ec = np.full(E, -1, np.int32)
    bl = np.zeros(E, np.float64)
    parent_of = {}
    for a in order:
        for c in a.children:
            if not c.is_retic:
                parent_of[id(c)] = a
    for a in tips + inner:
        if a is root:
            continue
        e = a.clv_index
        ep[e] = parent_of[id(a)].clv_index
        ec[e] = a.clv_index
        bl[e] = a.length
    base = n_tips + n_inner - 1
    rn = np.zeros(n_ret, np.int32)
    rf = np.zeros(n_ret, np.int32)
    rs = np.zeros(n_ret, np.int32)
    rp = np.zeros(n_ret, np.float64)
    for i, a in enumerate(rets):
        e1, e2 = base + 2 * i, base + 2 * i + 1
        ep[e1], ec[e1], bl[e1] = a.first_parent.clv_index, a.clv_index, a.first_len
        ep[e2], ec[e2], bl[e2] = a.second_parent.clv_index, a.clv_index, a.second_len
        rn[i], rf[i], rs[i], rp[i] = a.clv_index, e1, e2, a.first_prob
    labels = [""] * n_tips
    for a in tips:
        labels[a.clv_index] = a.label
    net = Network(n_tips, n_inner, n_ret, root.clv_index, ep, ec, bl, rn, rf, rs, rp, labels)
    net.validate()
    return net


# --------------------------------------------------------------------------
# Displayed trees (reference semantics, used by harness and generators)
# --------------------------------------------------------------------------


def active_edges(net: Network, config: int) -> np.ndarray:
    """Edge activity for tree config bits (bit r set = second parent,
    ReticulationConfigHelper.cpp:188-203)."""
    act = np.ones(net.n_edges, bool)
    for r in range(net.n_retic):
        if (config >> r) & 1:
            act[net.retic_first_edge[r]] = False
        else:
            act[net.retic_second_edge[r]] = False
    return act


def topo_order(net: Network) -> List[int]:
    """Parents before children."""
    indeg = np.bincount(net.edge_child, minlength=net.n_nodes)
    order, stack = [], [net.root]
    remaining = indeg.copy()
    while stack:
        v = stack.pop()
        order.append(v)
        for e in net.child_edges(v):
            c = int(net.edge_child[e])
            remaining[c] -= 1
            if remaining[c] == 0:
                stack.append(c)
    assert len(order) == net.n_nodes
    return order


def is_ancestor(net: Network, a: int, b: int) -> bool:
    """True if a == b or a reaches b by directed edges."""
    stack, seen = [a], set()
    while stack:
        v = stack.pop()
        if v == b:
            return True
        if v in seen:
            continue
        seen.add(v)
        stack.extend(net.children(v))
    return False


# --------------------------------------------------------------------------
# Seeded generators
# --------------------------------------------------------------------------


class _GNode:
    __slots__ = ("kids", "lens", "retic", "name")

    def __init__(self, name=""):
        self.kids = []  # list of child _GNode
        self.lens = []  # edge length to each child
        self.retic = None  # (first_parent, second_parent, prob) for reticulations
        self.name = name


def random_network(n_taxa: int, n_retic: int, rng: np.random.Generator,
                   mean_brlen: float = 0.1, prob_range=(0.2, 0.8)) -> Network:
    """Yule tree on ``n_taxa`` (rate 1, Exp(mean) lengths) plus ``n_retic``
    random arc insertions between two edges that are not ancestor/descendant
    of each other; reticulation first-parent prob ~ U(prob_range)."""
    if n_taxa < 2:
        raise ValueError("need at least two taxa")
    root = _GNode()
    leaves = [_GNode(), _GNode()]
    root.kids = list(leaves)
    root.lens = [rng.exponential(mean_brlen), rng.exponential(mean_brlen)]
    parent = {id(leaves[0]): root, id(leaves[1]): root}
    while len(leaves) < n_taxa:
        k = int(rng.integers(len(leaves)))
        leaf = leaves[k]
        a, b = _GNode(), _GNode()
        leaf.kids = [a, b]
        leaf.lens = [rng.exponential(mean_brlen), rng.exponential(mean_brlen)]
        parent[id(a)] = leaf
        parent[id(b)] = leaf
        leaves[k] = a
        leaves.append(b)
    for i, lf in enumerate(leaves):
        lf.name = f"T{i}"

    def all_edges():
        out, stack, seen = [], [root], set()
        while stack:
            u = stack.pop()
            if id(u) in seen:
                continue
            seen.add(id(u))
            for j, c in enumerate(u.kids):
                out.append((u, j))
                stack.append(c)
        return out

    def reaches(a, b):
        stack, seen = [a], set()
        while stack:
            v = stack.pop()
            if v is b:
                return True
            if id(v) in seen:
                continue
            seen.add(id(v))
            stack.extend(v.kids)
        return False

    for _ in range(n_retic):
        edges = all_edges()
        for _attempt in range(10000):
            i1, i2 = rng.choice(len(edges), 2, replace=False)
            (u1, j1), (u2, j2) = edges[i1], edges[i2]
            v1, v2 = u1.kids[j1], u2.kids[j2]
            if u1 is u2:
                continue
            # edge1 gets the new tree node, edge2 the reticulation; the two
            # edges must not be ancestor/descendant of each other
            if reaches(v1, u2) or reaches(v2, u1):
                continue
            break
        else:
            raise RuntimeError("could not place a reticulation")
        a = _GNode()  # splits edge (u1, v1)
        b = _GNode()  # reticulation splitting edge (u2, v2)
        l1, l2 = u1.lens[j1], u2.lens[j2]
        f1, f2 = rng.uniform(), rng.uniform()
        u1.kids[j1], u1.lens[j1] = a, l1 * f1
        a.kids, a.lens = [v1], [l1 * (1 - f1)]
        u2.kids[j2], u2.lens[j2] = b, l2 * f2
        b.kids, b.lens = [v2], [l2 * (1 - f2)]
        # a reticulation whose parent edge was split now hangs below the new node
        for new_parent, old_parent, child in ((a, u1, v1), (b, u2, v2)):
            if child.retic is not None:
                f, s_, p = child.retic
                child.retic = (new_parent if f is old_parent else f,
                               new_parent if s_ is old_parent else s_, p)
        a.kids.append(b)
        a.lens.append(rng.exponential(mean_brlen))
        b.retic = (u2, a, float(rng.uniform(*prob_range)))

    # emit extended Newick, then parse it so the index layout is canonical
    names = {}
    emitted = set()

    def rec(v, via_parent, length):
        lens = float(max(length, BRLEN_MIN))
        if v.retic is not None:
            if id(v) not in names:
                names[id(v)] = f"#H{len(names)}"
            first, second, p = v.retic
            prob = p if via_parent is first else 1.0 - p
            suffix = f"{names[id(v)]}:{lens!r}::{float(prob)!r}"
            if id(v) in emitted:
                return suffix
            emitted.add(id(v))
            return "(" + rec(v.kids[0], v, v.lens[0]) + ")" + suffix
        if not v.kids:
            return f"{v.name}:{lens!r}"
        inner = ",".join(rec(c, v, l) for c, l in zip(v.kids, v.lens))
        if via_parent is None:
            return "(" + inner + ")"
        return "(" + inner + f"):{lens!r}"

    # The first parent of a reticulation must be the one that appears first
    # in the string; emit, parse, then fix the probabilities by parent identity.
    nwk = rec(root, None, 0.0) + ";"
    net = parse_newick(nwk)
    return net


def tip_labels_sorted(net: Network) -> List[str]:
    return list(net.tip_labels)
