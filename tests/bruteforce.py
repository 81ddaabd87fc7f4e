"""Brute-force network likelihood for tiny inputs (test infrastructure).

Independent of both the oracle and the engine: every displayed tree is
scored by summing the joint probability over *all* state assignments of its
internal nodes (no pruning), with P = scipy.linalg.expm(Q r t) and discrete
Gamma rates from scipy.special.  Branch derivatives use the product rule on
the same sum (P' = rQ P).  Only for networks whose displayed trees have a
handful of internal nodes.

Semantics followed (reference file:line):
 * configuration bit r set = second parent (ReticulationConfigHelper.cpp:188-203)
 * dead nodes removed, displayed-tree root = first node with two alive
   children (NetworkFunctions.cpp:228-279)
 * root variants = configurations grouped by identical displayed tree
   (ImprovedLoglikelihood.cpp:194-408); their probability is the sum over
   members (ReticulationConfigSet.cpp:98-113)
 * AVERAGE / BEST per partition (ImprovedLoglikelihood.cpp:521-604)
"""
from __future__ import annotations

import itertools
import math

import numpy as np
from scipy.linalg import expm
from scipy.special import gammainc, gammaincinv


def gamma_rates(alpha, ncat):
    if ncat == 1:
        return np.ones(1)
    cuts = gammaincinv(alpha, np.arange(1, ncat) / ncat)
    cdf = np.concatenate([[0.0], gammainc(alpha + 1, cuts), [1.0]])
    return np.diff(cdf) * ncat


def rate_matrix(model):
    S = model.states
    Q = np.zeros((S, S))
    k = 0
    for i in range(S):
        for j in range(i + 1, S):
            Q[i, j] = model.subst_rates[k] * model.freqs[j]
            Q[j, i] = model.subst_rates[k] * model.freqs[i]
            k += 1
    np.fill_diagonal(Q, -Q.sum(1))
    return Q / -np.dot(model.freqs, np.diag(Q))


def displayed_tree(net, cfg):
    """(alive edges, dtroot) for one configuration."""
    act = np.ones(net.n_edges, bool)
    for r in range(net.n_retic):
        if (cfg >> r) & 1:
            act[net.retic_first_edge[r]] = False
        else:
            act[net.retic_second_edge[r]] = False
    alive = {}

    def is_alive(v):
        if v in alive:
            return alive[v]
        if v < net.n_tips:
            alive[v] = True
            return True
        res = False
        for e in range(net.n_edges):
            if net.edge_parent[e] == v and act[e] and is_alive(int(net.edge_child[e])):
                res = True
        alive[v] = res
        return res

    edges = frozenset(e for e in range(net.n_edges)
                      if act[e] and is_alive(int(net.edge_child[e])) and is_alive(int(net.edge_parent[e])))
    u = net.root
    while True:
        kids = [e for e in edges if net.edge_parent[e] == u]
        if len(kids) == 1:
            u = int(net.edge_child[kids[0]])
        else:
            break
    return edges, u


def _tree_edges_below(net, edges, root):
    out, stack = [], [root]
    while stack:
        v = stack.pop()
        for e in sorted(edges):
            if net.edge_parent[e] == v:
                out.append(e)
                stack.append(int(net.edge_child[e]))
    return out


def tree_site_lik(net, part, edges, root, brlen, deriv_edge=None):
    """Per-site likelihood (and d/db for ``deriv_edge``) of one displayed
    tree, summing over all internal state assignments."""
    model = part.model
    S = model.states
    Q = rate_matrix(model)
    rates = gamma_rates(model.alpha, model.ncat)
    tedges = _tree_edges_below(net, edges, root)
    internal = sorted({root} | {int(net.edge_child[e]) for e in tedges if net.edge_child[e] >= net.n_tips})
    pos = {v: i for i, v in enumerate(internal)}
    n_sites = part.masks.shape[1]
    L = np.zeros(n_sites)
    dL = np.zeros(n_sites)
    assigns = np.array(list(itertools.product(range(S), repeat=len(internal))))  # [A, nint]
    for c, r in enumerate(rates):
        Ps = {e: expm(Q * r * brlen[e]) for e in tedges}
        dPs = {e: (Q * r) @ Ps[e] for e in tedges}
        base = model.freqs[assigns[:, pos[root]]]  # [A]
        # contribution of internal->internal edges, independent of site
        inner_terms = {}
        tip_edges = []
        for e in tedges:
            u, v = int(net.edge_parent[e]), int(net.edge_child[e])
            if v >= net.n_tips:
                inner_terms[e] = (Ps[e][assigns[:, pos[u]], assigns[:, pos[v]]],
                                  dPs[e][assigns[:, pos[u]], assigns[:, pos[v]]])
            else:
                tip_edges.append(e)
        prod_inner = base.copy()
        for e, (pv, _) in inner_terms.items():
            prod_inner = prod_inner * pv
        for s in range(n_sites):
            tip_fac = {}
            for e in tip_edges:
                u, v = int(net.edge_parent[e]), int(net.edge_child[e])
                m = int(part.masks[v, s])
                allowed = [j for j in range(S) if (m >> j) & 1]
                tip_fac[e] = (Ps[e][:, allowed].sum(1)[assigns[:, pos[u]]],
                              dPs[e][:, allowed].sum(1)[assigns[:, pos[u]]])
            tot = prod_inner.copy()
            for e in tip_edges:
                tot = tot * tip_fac[e][0]
            L[s] += tot.sum() / len(rates)
            if deriv_edge is not None and deriv_edge in tedges:
                e = deriv_edge
                if e in inner_terms:
                    num = base.copy()
                    for e2, (pv, dv) in inner_terms.items():
                        num = num * (dv if e2 == e else pv)
                    for e2 in tip_edges:
                        num = num * tip_fac[e2][0]
                else:
                    num = prod_inner.copy()
                    for e2 in tip_edges:
                        num = num * (tip_fac[e2][1] if e2 == e else tip_fac[e2][0])
                dL[s] += num.sum() / len(rates)
    return L, dL


def tree_logl_pruning(net, part, edges, root, brlen):
    """lnL of one displayed tree by textbook Felsenstein pruning (numpy,
    scipy expm, no scaling: small inputs only).  Faster than
    tree_site_lik for repeated evaluation; checked against it in the tests."""
    model = part.model
    S = model.states
    Q = rate_matrix(model)
    rates = gamma_rates(model.alpha, model.ncat)
    bits = ((part.masks[:, :, None].astype(np.int64) >> np.arange(S)) & 1).astype(float)  # [tips][s][S]
    kids = {}
    for e in edges:
        kids.setdefault(int(net.edge_parent[e]), []).append(e)
    L = np.zeros(part.masks.shape[1])
    for r in rates:
        def clv(v):
            if v < net.n_tips:
                return bits[v]
            out = np.ones((part.masks.shape[1], S))
            for e in kids.get(v, []):
                out = out * (clv(int(net.edge_child[e])) @ expm(Q * r * brlen[e]).T)
            return out
        L += clv(root) @ model.freqs / len(rates)
    return float(np.dot(part.weights.astype(float), np.log(L)))


def classes(net, probs=None):
    """Root variants: list of (edge set, dtroot, member configs, prob)."""
    probs = net.retic_prob if probs is None else probs
    groups = {}
    for cfg in range(1 << net.n_retic):
        edges, root = displayed_tree(net, cfg)
        p = 1.0
        for r in range(net.n_retic):
            p *= (1 - probs[r]) if (cfg >> r) & 1 else probs[r]
        g = groups.setdefault(edges, [edges, root, [], 0.0])
        g[2].append(cfg)
        g[3] += p
    return list(groups.values())


def network_logl(net, parts, variant=0, brlen=None, min_tree_logprob=math.log(1e-6)):
    brlen = net.brlen if brlen is None else brlen
    total = 0.0
    per = []
    cls = classes(net)
    for part in parts:
        w = part.weights.astype(float)
        vals = []
        for edges, root, _cfgs, prob in cls:
            if math.log(prob) < min_tree_logprob:
                continue
            L, _ = tree_site_lik(net, part, edges, root, brlen)
            vals.append(math.log(prob) + float(np.dot(w, np.log(L))))
        if variant == 0:
            m = max(vals)
            lp = m + math.log(sum(math.exp(v - m) for v in vals))
        else:
            lp = max(vals)
        per.append(lp)
        total += lp
    return total, per


def network_gradient_brlen(net, parts, variant=0, brlen=None, min_tree_logprob=math.log(1e-6)):
    """True dlnL/db for every edge by the product rule on the brute-force sums."""
    brlen = net.brlen if brlen is None else brlen
    cls = [c for c in classes(net) if math.log(c[3]) >= min_tree_logprob]
    grad = np.zeros(net.n_edges)
    for part in parts:
        w = part.weights.astype(float)
        ll = []
        for edges, root, _cfgs, prob in cls:
            L, _ = tree_site_lik(net, part, edges, root, brlen)
            ll.append(math.log(prob) + float(np.dot(w, np.log(L))))
        ll = np.array(ll)
        if variant == 0:
            a = np.exp(ll - ll.max())
            a /= a.sum()
        else:
            a = np.zeros(len(ll))
            a[int(np.argmax(ll))] = 1.0
        for k, (edges, root, _cfgs, prob) in enumerate(cls):
            if a[k] == 0:
                continue
            for e in range(net.n_edges):
                L, dL = tree_site_lik(net, part, edges, root, brlen, deriv_edge=e)
                grad[e] += a[k] * float(np.dot(w, dL / L))
    return grad
