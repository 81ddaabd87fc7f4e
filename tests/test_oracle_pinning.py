"""Pins the oracle's reference-normalised branch derivatives and its
20-state arithmetic against checks the mathematics fixes, independent of
the oracle's own formulas.  CPU only.

* Reference mode (``orc_brlen_derivatives`` mode 0) restates
  computePartitionLhData (src/likelihood/LikelihoodDerivatives.cpp:30-188):
  AVERAGE gives the first and second derivative of
  ``f(b) = log sum_{t containing e} P(t) exp(lnL_t(b))`` -- a function of the
  per-tree lnLs only -- so it is checked against finite differences of f
  built from brute-force tree lnLs (tests/bruteforce.py: sums over all
  internal state assignments, scipy expm); BEST returns the derivatives of
  the single tree maximising ``tree_logl * tree_prob`` (:161-167), checked
  against finite differences of that tree's brute-force lnL on a fixture
  where this pick differs from the log-space maximum.
* 20-state lnL on the oracle itself: a two-taxon tree against scipy's
  matrix exponential, the F81-type 20-state closed form, and saturated
  branches (every site independent of the topology).
"""
import math

import numpy as np
import pytest
from scipy.linalg import expm

import bruteforce as BF
from fixtures import fixed_model, load_sample
from netrax_amd.network import parse_newick, random_network
from netrax_amd.simulate import Model, Partition, aa_masks
from oracle import oracle as O


def _tree_logls(net, part, brlen, classes, which):
    """lnL of the displayed trees `which` (indices into classes) by
    textbook pruning (numpy + scipy expm)."""
    out = {}
    for k in which:
        edges, root, _cfgs, _prob = classes[k]
        out[k] = BF.tree_logl_pruning(net, part, edges, root, brlen)
    return out


@pytest.mark.parametrize("name", ["two_reticulations", "reticulation_in_reticulation"])
def test_pruning_helper_matches_bruteforce_sum(name):
    """The fast per-tree helper used below equals the sum over all internal
    state assignments."""
    net, parts = load_sample(name, fixed_model())
    part = parts[0]
    w = part.weights.astype(float)
    for edges, root, _c, _p in BF.classes(net):
        L, _ = BF.tree_site_lik(net, part, edges, root, net.brlen)
        want = float(np.dot(w, np.log(L)))
        got = BF.tree_logl_pruning(net, part, edges, root, net.brlen)
        assert abs(got - want) <= 1e-12 * abs(want), (got, want)


def _fd_derivs(fun, b, h):
    """First and second derivative of fun at b by 5-point central stencils."""
    f = {k: fun(b + k * h) for k in (-2, -1, 0, 1, 2)}
    d1 = (f[-2] - 8 * f[-1] + 8 * f[1] - f[2]) / (12 * h)
    d2 = (-f[-2] + 16 * f[-1] - 30 * f[0] + 16 * f[1] - f[2]) / (12 * h * h)
    return d1, d2


def _used_classes_with_edge(net, classes, e, min_logprob=math.log(1e-6)):
    return [k for k, (edges, _r, _c, prob) in enumerate(classes)
            if math.log(prob) >= min_logprob and e in edges]


@pytest.mark.parametrize("name", ["two_reticulations", "three_reticulations", "interleaved_reticulations"])
def test_reference_mode_average_derivatives_match_fd_of_tree_mixture(name):
    net, parts = load_sample(name, fixed_model())
    part = parts[0]
    classes = BF.classes(net)
    logp = {k: math.log(c[3]) for k, c in enumerate(classes)}
    checked = 0
    for e in range(net.n_edges):
        members = _used_classes_with_edge(net, classes, e)
        if not members:
            continue
        b0 = float(net.brlen[e])
        h = min(1e-3, b0 / 4)

        def f(x):
            b = net.brlen.copy()
            b[e] = x
            ll = _tree_logls(net, part, b, classes, members)
            v = np.array([logp[k] + ll[k] for k in members])
            m = v.max()
            return m + math.log(np.exp(v - m).sum())

        fd1, fd2 = _fd_derivs(f, b0, h)
        d1, d2, p1, p2 = O.brlen_derivatives(net, parts, e, O.AVERAGE, mode=0)
        assert abs(d1 - fd1) <= 1e-7 * max(1.0, abs(fd1)), (name, e, d1, fd1)
        assert abs(d2 - fd2) <= 2e-5 * max(1.0, abs(fd2)), (name, e, d2, fd2)
        assert p1[0] == d1 and p2[0] == d2
        checked += 1
    assert checked >= net.n_edges - 2


def test_reference_mode_average_differs_from_true_derivative():
    """The reference normalisation is not the true derivative for a branch
    missing from some displayed trees (so the test above would catch an
    oracle that returned the true one)."""
    net, parts = load_sample("two_reticulations", fixed_model())
    classes = BF.classes(net)
    partial = [e for e in range(net.n_edges)
               if 0 < len(_used_classes_with_edge(net, classes, e)) < len(classes)]
    assert partial
    diffs = []
    for e in partial:
        d1r = O.brlen_derivatives(net, parts, e, O.AVERAGE, mode=0)[0]
        d1t = O.brlen_derivatives(net, parts, e, O.AVERAGE, mode=1)[0]
        diffs.append(abs(d1r - d1t) / max(1.0, abs(d1t)))
    assert max(diffs) > 1e-3


@pytest.mark.parametrize("name,probs", [("two_reticulations", [0.9, 0.85]),
                                        ("three_reticulations", [0.8, 0.3, 0.9])])
def test_reference_mode_best_picks_tree_logl_times_prob(name, probs):
    net, parts = load_sample(name, fixed_model())
    net.retic_prob = np.array(probs, float)
    part = parts[0]
    classes = BF.classes(net)
    logp = {k: math.log(c[3]) for k, c in enumerate(classes)}
    differs = 0
    for e in range(net.n_edges):
        members = _used_classes_with_edge(net, classes, e)
        if not members:
            continue
        ll = _tree_logls(net, part, net.brlen, classes, members)
        # LikelihoodDerivatives.cpp:161-167: first tree with the largest
        # tree_logl * tree_prob (strict >, members in class order)
        pick = members[0]
        for k in members[1:]:
            if ll[k] * math.exp(logp[k]) > ll[pick] * math.exp(logp[pick]):
                pick = k
        log_pick = max(members, key=lambda k: logp[k] + ll[k])
        differs += pick != log_pick
        b0 = float(net.brlen[e])
        h = min(1e-3, b0 / 4)

        def f(x):
            b = net.brlen.copy()
            b[e] = x
            return _tree_logls(net, part, b, classes, [pick])[pick]

        fd1, fd2 = _fd_derivs(f, b0, h)
        d1, d2, _, _ = O.brlen_derivatives(net, parts, e, O.BEST, mode=0)
        assert abs(d1 - fd1) <= 1e-7 * max(1.0, abs(fd1)), (name, e, pick, d1, fd1)
        assert abs(d2 - fd2) <= 2e-5 * max(1.0, abs(fd2)), (name, e, pick, d2, fd2)
    # the fixture separates the reference's pick from the log-space maximum
    assert differs >= 2


def _aa_model(rng, ncat=4, alpha=0.8):
    return Model(20, rng.gamma(1.0, 1.0, 190), rng.dirichlet(np.ones(20)), alpha, ncat)


def _aa_masks(rng, n_tips, n_sites):
    chars = np.array(list("ARNDCQEGHILKMFPSTWYV-XBZ"))
    p = np.r_[np.full(20, 0.9 / 20), np.full(4, 0.1 / 4)]
    seqs = ["".join(rng.choice(chars, size=n_sites, p=p)) for _ in range(n_tips)]
    return np.stack([aa_masks(s) for s in seqs])


def test_protein_two_taxon_matches_scipy_expm():
    rng = np.random.default_rng(21)
    net = parse_newick("(A:0.13,B:0.42);")
    masks = _aa_masks(rng, 2, 60)
    for ncat, alpha in ((1, 1.0), (4, 0.6)):
        m = _aa_model(rng, ncat, alpha)
        w = rng.integers(1, 4, size=60).astype(np.uint32)
        part = Partition(m, masks, w)
        Q = BF.rate_matrix(m)
        rates = BF.gamma_rates(alpha, ncat)
        bits = ((masks[:, :, None] >> np.arange(20)) & 1).astype(float)  # [2][sites][20]
        L = np.zeros(60)
        e0, e1 = (0, 1) if net.edge_child[0] == 0 else (1, 0)  # edges to tips A and B
        for r in rates:
            PA, PB = expm(Q * r * net.brlen[e0]), expm(Q * r * net.brlen[e1])
            L += (m.freqs[None, :] * (bits[0] @ PA.T) * (bits[1] @ PB.T)).sum(1) / ncat
        want = float(np.dot(w, np.log(L)))
        got = O.loglikelihood(net, [part])[0]
        assert abs(got - want) <= 1e-11 * abs(want), (ncat, got, want)


def test_protein_f81_closed_form():
    """Equal exchangeabilities: P(t) = e^{-mu t} I + (1 - e^{-mu t}) 1 pi^T
    with mu = 1 / (1 - sum pi^2) after normalising the mean rate to 1."""
    rng = np.random.default_rng(22)
    pi = rng.dirichlet(np.ones(20))
    m = Model(20, np.ones(190), pi, 1.0, 1)
    mu = 1.0 / (1.0 - float(np.dot(pi, pi)))
    net = parse_newick("(A:0.07,B:0.31);")
    masks = _aa_masks(rng, 2, 40)
    part = Partition(m, masks, np.ones(40, np.uint32))
    bits = ((masks[:, :, None] >> np.arange(20)) & 1).astype(float)

    def P(t):
        a = math.exp(-mu * t)
        return a * np.eye(20) + (1 - a) * np.outer(np.ones(20), pi)

    e0, e1 = (0, 1) if net.edge_child[0] == 0 else (1, 0)
    L = (pi[None, :] * (bits[0] @ P(net.brlen[e0]).T) * (bits[1] @ P(net.brlen[e1]).T)).sum(1)
    want = float(np.log(L).sum())
    got = O.loglikelihood(net, [part])[0]
    assert abs(got - want) <= 1e-12 * abs(want), (got, want)


@pytest.mark.parametrize("variant", [O.AVERAGE, O.BEST])
def test_protein_saturated_network_is_independent_sites(variant):
    """Saturated branches make every site prod_tips (pi . mask): the same
    for every displayed tree, so AVERAGE gives it exactly and BEST adds log
    of the most probable tree's probability.  120 tips force the 2^256
    rescaling at 20 states."""
    rng = np.random.default_rng(23)
    n = 120
    net = random_network(n, 2, rng)
    net.brlen[:] = 5000.0
    m = Model(20, np.ones(190), rng.dirichlet(np.ones(20)), 1.0, 1)
    masks = _aa_masks(rng, n, 25)
    part = Partition(m, masks, np.ones(25, np.uint32))
    bits = ((masks[:, :, None] >> np.arange(20)) & 1).astype(float)
    ll = float(np.log(bits @ m.freqs).sum())
    if variant == O.BEST:
        ll += max(math.log(c[3]) for c in BF.classes(net))
    got = O.loglikelihood(net, [part], variant)[0]
    assert abs(got - ll) <= 1e-9 * abs(ll), (got, ll)
