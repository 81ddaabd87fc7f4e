"""Pins the CPU oracle (oracle/nrx_oracle.c) before it is trusted as the
GPU engine's checker: against the brute-force golden vectors, closed forms,
library routines and invariants.  CPU only."""
import math

import numpy as np
import pytest
from scipy.linalg import expm
from scipy.special import gammainc, gammaincinv

from fixtures import dead_retic_sample, fixed_model, load_golden, load_sample
from netrax_amd.network import parse_newick
from netrax_amd.simulate import Model, Partition, dna_masks, make_workload
from oracle import oracle as O

GOLD = load_golden()


def _sample(name, model):
    if name == "dead_retic":
        return dead_retic_sample(model)
    return load_sample(name, model)


def _golden_model():
    m = GOLD["model"]
    return Model(4, np.array(m["subst_rates"]), np.array(m["freqs"]), m["alpha"], m["ncat"])


@pytest.mark.parametrize("name", sorted(GOLD["cases"]))
def test_oracle_matches_bruteforce_lnl(name):
    case = GOLD["cases"][name]
    net, parts = _sample(name, _golden_model())
    assert [int(x) for x in net.edge_parent] == case["edge_parent"]
    assert O.num_classes(net) == case["n_classes"]
    for var, key in ((O.AVERAGE, "average"), (O.BEST, "best")):
        got = O.loglikelihood(net, parts, var)[0]
        want = case[f"lnl_{key}"]
        assert abs(got - want) <= 1e-11 * abs(want), (name, key, got, want)


@pytest.mark.parametrize("name", sorted(k for k, v in GOLD["cases"].items() if "grad_average" in v))
def test_oracle_matches_bruteforce_gradient(name):
    case = GOLD["cases"][name]
    net, parts = _sample(name, _golden_model())
    _, g, _ = O.gradient(net, parts, O.AVERAGE)
    want = np.array(case["grad_average"])
    assert np.allclose(g, want, rtol=1e-9, atol=1e-10), (g, want)


def test_gamma_rates_match_scipy():
    for alpha in (0.05, 0.3, 0.7, 1.0, 2.5, 10.0):
        for k in (2, 4, 8):
            cuts = gammaincinv(alpha, np.arange(1, k) / k)
            want = np.diff(np.concatenate([[0], gammainc(alpha + 1, cuts), [1]])) * k
            got = O.gamma_rates(alpha, k)
            assert np.allclose(got, want, rtol=1e-10, atol=1e-13), (alpha, k)
            assert abs(got.mean() - 1.0) < 1e-12


def test_pmatrix_properties():
    rng = np.random.default_rng(5)
    for S in (4, 20):
        nr = S * (S - 1) // 2
        m = Model(S, rng.gamma(1.0, 1.0, nr), rng.dirichlet(np.ones(S)), 1.0)
        Q = m.rate_matrix()
        for r, t in ((0.1, 0.01), (1.3, 0.37), (3.0, 2.0), (0.5, 50.0)):
            P, dP, d2P = O.pmatrix(m, r, t)
            assert np.allclose(P, expm(Q * r * t), atol=1e-13)
            assert np.allclose(P.sum(1), 1.0, atol=1e-13)
            pi = m.freqs
            assert np.allclose(pi[:, None] * P, (pi[:, None] * P).T, atol=1e-13)  # reversibility
            h = 1e-6
            Pp, Pm = O.pmatrix(m, r, t + h)[0], O.pmatrix(m, r, t - h)[0]
            assert np.allclose(dP, (Pp - Pm) / (2 * h), atol=1e-7)
            h2 = 1e-4
            Pp2, Pm2 = O.pmatrix(m, r, t + h2)[0], O.pmatrix(m, r, t - h2)[0]
            assert np.allclose(d2P, (Pp2 - 2 * P + Pm2) / h2**2, atol=1e-5)
        P1 = O.pmatrix(m, 1.0, 0.2)[0]
        P2 = O.pmatrix(m, 1.0, 0.5)[0]
        assert np.allclose(P1 @ P2, O.pmatrix(m, 1.0, 0.7)[0], atol=1e-13)
        assert np.allclose(O.pmatrix(m, 1.0, 0.0)[0], np.eye(S), atol=1e-15)


def _jc(ncat=1, alpha=1.0):
    return Model(4, np.ones(6), np.full(4, 0.25), alpha, ncat)


def test_jc69_two_taxon_closed_form():
    net = parse_newick("(A:0.1,B:0.25);")
    seqA, seqB = "ACGTAACC", "ACGTCAGC"
    masks = np.stack([dna_masks(seqA), dna_masks(seqB)])
    for ncat, alpha in ((1, 1.0), (4, 0.5)):
        m = _jc(ncat, alpha)
        part = Partition(m, masks, np.ones(8, np.uint32))
        rates = O.gamma_rates(alpha, ncat)
        t = 0.35
        same = np.mean([0.25 * (0.25 + 0.75 * math.exp(-4.0 / 3.0 * r * t)) for r in rates])
        diff = np.mean([0.25 * (0.25 - 0.25 * math.exp(-4.0 / 3.0 * r * t)) for r in rates])
        n_same = sum(a == b for a, b in zip(seqA, seqB))
        want = n_same * math.log(same) + (8 - n_same) * math.log(diff)
        got = O.loglikelihood(net, [part])[0]
        assert abs(got - want) < 1e-12 * abs(want)


def test_scaling_matches_independent_sites_limit():
    """300 tips on saturated branches (P = stationary rows): every site
    likelihood is prod_tips pi_x ~ 1e-180, far below 2^-256 per node, so
    the 2^256 rescaling must kick in and the lnL has a closed form."""
    rng = np.random.default_rng(11)
    n = 300
    net = __import__("netrax_amd.network", fromlist=["random_network"]).random_network(n, 0, rng)
    net.brlen[:] = 100.0
    m = Model(4, np.ones(6), np.array([0.1, 0.2, 0.3, 0.4]), 1.0, 1)
    states = rng.integers(4, size=(n, 40))
    masks = (1 << states).astype(np.uint32)
    part = Partition(m, masks, np.ones(40, np.uint32))
    want = float(np.log(m.freqs)[states].sum())
    got = O.loglikelihood(net, [part])[0]
    assert abs(got - want) < 1e-9 * abs(want), (got, want)


def test_dead_reticulation_merges_configurations():
    net, parts = dead_retic_sample(fixed_model())
    assert O.num_classes(net) == 3  # 4 configurations, one pair shares a tree
    # AVERAGE is invariant to the grouping, BEST is not: the merged class
    # carries the summed probability
    _, _, tl, lp = O.loglikelihood(net, parts, O.AVERAGE, tree_logls=True)
    assert np.isclose(np.exp(lp).sum(), 1.0)


def test_oracle_gradient_matches_finite_differences():
    net, parts = load_sample("two_reticulations", fixed_model())
    for var in (O.AVERAGE, O.BEST):
        lnl, g, gp = O.gradient(net, parts, var)
        h = 1e-6
        for e in range(net.n_edges):
            b = net.brlen.copy()
            b[e] += h
            up = O.loglikelihood(net, parts, var, brlen=b)[0]
            b[e] -= 2 * h
            dn = O.loglikelihood(net, parts, var, brlen=b)[0]
            assert abs((up - dn) / (2 * h) - g[e]) < 1e-5 * max(1.0, abs(g[e])), (var, e)
        for r in range(net.n_retic):
            p = net.retic_prob.copy()
            p[r] += h
            up = O.loglikelihood(net, parts, var, probs=p)[0]
            p[r] -= 2 * h
            dn = O.loglikelihood(net, parts, var, probs=p)[0]
            assert abs((up - dn) / (2 * h) - gp[r]) < 1e-5 * max(1.0, abs(gp[r])), (var, r)


def test_reference_derivative_normalisation():
    """Reference mode divides by the mass B of trees containing the branch,
    the true derivative by the full mass A + B: d1_true = d1_ref * B/(A+B);
    both agree for branches present in every displayed tree."""
    net, parts = load_sample("two_reticulations", fixed_model())
    lnl, _, tl, lp = O.loglikelihood(net, parts, O.AVERAGE, tree_logls=True)
    for e in range(net.n_edges):
        d1r, d2r, _, _ = O.brlen_derivatives(net, parts, e, O.AVERAGE, mode=0)
        d1t, d2t, _, _ = O.brlen_derivatives(net, parts, e, O.AVERAGE, mode=1)
        lnl_t, g, _ = O.gradient(net, parts, O.AVERAGE)
        assert abs(d1t - g[e]) < 1e-9 * max(1, abs(g[e]))
        # mass of trees whose alive edge set contains e
        import bruteforce as BF
        cls = BF.classes(net)
        has = []
        for edges, _root, _c, prob in cls:
            has.append(e in edges)
        # classes come in the same first-configuration order in both codes
        w = np.exp(lp + tl[0] - (lp + tl[0]).max())
        frac = w[np.array(has)].sum() / w.sum()
        assert abs(d1t - d1r * frac) < 1e-8 * max(1.0, abs(d1t)), (e, d1t, d1r, frac)


def test_second_derivative_matches_fd():
    net, parts = load_sample("small", fixed_model())
    h = 1e-5
    for e in range(net.n_edges):
        d1, d2, _, _ = O.brlen_derivatives(net, parts, e, O.AVERAGE, mode=1)
        b = net.brlen.copy()
        b[e] += h
        gp = O.gradient(net, parts, O.AVERAGE, brlen=b)[1][e]
        b[e] -= 2 * h
        gm = O.gradient(net, parts, O.AVERAGE, brlen=b)[1][e]
        assert abs((gp - gm) / (2 * h) - d2) < 1e-4 * max(1.0, abs(d2)), (e, d2, (gp - gm) / (2 * h))


def test_pattern_weights_are_multiplicities():
    """Weights scale each displayed tree's lnL; the mixture is taken over
    whole-partition tree lnLs, so the total does not simply double."""
    w = make_workload("w", 10, 2, [300], seed=3, compress=False)
    net, part = w.network, w.partitions[0]
    tot, _, tl, lp = O.loglikelihood(net, [part], tree_logls=True)
    doubled = Partition(part.model, part.masks, part.weights * 2)
    tot2, _, tl2, _ = O.loglikelihood(net, [doubled], tree_logls=True)
    assert np.allclose(tl2, 2 * tl, rtol=1e-12)
    m = (lp + 2 * tl[0]).max()
    assert abs(tot2 - (m + np.log(np.exp(lp + 2 * tl[0] - m).sum()))) < 1e-9 * abs(tot2)
    assert abs(tot2 - 2 * tot) > 1e-6  # reticulations make the mixture non-linear


def test_min_interesting_tree_logprob_drops_unlikely_trees():
    net, parts = load_sample("two_reticulations", fixed_model())
    probs = np.array([1e-4, 0.5])
    full = O.loglikelihood(net, parts, O.AVERAGE, probs=probs, min_tree_logprob=-1e300)[0]
    cut = O.loglikelihood(net, parts, O.AVERAGE, probs=probs, min_tree_logprob=math.log(1e-3))[0]
    assert cut < full
