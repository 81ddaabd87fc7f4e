"""Reticulation-probability and model optimisation on the GPU engine
(ReticulationOptimization.cpp:68-119, ModelOptimization.cpp:25-97),
against the reference's numerical route run on the oracle:

* d lnL/dp and d2 lnL/dp2 of nrx_prob_derivatives against nrx_gradient's
  d lnL/dp and central differences of the oracle's lnL;
* nrx_optimize_reticulations against coordinate-wise bounded Brent searches
  (scipy) on the oracle's lnL, cycled as optimize_reticulations cycles:
  the engine's optimum is at least as good and is stationary;
* optimize_params against the same three stages (exchangeabilities,
  frequencies, Gamma shape) run by L-BFGS-B with finite-difference
  gradients on the oracle's lnL.
"""
import dataclasses
import sys

import numpy as np
import pytest

from fixtures import fixed_model, load_sample
from netrax_amd import AnnotatedNetwork, LikelihoodVariant, NetraxOptions, Precision
from netrax_amd import optimization as OPT
from netrax_amd.simulate import make_workload
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def eng(net, parts, variant=0, precision=Precision.FP64):
    return AnnotatedNetwork(net, parts, NetraxOptions(likelihood_variant=LikelihoodVariant(variant),
                                                      precision=precision))


def _oracle_brent_cycle(net, parts, variant, probs, max_iters=10, lo=1e-6, hi=1 - 1e-6):
    from scipy.optimize import minimize_scalar

    probs = np.array(probs, float)

    def f(r, x):
        q = probs.copy()
        q[r] = x
        return -O.loglikelihood(net, parts, variant, probs=q)[0]

    act = -f(0, probs[0])
    for _ in range(max_iters):
        for r in range(len(probs)):
            res = minimize_scalar(lambda x: f(r, x), bounds=(lo, hi), method="bounded",
                                  options=dict(xatol=1e-9))
            if -res.fun > -f(r, probs[r]):
                probs[r] = res.x
        loop = -f(0, probs[0])
        if abs(loop - act) < 1e-10:
            break
        act = loop
    return act, probs


@pytest.mark.parametrize("precision", [Precision.FP64, Precision.FP32])
def test_prob_derivatives_match_gradient_and_finite_differences(precision):
    w = make_workload("rp", 14, 3, [900], seed=21)
    net, parts = w.network, w.partitions
    ann = eng(net, parts, precision=precision)
    lnl, d1, d2 = ann.prob_derivatives()
    l2, _db, dp = ann.gradient()
    tol = 1e-9 if precision == Precision.FP64 else 1e-5
    assert np.allclose(d1, dp, rtol=tol, atol=tol * max(1.0, np.abs(dp).max()))
    assert np.all(d2 <= 0.0)
    for r in range(net.n_retic):
        h = 1e-5
        vals = []
        for sgn in (1, -1):
            q = np.array(net.retic_prob, float)
            q[r] += sgn * h
            vals.append(O.loglikelihood(net, parts, O.AVERAGE, probs=q)[0])
        fd1 = (vals[0] - vals[1]) / (2 * h)
        # FP32 tree lnLs carry ~1e-7 relative error, which moves the tree
        # weights by exp(1e-7 |lnL|)
        rt = 1e-5 if precision == Precision.FP64 else 1e-2
        assert abs(d1[r] - fd1) <= rt * max(1.0, abs(fd1)), (r, d1[r], fd1)
        # d2 from the oracle's d lnL/dp at p +- h
        gs = []
        for sgn in (1, -1):
            q = np.array(net.retic_prob, float)
            q[r] += sgn * 1e-4
            gs.append(O.gradient(net, parts, O.AVERAGE, probs=q)[2][r])
        fd2 = (gs[0] - gs[1]) / 2e-4
        if precision == Precision.FP64:
            assert abs(d2[r] - fd2) <= 1e-4 * max(1.0, abs(fd2)), (r, d2[r], fd2)
    ann.close()


@pytest.mark.parametrize("case", ["two_reticulations", "three_reticulations", "workload"])
def test_optimize_reticulations_matches_brent_on_oracle(case):
    if case == "workload":
        w = make_workload("ro", 16, 4, [1200, 600], seed=22)
        net, parts = w.network, w.partitions
    else:
        net, parts = load_sample(case, fixed_model(4))
    rng = np.random.default_rng(3)
    net.retic_prob = rng.uniform(0.05, 0.95, net.n_retic)
    start = np.array(net.retic_prob)
    say(case, "start probs", np.round(start, 4))
    ol, oprobs = _oracle_brent_cycle(net, parts, O.AVERAGE, start)
    ann = eng(net, parts)
    l0 = ann.loglikelihood()
    l1 = OPT.optimize_reticulations(ann, 10)
    probs = ann.reticulation_probs()
    say(case, "lnL", l0, "->", l1, "oracle Brent", ol, "probs", np.round(probs, 5), "oracle", np.round(oprobs, 5))
    assert l1 >= l0
    assert l1 >= ol - 1e-7 * abs(ol), (l1, ol)
    # the engine's lnL at its probabilities is the oracle's
    assert abs(O.loglikelihood(net, parts, O.AVERAGE, probs=probs)[0] - l1) <= 1e-10 * abs(l1)
    # stationary, or at a bound with the slope pointing out
    _, d1, _ = ann.prob_derivatives()
    for r in range(net.n_retic):
        if 1e-6 + 1e-6 < probs[r] < 1 - 2e-6:
            assert abs(d1[r]) <= 1e-3 * max(1.0, np.abs(d1).max()) or abs(probs[r] - oprobs[r]) < 1e-4, \
                (r, probs[r], d1[r])
        elif probs[r] <= 2e-6:
            assert d1[r] <= 0
        else:
            assert d1[r] >= 0
    np.testing.assert_allclose(net.retic_prob, probs)
    ann.close()


def test_optimize_reticulations_best_variant_does_not_lose():
    net, parts = load_sample("three_reticulations", fixed_model(6))
    net.retic_prob = np.array([0.3, 0.6, 0.45])[:net.n_retic]
    ann = eng(net, parts, variant=1)
    l0 = ann.loglikelihood()
    l1 = ann.optimize_reticulations(10)
    assert l1 >= l0
    probs = ann.reticulation_probs()
    assert abs(O.loglikelihood(net, parts, O.BEST, probs=probs)[0] - l1) <= 1e-10 * abs(l1)
    ann.close()


def _oracle_stages(net, parts):
    """The three stages with numerical gradients on the oracle's lnL."""
    from scipy.optimize import minimize

    parts = [dataclasses.replace(p) for p in parts]
    P = len(parts)

    def lnl(ms):
        ps = [dataclasses.replace(parts[i], model=ms[i]) for i in range(P)]
        return O.loglikelihood(net, ps, O.AVERAGE)[0]

    def run(x0, bounds, models_of):
        res = minimize(lambda x: -lnl(models_of(x)), x0, method="L-BFGS-B", bounds=bounds,
                       options=dict(ftol=1e7 * np.finfo(float).eps, eps=1e-7, maxiter=300))
        ms = models_of(res.x)
        for i in range(P):
            parts[i] = dataclasses.replace(parts[i], model=ms[i])
        return -res.fun

    nr = len(parts[0].model.subst_rates)
    fixed = [p.model.subst_rates[-1] for p in parts]
    base = [p.model for p in parts]
    run(np.concatenate([p.model.subst_rates[:-1] for p in parts]), [(1e-3, 1e3)] * (P * (nr - 1)),
        lambda x: [OPT._with(base[i], subst_rates=np.append(x.reshape(P, nr - 1)[i], fixed[i])) for i in range(P)])
    S = len(parts[0].model.freqs)
    base = [p.model for p in parts]
    ref = [int(np.argmax(p.model.freqs)) for p in parts]

    def fr(x, i):
        full = np.insert(x.reshape(P, S - 1)[i], ref[i], 1.0)
        return full / full.sum()

    run(np.concatenate([np.delete(p.model.freqs / p.model.freqs[ref[i]], ref[i]) for i, p in enumerate(parts)]),
        [(1e-3, 1e2)] * (P * (S - 1)), lambda x: [OPT._with(base[i], freqs=fr(x, i)) for i in range(P)])
    base = [p.model for p in parts]
    return run(np.array([p.model.alpha for p in parts]), [(0.0201, 100.0)] * P,
               lambda x: [OPT._with(base[i], alpha=x[i]) for i in range(P)])


def test_optimize_params_matches_numerical_stages_on_oracle():
    w = make_workload("mo", 12, 2, [1500, 800], seed=23)
    net, parts = w.network, w.partitions
    # start away from the generating model
    rng = np.random.default_rng(8)
    for p in parts:
        p.model = OPT._with(p.model, subst_rates=p.model.subst_rates * rng.uniform(0.5, 2.0, 6),
                            freqs=np.full(4, 0.25), alpha=1.0)
    start = [dataclasses.replace(p) for p in parts]
    ol = _oracle_stages(net, start)
    ann = eng(net, parts)
    l0 = ann.loglikelihood()
    l1 = OPT.optimize_params(ann)
    say("model optimisation lnL", l0, "->", l1, "oracle stages", ol)
    assert l1 > l0
    assert l1 >= ol - 1e-6 * abs(ol), (l1, ol)
    # the models written back score the same on the oracle
    assert abs(O.loglikelihood(net, ann.partitions, O.AVERAGE)[0] - l1) <= 1e-10 * abs(l1)
    # the last stage (the Gamma shape) ends stationary
    _, g = ann.model_gradient()
    for i, p in enumerate(ann.partitions):
        if 0.03 < p.model.alpha < 90:
            assert abs(g[i]["alpha"]) <= 5e-2, (i, g[i]["alpha"])
    ann.close()
