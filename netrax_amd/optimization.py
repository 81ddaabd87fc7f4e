"""Branch-length optimisation over the engine's branch derivatives.

Mirrors the reference's Newton-Raphson branch sweep
(src/optimization/BranchLengthOptimization.cpp):

* ``optimize_branch_newton_raphson`` (:202-241): maximise lnL in one branch
  length inside [brlen_min, brlen_max] with tolerance brlen_min / 10, the
  derivatives coming from ``computeLoglikelihoodDerivatives`` evaluated on
  the branch's sumtables (the engine keeps them on the device while only
  that branch moves);
* ``optimize_branch`` (:345-420): restore the old length when the new one
  does not improve the network lnL (:317-330);
* ``optimize_branches`` (:422-477, 536-583): one pass over the candidate
  branches, each optimised once, and the "Overall loglikelihood got worse"
  check;
* ``optimizeBranches`` (src/optimization/Optimization.cpp:17-37):
  max_iters = brlen_smooth_factor * RAXML_BRLEN_SMOOTHINGS.

The one-dimensional solver is a safeguarded Newton iteration (Newton steps
kept inside a shrinking bracket of the derivative's sign change, bisection
when a step leaves it or the curvature is not negative), written here; it
is not pll-modules' implementation.

Reticulation probabilities and model parameters
(src/optimization/ReticulationOptimization.cpp:68-119,
src/optimization/ModelOptimization.cpp:25-97): the reference maximises the
probabilities one by one with Brent's method on lnL evaluations and the
model through pll-modules' numerical optimisers.  Here both use exact
derivatives: the probabilities by the engine's Newton iteration on
d lnL/dp (every CLV reused, only the displayed-tree weights change), the
model by L-BFGS-B (scipy) on the analytic model gradient, in the
reference's order -- exchangeabilities, then frequencies, then the Gamma
shape -- each stage checked not to lower lnL.
"""
from __future__ import annotations

import dataclasses
from typing import Iterable, Optional

import numpy as np

from . import _lib
from .likelihood import AnnotatedNetwork, computeLoglikelihood

RAXML_BRLEN_SMOOTHINGS = 32
# parameter bounds of pll-modules' model optimisers (pll_optimize.h), the
# dependency the reference calls for them
MIN_SUBST_RATE, MAX_SUBST_RATE = 1e-3, 1e3
MIN_FREQ_RATIO, MAX_FREQ_RATIO = 1e-3, 1e2
MIN_ALPHA, MAX_ALPHA = 0.0201, 100.0
LOGLH_TOLERANCE = 1e-9  # relative lnL loss tolerated by a stage


def newton_maximise(deriv, x0: float, lo: float, hi: float, tol: float, max_iters: int) -> float:
    """Maximises a smooth function of one variable on [lo, hi] given
    ``deriv(x) -> (f', f'')``."""
    x = min(max(x0, lo), hi)
    a, b = lo, hi  # the maximiser lies in [a, b]
    for _ in range(max_iters):
        d1, d2 = deriv(x)
        if d1 > 0:
            a = x
        elif d1 < 0:
            b = x
        else:
            break
        if d2 < 0:
            nx = x - d1 / d2
        else:
            nx = None
        if nx is None or not (a < nx < b):
            # bisection; toward a far bound, grow geometrically instead of
            # jumping across the whole bracket
            if d1 > 0 and b - x > 4.0 * (x + tol):
                nx = min(b, 2.0 * x + tol)
            elif d1 < 0 and x - a > 4.0 * tol and x > 4.0 * a:
                nx = max(a, 0.5 * x)
            else:
                nx = 0.5 * (a + b)
        if abs(nx - x) < tol:
            x = nx
            break
        x = nx
    return min(max(x, lo), hi)


def optimize_branch_newton_raphson(ann_network: AnnotatedNetwork, pmatrix_index: int,
                                   max_iters: int = RAXML_BRLEN_SMOOTHINGS,
                                   mode: int = _lib.NRX_DERIV_REFERENCE) -> float:
    opts = ann_network.options
    tol = opts.brlen_min / 10.0 if opts.brlen_min > 0 else 1e-7

    def deriv(x):
        d = ann_network.brlen_derivatives(pmatrix_index, brlen=x, mode=mode)
        return d.logl_prime, d.logl_prime_prime

    x0 = float(ann_network.network.brlen[pmatrix_index])
    return newton_maximise(deriv, x0, opts.brlen_min, opts.brlen_max, tol, max_iters)


def optimize_branch(ann_network: AnnotatedNetwork, pmatrix_index: int,
                    max_iters: int = RAXML_BRLEN_SMOOTHINGS,
                    mode: int = _lib.NRX_DERIV_REFERENCE) -> float:
    old_logl = computeLoglikelihood(ann_network)
    old_brlen = float(ann_network.network.brlen[pmatrix_index])
    new_brlen = optimize_branch_newton_raphson(ann_network, pmatrix_index, max_iters, mode)
    ann_network.set_branch_length(pmatrix_index, new_brlen)
    new_logl = computeLoglikelihood(ann_network)
    if new_logl < old_logl:
        ann_network.set_branch_length(pmatrix_index, old_brlen)
        new_logl = computeLoglikelihood(ann_network)
    return new_logl


def optimize_branches(ann_network: AnnotatedNetwork, max_iters: int = RAXML_BRLEN_SMOOTHINGS,
                      candidates: Optional[Iterable[int]] = None,
                      mode: int = _lib.NRX_DERIV_REFERENCE) -> float:
    start_logl = computeLoglikelihood(ann_network)
    logl = start_logl
    cands = sorted(range(ann_network.network.n_edges) if candidates is None else set(candidates))
    for e in cands:
        logl = optimize_branch(ann_network, e, max_iters, mode)
    if logl < start_logl and abs(logl - start_logl) >= 1e-3:
        raise RuntimeError("Overall loglikelihood got worse")
    return logl


def optimize_branches_device(ann_network: AnnotatedNetwork, max_iters: int = RAXML_BRLEN_SMOOTHINGS,
                             candidates: Optional[Iterable[int]] = None,
                             mode: int = _lib.NRX_DERIV_REFERENCE) -> float:
    """optimize_branches with each branch's Newton iteration, bracket and
    acceptance check on the device (one host wait per branch)."""
    return ann_network.optimize_branches(candidates, max_iters, mode)


def optimizeBranches(ann_network: AnnotatedNetwork, brlen_smooth_factor: float = 1.0) -> float:
    """src/optimization/Optimization.cpp:17-37 without the BIC bookkeeping."""
    return optimize_branches(ann_network, int(brlen_smooth_factor * RAXML_BRLEN_SMOOTHINGS))


def optimize_reticulations(ann_network: AnnotatedNetwork, max_iters: int = 10) -> float:
    """src/optimization/ReticulationOptimization.cpp:103-119 with the
    per-reticulation Brent search replaced by Newton on the exact
    derivative (nrx_optimize_reticulations)."""
    if ann_network.network.n_retic == 0:
        return computeLoglikelihood(ann_network)
    return ann_network.optimize_reticulations(max_iters)


def optimizeReticulationProbs(ann_network: AnnotatedNetwork) -> float:
    """src/optimization/Optimization.cpp:91-107 without the BIC bookkeeping
    (the parameter count does not change, so BIC moves with lnL)."""
    old = computeLoglikelihood(ann_network)
    new = optimize_reticulations(ann_network, 10)
    if new < old and old - new > 1e-3:
        raise RuntimeError("BIC got worse after optimizing reticulation probs")
    return new


def _with(model, **kw):
    return dataclasses.replace(model, subst_rates=np.array(kw.get("subst_rates", model.subst_rates), float),
                               freqs=np.array(kw.get("freqs", model.freqs), float),
                               alpha=float(kw.get("alpha", model.alpha)))


def _lbfgsb(ann_network: AnnotatedNetwork, x0, bounds, apply, grad_of) -> float:
    """Minimises -lnL over x with L-BFGS-B; ``apply(x)`` sets the models,
    ``grad_of(grads)`` maps the model gradient to d lnL/dx.  Leaves the
    best point seen applied and returns its lnL."""
    from scipy.optimize import minimize

    best = [None, -np.inf]

    def fun(x):
        apply(x)
        lnl, grads = ann_network.model_gradient()
        if lnl > best[1]:
            best[0], best[1] = np.array(x, float), lnl
        return -lnl, -grad_of(grads)

    minimize(fun, np.asarray(x0, float), jac=True, method="L-BFGS-B", bounds=bounds,
             options=dict(ftol=1e7 * np.finfo(float).eps, gtol=1e-6, maxiter=200))
    apply(best[0])
    return best[1]


def _check_improvement(old: float, new: float, where: str) -> None:
    """src/optimization/ModelOptimization.cpp:10-23."""
    if not (old - new < -new * LOGLH_TOLERANCE):
        raise RuntimeError("[%s] Worse log-likelihood after optimization!\nOld: %r\nNew: %r\n" % (where, old, new))


def optimize_params(ann_network: AnnotatedNetwork) -> float:
    """src/optimization/ModelOptimization.cpp:25-97 for GTR+Gamma: the
    exchangeabilities (the last one fixed: lnL does not change when all are
    scaled), then the frequencies (as ratios to the most frequent state,
    renormalised), then the Gamma shape, each by L-BFGS-B on the analytic
    gradient of all partitions at once.  Updates ``partitions[i].model``."""
    parts = ann_network.partitions
    P = len(parts)

    def set_all(models):
        for i, m in enumerate(models):
            parts[i].model = m
            ann_network.set_model(i, m)

    cur = computeLoglikelihood(ann_network)

    # exchangeabilities
    nr = len(parts[0].model.subst_rates)
    fixed = [float(p.model.subst_rates[-1]) for p in parts]
    base = [p.model for p in parts]

    def apply_rates(x):
        x = np.asarray(x).reshape(P, nr - 1)
        set_all([_with(base[i], subst_rates=np.append(x[i], fixed[i])) for i in range(P)])

    x0 = np.concatenate([np.clip(p.model.subst_rates[:-1], MIN_SUBST_RATE, MAX_SUBST_RATE) for p in parts])
    new = _lbfgsb(ann_network, x0, [(MIN_SUBST_RATE, MAX_SUBST_RATE)] * len(x0), apply_rates,
                  lambda g: np.concatenate([g[i]["subst"][:-1] for i in range(P)]))
    if new < cur:
        set_all(base)
        new = computeLoglikelihood(ann_network)
    _check_improvement(cur, new, "RATES")
    cur = new

    # frequencies as ratios x_k = f_k / f_ref, f = x / sum(x)
    S = len(parts[0].model.freqs)
    base = [p.model for p in parts]
    ref = [int(np.argmax(p.model.freqs)) for p in parts]

    def freqs_of(x, i):
        full = np.insert(np.asarray(x), ref[i], 1.0)
        return full / full.sum(), full.sum()

    def apply_freqs(x):
        x = np.asarray(x).reshape(P, S - 1)
        set_all([_with(base[i], freqs=freqs_of(x[i], i)[0]) for i in range(P)])

    def freq_grad(g):
        out = []
        for i in range(P):
            f = parts[i].model.freqs
            tot = 1.0 / f[ref[i]]  # sum of the ratios
            gf = g[i]["freqs"]
            gx = (gf - np.dot(gf, f)) / tot  # d lnL / d x_k for every k
            out.append(np.delete(gx, ref[i]))
        return np.concatenate(out)

    x0 = np.concatenate([np.clip(np.delete(p.model.freqs / p.model.freqs[ref[i]], ref[i]), MIN_FREQ_RATIO,
                                 MAX_FREQ_RATIO) for i, p in enumerate(parts)])
    new = _lbfgsb(ann_network, x0, [(MIN_FREQ_RATIO, MAX_FREQ_RATIO)] * len(x0), apply_freqs, freq_grad)
    if new < cur:
        set_all(base)
        new = computeLoglikelihood(ann_network)
    _check_improvement(cur, new, "FREQS")
    cur = new

    # Gamma shape
    if parts[0].model.ncat > 1:
        base = [p.model for p in parts]

        def apply_alpha(x):
            set_all([_with(base[i], alpha=x[i]) for i in range(P)])

        x0 = np.clip([p.model.alpha for p in parts], MIN_ALPHA, MAX_ALPHA)
        new = _lbfgsb(ann_network, x0, [(MIN_ALPHA, MAX_ALPHA)] * P, apply_alpha,
                      lambda g: np.array([g[i]["alpha"] for i in range(P)]))
        if new < cur:
            set_all(base)
            new = computeLoglikelihood(ann_network)
        _check_improvement(cur, new, "ALPHA")
    return new


def optimizeModel(ann_network: AnnotatedNetwork) -> float:
    """src/optimization/Optimization.cpp:72-83 without the BIC bookkeeping
    (the parameter count does not change, so BIC moves with lnL)."""
    return optimize_params(ann_network)
