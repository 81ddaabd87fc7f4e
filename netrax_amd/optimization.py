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
"""
from __future__ import annotations

from typing import Iterable, Optional

from . import _lib
from .likelihood import AnnotatedNetwork, computeLoglikelihood

RAXML_BRLEN_SMOOTHINGS = 32


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
