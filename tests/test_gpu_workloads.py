"""Full-size workloads and edge cases of the batched (FP32) form on the GPU:

* the 256-candidate search-step batch (20-60 taxa, 0-5 reticulations,
  2k-50k sites), scored by one batch evaluation, against every engine
  scored alone and a stratified sample against the oracle;
* the Newton-Raphson branch sweep at 50 taxa / 4 reticulations / 50k sites
  (BranchLengthOptimization.cpp:422-477), each accepted length checked to
  be a stationary point of the oracle's reference-normalised derivative;
* FP32 at clamped zero branch lengths (P = I at 1e-6);
* FP32 with 1, 2 and 8 rate categories, every kernel shape, single engine
  and batch, against the oracle;
* batch / derivative argument handling.
"""
import math
import sys

import numpy as np
import pytest

from fixtures import fixed_model, load_sample
from netrax_amd import (AnnotatedNetwork, LikelihoodVariant, NetraxOptions, NetworkBatch, Precision,
                        computeLoglikelihood)
from oracle import oracle as O

pytestmark = pytest.mark.gpu

FP32_RTOL = 1e-6   # lnL, relative to the FP64 oracle
GRAD_TOL = 1e-4    # gradient, relative to the largest gradient entry


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def eng(net, parts, variant=0, precision=Precision.FP32, **kw):
    return AnnotatedNetwork(net, parts, NetraxOptions(likelihood_variant=LikelihoodVariant(variant),
                                                      precision=precision, **kw))


def _grad_close(g, og, tol=GRAD_TOL):
    scale = max(1.0, float(np.abs(og).max(initial=0.0)))
    return float(np.abs(np.asarray(g) - np.asarray(og)).max(initial=0.0)) <= tol * scale


def test_candidate_batch_256_matches_engines_alone_and_oracle():
    from netrax_amd.simulate import candidate_batch
    say("256 candidates: building")
    O.set_threads(16)
    cands = candidate_batch(8, 32, seed=5)
    anns = [eng(n, p, use_graph=True) for n, p in cands]
    best, _ = anns[0].autotune(reps=1)
    for a in anns:
        a.set_tuning(*best[:3])
    batch = NetworkBatch(anns)
    say("256 candidates: batch evaluation")
    lnl, dbs, dps, ms = batch.evaluate(0)
    assert ms > 0 and np.all(np.isfinite(lnl))
    # every engine scored alone (same pass slots as in the batch): identical
    for i, a in enumerate(anns):
        l1, g1, p1 = a.gradient()
        assert abs(lnl[i] - l1) <= 1e-12 * abs(l1), (i, lnl[i], l1)
        assert np.allclose(dbs[i], g1, rtol=1e-10, atol=1e-10 * max(1.0, np.abs(g1).max()))
        assert np.allclose(dps[i], p1, rtol=1e-10, atol=1e-10)
    # a stratified sample against the oracle: one candidate of every base,
    # plus the largest site count and the most reticulations
    info = [(n.n_tips, n.n_retic, p[0].n_sites) for n, p in cands]
    pick = {b * 32 + (7 * b) % 32 for b in range(8)}
    pick.add(max(range(len(cands)), key=lambda i: (info[i][2], info[i][1])))
    pick.add(max(range(len(cands)), key=lambda i: (info[i][1], info[i][2])))
    pick.add(min(range(len(cands)), key=lambda i: (info[i][1], info[i][2])))
    assert len(pick) >= 8
    for i in sorted(pick):
        n, p = cands[i]
        say("candidate", i, "taxa/retic/patterns", info[i])
        ol, og, op = O.gradient(n, p, O.AVERAGE)
        assert abs(lnl[i] - ol) <= FP32_RTOL * abs(ol), (i, lnl[i], ol)
        assert _grad_close(dbs[i], og), (i, np.abs(dbs[i] - og).max())
        assert _grad_close(dps[i], op), (i, dps[i], op)
    batch.close()
    for a in anns:
        a.close()


def test_newton_sweep_50taxa_50k_reaches_oracle_stationary_points():
    """optimize_branches on 10 branches of the 50-taxon workload: lnL does
    not decrease, and every accepted length is where the oracle's
    reference-normalised first derivative vanishes (a Newton step from it
    is below the sweep's tolerance) or a bound the derivative pushes
    against.  Also the engine's mode-0 derivatives against the oracle."""
    from netrax_amd.optimization import optimize_branch
    from netrax_amd.simulate import workload_50taxa_50k
    O.set_threads(16)
    w = workload_50taxa_50k()
    net = w.network
    rng = np.random.default_rng(4)
    net.brlen = net.brlen * rng.uniform(0.5, 2.0, net.n_edges)
    ann = eng(net, w.partitions, precision=Precision.FP64)
    start = computeLoglikelihood(ann)
    edges = sorted(rng.choice(net.n_edges, 10, replace=False).tolist())
    logl = start
    for e in edges:
        before = net.brlen.copy()
        new = optimize_branch(ann, e)
        assert new >= logl - 1e-6 * abs(logl), (e, new, logl)
        logl = new
        after = net.brlen.copy()
        assert np.array_equal(np.delete(after, e), np.delete(before, e))
        if after[e] == before[e]:
            continue  # step rejected, old length kept
        d1, d2, _, _ = O.brlen_derivatives(net, w.partitions, e, O.AVERAGE, mode=0, brlen=after)
        say("edge", e, "length", before[e], "->", after[e], "oracle d1", d1, "d2", d2)
        at_lo = after[e] <= 1e-6 * (1 + 1e-12) and d1 < 0
        at_hi = after[e] >= 100.0 * (1 - 1e-12) and d1 > 0
        assert at_lo or at_hi or (d2 < 0 and abs(d1 / d2) <= 1e-6), (e, after[e], d1, d2)
        olnl = O.loglikelihood(net, w.partitions, brlen=after)[0]
        assert abs(new - olnl) <= 1e-9 * abs(olnl)
    assert logl > start
    for e in edges[:3]:
        d = ann.brlen_derivatives(e, mode=0)
        o1, o2, _, _ = O.brlen_derivatives(net, w.partitions, e, O.AVERAGE, mode=0)
        assert abs(d.logl_prime - o1) <= 1e-7 * max(1.0, abs(o1)), (e, d.logl_prime, o1)
        assert abs(d.logl_prime_prime - o2) <= 1e-7 * max(1.0, abs(o2)), (e, d.logl_prime_prime, o2)
    ann.close()


def test_fp32_celine_zero_branch_lengths_match_oracle():
    """celine without lengths: every branch clamped to 1e-6, P = I to FP32
    precision; lnL, dlnL/dbrlen (P' = r Q P at t = 1e-6) and dlnL/dp."""
    say("fp32 celine, clamped zero lengths")
    net, parts = load_sample("celine", fixed_model(), with_lengths=False)
    for var in (0, 1):
        ann = eng(net, parts, var)
        lnl, g, gp = ann.gradient()
        ol, og, op = O.gradient(net, parts, var, brlen=np.maximum(net.brlen, 1e-6))
        assert abs(lnl - ol) <= FP32_RTOL * abs(ol), (var, lnl, ol)
        assert _grad_close(g, og), (var, np.abs(g - og).max(), np.abs(og).max())
        assert _grad_close(gp, op), (var, gp, op)
        ann.close()


@pytest.mark.parametrize("ncat", [1, 2, 8])
def test_fp32_other_category_counts_every_shape_and_batch(ncat):
    """The FP32 kernels at 1, 2 and 8 categories (per-site LDS exponent
    slots, category reductions over C lanes, C-dependent tile shapes and
    gradient-partial layouts) against the oracle, in every shape, alone
    and in a batch."""
    say("fp32 categories", ncat)
    net, parts = load_sample("three_reticulations", fixed_model(ncat=ncat))
    from netrax_amd.simulate import make_workload
    w = make_workload("c", 24, 3, [700, 1300], seed=30 + ncat)
    for p in w.partitions:
        p.model.ncat = ncat
    cases = [(net, parts), (w.network, w.partitions)]
    for n, p in cases:
        for var in (0, 1):
            ol, og, op = O.gradient(n, p, var)
            ann = eng(n, p, var)
            _, shapes = ann.autotune(variant=var, reps=1)
            assert shapes
            for gw, ts, h, _ms in shapes:
                ann.set_tuning(gw, ts, h)
                lnl, g, gp = ann.gradient()
                assert abs(lnl - ol) <= FP32_RTOL * abs(ol), (ncat, var, gw, ts, h, lnl, ol)
                assert _grad_close(g, og), (ncat, var, gw, ts, h, np.abs(g - og).max())
                assert _grad_close(gp, op), (ncat, var, gw, ts, h)
            ann.close()
    # batch of both networks (one shape shared)
    anns = [eng(n, p) for n, p in cases]
    batch = NetworkBatch(anns)
    lnl, dbs, dps, _ = batch.evaluate(0)
    for i, (n, p) in enumerate(cases):
        ol, og, op = O.gradient(n, p, 0)
        assert abs(lnl[i] - ol) <= FP32_RTOL * abs(ol), (ncat, i)
        assert _grad_close(dbs[i], og) and _grad_close(dps[i], op), (ncat, i)
    batch.close()
    for a in anns:
        a.close()


def test_batch_refuses_repeated_and_closed_engines():
    from netrax_amd import _lib
    import ctypes as C
    net, parts = load_sample("two_reticulations", fixed_model())
    a, b = eng(net, parts), eng(net, parts)
    with pytest.raises(ValueError):
        NetworkBatch([a, b, a])
    # the C ABI refuses it too, before any device work
    lib = _lib.load()
    arr = (C.c_void_p * 2)(a.handle.value, a.handle.value)
    h = C.c_void_p()
    assert lib.nrx_batch_create(arr, 2, C.byref(h)) == _lib.NRX_ERR_ARG
    assert b"twice" in lib.nrx_last_error()
    batch = NetworkBatch([a, b])
    batch.evaluate(0)
    b.close()
    with pytest.raises(RuntimeError):
        batch.evaluate(0)
    batch.close()
    a.close()


@pytest.mark.parametrize("prec", [Precision.FP64, Precision.FP32])
def test_brlen_derivatives_clamp_lengths_below_the_minimum(prec):
    net, parts = load_sample("two_reticulations", fixed_model())
    ann = eng(net, parts, precision=prec)
    for e in (0, 5, 9):
        d0 = ann.brlen_derivatives(e, brlen=0.0, mode=1)
        dm = ann.brlen_derivatives(e, brlen=1e-6, mode=1)
        assert d0.logl_prime == dm.logl_prime and d0.logl_prime_prime == dm.logl_prime_prime
        dh = ann.brlen_derivatives(e, brlen=1e4, mode=1)
        dmax = ann.brlen_derivatives(e, brlen=100.0, mode=1)
        assert dh.logl_prime == dmax.logl_prime
    ann.close()


@pytest.mark.parametrize("name,prec", [("50taxa_50k", Precision.FP64), ("50taxa_50k", Precision.FP32),
                                       ("two_reticulations", Precision.FP64)])
def test_device_newton_sweep_matches_host_sweep(name, prec):
    """nrx_optimize_branches (Newton iteration, bracket and acceptance on
    the device) gives the lengths and lnL of the host-driven sweep
    (optimize_branches) for both variants and derivative modes."""
    from netrax_amd.optimization import optimize_branches
    from netrax_amd.simulate import workload_50taxa_50k
    if name == "two_reticulations":
        net0, parts = load_sample(name, fixed_model())
        edges = list(range(net0.n_edges))
        cases = [(0, 0), (0, 1), (1, 0), (1, 1)]
    else:
        w = workload_50taxa_50k()
        net0, parts = w.network, w.partitions
        edges = sorted(np.random.default_rng(8).choice(net0.n_edges, 10, replace=False).tolist())
        cases = [(0, 0)] if prec == Precision.FP32 else [(0, 0), (1, 1)]
    rng = np.random.default_rng(12)
    start = net0.brlen * rng.uniform(0.5, 2.0, net0.n_edges)
    for variant, mode in cases:
        say("device sweep", name, prec.name, "variant", variant, "mode", mode)
        nh, nd = net0.copy(), net0.copy()
        nh.brlen = start.copy()
        nd.brlen = start.copy()
        host = eng(nh, parts, variant, precision=prec)
        dev = eng(nd, parts, variant, precision=prec)
        before = computeLoglikelihood(host)
        lh = optimize_branches(host, candidates=edges, mode=mode)
        ld = dev.optimize_branches(edges, mode=mode)
        assert ld >= before - 1e-6 * abs(before)
        assert np.allclose(dev.network.brlen, host.network.brlen, rtol=1e-6, atol=1e-9), \
            np.abs(dev.network.brlen - host.network.brlen).max()
        assert abs(ld - lh) <= 1e-9 * abs(lh), (ld, lh)
        assert np.array_equal(dev.branch_lengths(), dev.network.brlen)
        host.close()
        dev.close()


def _move_edges(net):
    """Branches around the reticulations (those a move touches), plus the
    first edge, in index order."""
    es = {0}
    for r in range(net.n_retic):
        es.add(int(net.retic_first_edge[r]))
        es.add(int(net.retic_second_edge[r]))
    return sorted(es)


def test_batch_optimize_matches_engines_optimised_alone_and_oracle():
    from netrax_amd.simulate import candidate_batch
    cands = candidate_batch(2, 8, seed=7, sites=(2000, 8000))
    say("batch optimise: 16 candidates")
    edges = [_move_edges(n) for n, _p in cands]
    retics = [list(range(n.n_retic)) for n, _p in cands]
    orig = [n.copy() for n, _p in cands]  # the engines update their networks
    start = []
    anns = []
    for n, p in cands:
        a = eng(n, p, use_graph=True)
        start.append(a.loglikelihood())
        anns.append(a)
    batch = NetworkBatch(anns)
    lnl, st = batch.optimize(edges, retics, brlen_iters=32, prob_iters=10, n_threads=8)
    assert np.all(st == 0), st
    say("batch optimise: lnL gain", np.round(lnl - np.array(start), 3))
    # the same steps on twin engines, one at a time
    for i, (n, p) in enumerate(cands):
        twin = eng(orig[i].copy(), p, use_graph=True)
        twin.optimize_branches(edges[i], max_iters=32)
        lt = twin.optimize_reticulations(10) if n.n_retic else twin.loglikelihood()
        assert abs(lt - lnl[i]) <= 1e-12 * abs(lt), (i, lt, lnl[i])
        np.testing.assert_array_equal(twin.branch_lengths(), anns[i].branch_lengths())
        assert lnl[i] >= start[i] - FP32_RTOL * abs(start[i]), (i, lnl[i], start[i])
        twin.close()
    # the optimised parameters score the same on the oracle
    for i in (0, 7, 8, 15):
        n, p = cands[i]
        ol = O.loglikelihood(n, p, O.AVERAGE, brlen=anns[i].branch_lengths(),
                             probs=anns[i].reticulation_probs() if n.n_retic else None)[0]
        assert abs(lnl[i] - ol) <= FP32_RTOL * abs(ol), (i, lnl[i], ol)
    batch.close()
    for a in anns:
        a.close()
