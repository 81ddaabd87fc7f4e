"""hipGraph replays of the FP32 form (single engine and network batch):
every replay, with eager work (P-matrix updates, host synchronisation)
between replays, is bit-identical to the same evaluation launched eagerly."""
import numpy as np
import pytest

from netrax_amd import (AnnotatedNetwork, LikelihoodVariant, NetraxOptions, NetworkBatch,
                        Precision)
from netrax_amd.simulate import candidate_batch, make_workload, workload_30taxa_10k

pytestmark = pytest.mark.gpu


def _pair(w, variant=0):
    mk = lambda g: AnnotatedNetwork(w.network, w.partitions, NetraxOptions(
        likelihood_variant=LikelihoodVariant(variant), precision=Precision.FP32, use_graph=g))
    return mk(True), mk(False)


def _replays(w):
    g, e = _pair(w)
    rng = np.random.default_rng(7)
    b0 = w.network.brlen.copy()
    for it in range(4):
        b = b0 * rng.uniform(0.5, 1.5, size=b0.shape)
        for ann in (g, e):
            ann.set_branch_lengths(b)
        var = it % 2
        lg, gg, pg = g.gradient(variant=var)
        le, ge, pe = e.gradient(variant=var)
        assert lg == le and np.array_equal(gg, ge) and np.array_equal(pg, pe), (it, lg, le)
        # lnL only (no adjoint): a second captured graph
        assert g.loglikelihood(incremental=0) == e.loglikelihood(incremental=0)
        # moved branches evaluated with the P matrices left as they are (a
        # graph without the P kernels), then with them updated
        for ann in (g, e):
            ann.set_branch_lengths(b * 1.1)
        stale = g.loglikelihood(incremental=0, update_pmatrices=0)
        assert stale == e.loglikelihood(incremental=0, update_pmatrices=0)
        fresh = g.loglikelihood(incremental=0)
        assert fresh == e.loglikelihood(incremental=0) and fresh != stale
    if w.network.n_retic:
        # a reticulation probability that drops trees below the threshold
        # changes the schedule: the graphs are captured again
        p = w.network.retic_prob.copy()
        p[0] = 1e-7
        for ann in (g, e):
            ann.set_reticulation_probs(p)
        lg, gg, _ = g.gradient()
        le, ge, _ = e.gradient()
        assert lg == le and np.array_equal(gg, ge)
    g.close()
    e.close()


def test_engine_graph_replays_match_eager():
    print("graph replays 30 taxa", flush=True)
    _replays(workload_30taxa_10k())


def test_protein_fp32_graph_replays_match_eager():
    print("graph replays protein", flush=True)
    _replays(make_workload("aa", 20, 2, [900], seed=8, protein=True))


def test_batch_graph_replays_match_eager():
    print("batch graph replays", flush=True)
    cands = candidate_batch(2, 3, seed=9, sites=(300, 3000))
    mk = lambda g: [AnnotatedNetwork(n, p, NetraxOptions(precision=Precision.FP32, use_graph=g))
                    for n, p in cands]
    ag, ae = mk(True), mk(False)
    bg, be = NetworkBatch(ag), NetworkBatch(ae)
    rng = np.random.default_rng(3)
    for it in range(3):
        for (n, _), x, y in zip(cands, ag, ae):
            b = n.brlen * rng.uniform(0.5, 1.5, size=n.brlen.shape)
            x.set_branch_lengths(b)
            y.set_branch_lengths(b)
        var = it % 2
        rg = bg.evaluate(var)
        re = be.evaluate(var)
        assert np.array_equal(rg[0], re[0]), it
        for k in (1, 2):
            for u, v in zip(rg[k], re[k]):
                assert np.array_equal(u, v), (it, k)
        assert np.array_equal(bg.evaluate(var, gradient=False)[0], rg[0])
    bg.close()
    be.close()
