"""Autotuned kernel shapes of the FP32 form: every candidate shape (ops per
group x sites per workgroup) gives the same lnL and gradient as the FP64
form, the autotuner keeps the fastest, and batches run with a tuned shape."""
import numpy as np
import pytest

from netrax_amd import AnnotatedNetwork, NetraxOptions, NetworkBatch, Precision
from netrax_amd.simulate import candidate_batch, make_workload, workload_30taxa_10k

pytestmark = pytest.mark.gpu


def _check_all_shapes(w):
    a64 = AnnotatedNetwork(w.network, w.partitions, NetraxOptions())
    a32 = AnnotatedNetwork(w.network, w.partitions, NetraxOptions(precision=Precision.FP32))
    l64, g64, p64 = a64.gradient()
    best, cands = a32.autotune(reps=2)
    print("autotune", best, cands, flush=True)
    assert len(cands) >= 3
    assert best[3] == min(c[3] for c in cands)
    assert a32.tuning() == best[:3]
    for gw, ts, h, ms in cands:
        a32.set_tuning(gw, ts, h)
        l32, g32, p32 = a32.gradient()
        assert abs(l32 - l64) <= 1e-6 * abs(l64), (gw, ts, h, l32, l64)
        scale = max(1.0, np.abs(g64).max())
        assert np.abs(g32 - g64).max() <= 1e-4 * scale, (gw, ts, h)
        assert np.abs(p32 - p64).max(initial=0.0) <= 1e-4 * max([1.0, *np.abs(p64)])
    with pytest.raises(Exception):
        a32.set_tuning(5, 64)
    a32.close()
    a64.close()


def test_every_shape_matches_fp64_and_autotune_picks_fastest():
    print("shapes 30 taxa", flush=True)
    _check_all_shapes(workload_30taxa_10k())


def test_protein_fp32_every_shape_matches_fp64():
    print("shapes protein", flush=True)
    _check_all_shapes(make_workload("aa", 20, 2, [700], seed=8, protein=True))


def test_batch_runs_with_a_tuned_shape():
    print("batch with shapes (2, 32, 1) and (3, 32, 2)", flush=True)
    cands = candidate_batch(2, 2, seed=10, sites=(300, 2000))
    anns = [AnnotatedNetwork(n, p, NetraxOptions(precision=Precision.FP32)) for n, p in cands]
    for shape in ((2, 32, 1), (3, 32, 2)):
        for a in anns:
            a.set_tuning(*shape)
        batch = NetworkBatch(anns)
        lnl, dbs, dps, _ = batch.evaluate(0)
        for i, a in enumerate(anns):
            l, g, p = a.gradient()
            assert abs(l - lnl[i]) <= 1e-9 * abs(l), shape
            assert np.allclose(g, dbs[i], rtol=1e-6, atol=1e-6 * max(1.0, np.abs(g).max())), shape
        batch.close()
    anns[0].set_tuning(4, 64, 2)
    with pytest.raises(Exception):
        NetworkBatch(anns)


def test_batch_gives_up_lds_pass_slots_and_results_stay_the_same():
    """A batch launch holds many more workgroups than one engine's, so each
    engine keeps only as many LDS pass slots as leave the workgroups
    resident per CU unchanged.  Deliveries moved back to HBM are normalised
    the same way, so the batch and the engines alone (before and after)
    agree to the batch tolerance."""
    print("batch pass slots: 4 networks, 50-60 taxa, 30k sites", flush=True)
    cands = candidate_batch(1, 4, seed=21, taxa=(50, 60), retic=(2, 4), sites=(30000, 30000))
    anns = [AnnotatedNetwork(n, p, NetraxOptions(precision=Precision.FP32)) for n, p in cands]
    # 32-site workgroups: small enough that LDS, not registers, limits how
    # many are resident per CU
    for a in anns:
        a.set_tuning(2, 32)
    solo = [a.gradient() for a in anns]
    before = [a.stats().n_lds_handoffs for a in anns]
    assert min(before) > 0
    batch = NetworkBatch(anns)
    after = [a.stats().n_lds_handoffs for a in anns]
    print("hand-offs before", before, "after", after, flush=True)
    assert all(x <= y for x, y in zip(after, before))
    assert sum(after) < sum(before)  # over 768 tiles fill every CU: fewer slots
    lnl, dbs, _, _ = batch.evaluate(0)
    for i, a in enumerate(anns):
        l0, g0, _ = solo[i]
        assert abs(lnl[i] - l0) <= 1e-9 * abs(l0), (i, lnl[i], l0)
        assert np.allclose(dbs[i], g0, rtol=1e-6, atol=1e-6 * max(1.0, np.abs(g0).max()))
        l1, g1, _ = a.gradient()
        assert abs(l1 - l0) <= 1e-9 * abs(l0), (i, l1, l0)
        assert np.allclose(g1, g0, rtol=1e-6, atol=1e-6 * max(1.0, np.abs(g0).max()))
    batch.close()
    for a in anns:
        a.close()
