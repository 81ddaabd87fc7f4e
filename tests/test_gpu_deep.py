"""A deep network: a 40-taxon caterpillar with two reticulations, so that
root-to-branch paths run to ~40 steps (past the P blocks the path kernels
stage in LDS, 24 steps) and the pruning schedule has ~40 levels of one or
two ops.  Branch derivatives (both modes, both variants, FP64 and FP32,
allocating and allocation-free), the device Newton sweep and the gradient
against the oracle."""
import ctypes as C
import sys

import numpy as np
import pytest

from netrax_amd import AnnotatedNetwork, NetraxOptions, Precision, _lib
from netrax_amd.network import parse_newick
from netrax_amd.optimization import optimize_branches
from netrax_amd.simulate import Partition, compress_patterns, random_dna_model, simulate_states, states_to_masks
from oracle import oracle as O

pytestmark = pytest.mark.gpu
N_TAXA = 40


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _caterpillar():
    """((((t0,t1),t2),t3)...), with t4's and t11's parents made reticulations
    whose second parents sit next to t30 and t36."""
    s = "t0"
    for i in range(1, N_TAXA):
        leaf = f"t{i}"
        if i == 4:
            leaf = "(t4)R1#H1"
        elif i == 11:
            leaf = "(t11)R2#H2"
        elif i == 30:
            leaf = "(t30,R1#H1)"
        elif i == 36:
            leaf = "(t36,R2#H2)"
        s = f"({s},{leaf})"
    net = parse_newick(s + ";")
    rng = np.random.default_rng(17)
    net.brlen = rng.uniform(0.02, 0.25, net.n_edges)
    net.retic_prob = np.array([0.35, 0.7])[:net.n_retic]
    return net


def _workload(sites=1500):
    net = _caterpillar()
    rng = np.random.default_rng(18)
    model = random_dna_model(rng)
    masks, weights = compress_patterns(states_to_masks(simulate_states(net, model, sites, rng)))
    return net, [Partition(model, masks, weights)]


def _deep_edges(net, k=5):
    """The k edges whose paths from the network root are longest: edges
    below the caterpillar's deepest inner nodes."""
    depth = np.zeros(net.n_nodes, int)
    # depth by repeated relaxation over the edge list (parent -> child)
    for _ in range(net.n_nodes):
        for e in range(net.n_edges):
            p, c = int(net.edge_parent[e]), int(net.edge_child[e])
            depth[c] = max(depth[c], depth[p] + 1)
    d = np.array([depth[int(net.edge_child[e])] for e in range(net.n_edges)])
    return [int(e) for e in np.argsort(-d, kind="stable")[:k]], int(d.max())


def test_deep_network_shape():
    net, _ = _workload(200)
    assert net.n_retic == 2 and net.n_tips == N_TAXA
    _, dmax = _deep_edges(net)
    assert dmax > 30, dmax


@pytest.mark.parametrize("prec", [Precision.FP64, Precision.FP32])
def test_deep_network_derivatives_match_oracle(prec):
    net, parts = _workload()
    edges, dmax = _deep_edges(net)
    say("deep network", prec.name, "depth", dmax, "edges", edges)
    tol = 1e-7 if prec == Precision.FP64 else 2e-4
    lib = _lib.load()
    for var in (0, 1):
        ann = AnnotatedNetwork(net.copy(), parts, NetraxOptions(precision=prec, likelihood_variant=var))
        lnl, g, gp = ann.gradient()
        ol, og, ogp = O.gradient(net, parts, var)[:3]
        assert abs(lnl - ol) <= (1e-9 if prec == Precision.FP64 else 2e-6) * abs(ol), (lnl, ol)
        gt = 1e-7 if prec == Precision.FP64 else 2e-4
        assert np.allclose(g, og, rtol=gt, atol=gt * max(1.0, np.abs(og).max()))
        for e in edges:
            for mode in (0, 1):
                d = ann.brlen_derivatives(e, mode=mode)
                o1, o2, _, _ = O.brlen_derivatives(net, parts, e, var, mode=mode)
                assert abs(d.logl_prime - o1) <= tol * max(1.0, abs(o1)), (var, e, mode, d.logl_prime, o1)
                assert abs(d.logl_prime_prime - o2) <= tol * max(1.0, abs(o2)), (var, e, mode)
                # the allocation-free entry point walks the same long paths
                d1, d2 = C.c_double(), C.c_double()
                p1, p2 = np.zeros(1), np.zeros(1)
                pd = C.POINTER(C.c_double)
                rc = lib.nrx_brlen_derivatives_noalloc(ann.handle, var, e, mode, float(ann.network.brlen[e]),
                                                       C.byref(d1), C.byref(d2), p1.ctypes.data_as(pd),
                                                       p2.ctypes.data_as(pd))
                assert rc == _lib.NRX_OK, lib.nrx_last_error()
                assert abs(d1.value - d.logl_prime) <= 1e-12 * max(1.0, abs(d.logl_prime))
                assert abs(d2.value - d.logl_prime_prime) <= 1e-12 * max(1.0, abs(d.logl_prime_prime))
        ann.close()


def test_deep_network_device_sweep_matches_host_sweep():
    net, parts = _workload()
    edges, _ = _deep_edges(net, 8)
    rng = np.random.default_rng(4)
    start = net.brlen * rng.uniform(0.5, 2.0, net.n_edges)
    out = []
    for device in (True, False):
        n = net.copy()
        n.brlen = start.copy()
        ann = AnnotatedNetwork(n, parts, NetraxOptions())
        l0 = ann.loglikelihood()
        l1 = ann.optimize_branches(edges) if device else optimize_branches(ann, candidates=edges)
        out.append((l0, l1, ann.network.brlen.copy()))
        ann.close()
    (l0, ld, bd), (_, lh, bh) = out
    say("deep sweep", l0, "->", ld, lh)
    assert ld >= l0 - 1e-6 * abs(l0)
    assert np.allclose(bd, bh, rtol=1e-6, atol=1e-9)
    assert abs(ld - lh) <= 1e-9 * abs(lh)


def test_wide_network_gradient_matches_oracle():
    """30 taxa, 7 reticulations (128 displayed trees): pruning levels of up
    to 96 ops, past the 48 whose P blocks the plain reverse sweep stages in
    LDS per level, so the rest read P from memory."""
    from netrax_amd.simulate import make_workload
    w = make_workload("wide", 30, 7, [800], seed=2)
    net, parts = w.network, w.partitions
    for prec, tol in ((Precision.FP64, 1e-9), (Precision.FP32, 2e-6)):
        for var in (0, 1):
            say("wide network", prec.name, "variant", var)
            ann = AnnotatedNetwork(net.copy(), parts, NetraxOptions(precision=prec, likelihood_variant=var))
            lnl, g, gp = ann.gradient()
            ol, og, ogp = O.gradient(net, parts, var)[:3]
            assert abs(lnl - ol) <= tol * abs(ol), (prec.name, var, lnl, ol)
            gt = 1e-7 if prec == Precision.FP64 else 2e-4
            assert np.allclose(g, og, rtol=gt, atol=gt * max(1.0, np.abs(og).max())), (prec.name, var)
            assert np.allclose(gp, ogp, rtol=gt, atol=gt * max(1.0, np.abs(ogp).max())), (prec.name, var)
            ann.close()


@pytest.mark.parametrize("variant", [0, 1])
def test_wide_network_device_sweep_matches_host_sweep(variant):
    """128 displayed-tree classes: the device Newton step and acceptance
    spread them over a 64-lane wave, two per lane."""
    from netrax_amd.simulate import make_workload
    w = make_workload("wide", 30, 7, [800], seed=2)
    net, parts = w.network, w.partitions
    rng = np.random.default_rng(6)
    edges = sorted(int(e) for e in rng.choice(net.n_edges, 8, replace=False))
    start = net.brlen * rng.uniform(0.5, 2.0, net.n_edges)
    out = []
    for device in (True, False):
        n = net.copy()
        n.brlen = start.copy()
        ann = AnnotatedNetwork(n, parts, NetraxOptions(likelihood_variant=variant))
        l0 = ann.loglikelihood()
        l1 = ann.optimize_branches(edges) if device else optimize_branches(ann, candidates=edges)
        out.append((l0, l1, ann.network.brlen.copy()))
        ann.close()
    (l0, ld, bd), (_, lh, bh) = out
    say("wide sweep, variant", variant, l0, "->", ld, lh)
    assert ld >= l0 - 1e-6 * abs(l0)
    assert np.allclose(bd, bh, rtol=1e-6, atol=1e-9)
    assert abs(ld - lh) <= 1e-9 * abs(lh)


@pytest.mark.parametrize("shape", [(16, 9, 5), (24, 11, 7)])
def test_many_reticulations_match_oracle(shape):
    """9 and 11 reticulations (512 and 2,048 configurations, 648 distinct
    displayed-tree classes for the larger): the mixture's inputs no longer
    fit its LDS staging and run from memory; lnL, branch and probability
    gradient of both forms and variants against the oracle."""
    from netrax_amd.simulate import make_workload
    nt, nr, seed = shape
    w = make_workload("many", nt, nr, [300], seed=seed)
    net, parts = w.network, w.partitions
    for prec, tol in ((Precision.FP64, 1e-9), (Precision.FP32, 2e-6)):
        for var in (0, 1):
            say("many reticulations", nt, nr, prec.name, "variant", var)
            ann = AnnotatedNetwork(net.copy(), parts, NetraxOptions(precision=prec, likelihood_variant=var))
            st = ann.stats()
            staged = st.n_root_clvs * len(parts) + 2 * st.n_classes + st.n_classes * len(parts) \
                + st.n_classes * net.n_retic
            say("classes", st.n_classes, "root CLVs", st.n_root_clvs, "mixture doubles", staged)
            if nr == 11:
                assert staged > 6144  # past the mixture's LDS staging
            lnl, g, gp = ann.gradient()
            ol, og, ogp = O.gradient(net, parts, var)[:3]
            assert abs(lnl - ol) <= tol * abs(ol), (prec.name, var, lnl, ol)
            gt = 1e-7 if prec == Precision.FP64 else 2e-4
            assert np.allclose(g, og, rtol=gt, atol=gt * max(1.0, np.abs(og).max())), (prec.name, var)
            assert np.allclose(gp, ogp, rtol=gt, atol=gt * max(1.0, np.abs(ogp).max())), (prec.name, var)
            ann.close()


def test_many_classes_device_sweep_matches_host_sweep():
    """2,304 displayed-tree classes (12 reticulations): the Newton kernels'
    class table alone exceeds their LDS staging, so it, the tree lnLs and the
    path sums are read from memory."""
    from netrax_amd.simulate import make_workload
    w = make_workload("many", 24, 12, [300], seed=8)
    net, parts = w.network, w.partitions
    rng = np.random.default_rng(8)
    edges = sorted(int(e) for e in rng.choice(net.n_edges, 5, replace=False))
    start = net.brlen * rng.uniform(0.5, 2.0, net.n_edges)
    out = []
    for device in (True, False):
        n = net.copy()
        n.brlen = start.copy()
        ann = AnnotatedNetwork(n, parts, NetraxOptions())
        st = ann.stats()
        assert 24 * st.n_classes > 32768  # the class table alone is past the Newton kernels' LDS staging
        l0 = ann.loglikelihood()
        l1 = ann.optimize_branches(edges) if device else optimize_branches(ann, candidates=edges)
        out.append((l0, l1, ann.network.brlen.copy()))
        ann.close()
    (l0, ld, bd), (_, lh, bh) = out
    say("many-class sweep", l0, "->", ld, lh)
    assert ld >= l0 - 1e-6 * abs(l0)
    assert np.allclose(bd, bh, rtol=1e-6, atol=1e-9)
    assert abs(ld - lh) <= 1e-9 * abs(lh)


@pytest.mark.parametrize("protein", [False, True])
def test_ragged_partitions_match_oracle(protein):
    """Partitions of 1, 65 and 7 sites (each padded to whole 64-site tiles,
    one of them a single site): lnL and gradient of both forms against the
    oracle."""
    from netrax_amd.simulate import make_workload
    w = make_workload("ragged", 12, 2, [1, 65, 7], seed=9, protein=protein)
    net, parts = w.network, w.partitions
    assert min(p.n_sites for p in parts) == 1
    for prec, tol in ((Precision.FP64, 1e-9), (Precision.FP32, 5e-6)):
        for var in (0, 1):
            ann = AnnotatedNetwork(net.copy(), parts, NetraxOptions(precision=prec, likelihood_variant=var))
            lnl, g, gp = ann.gradient()
            ol, og, ogp = O.gradient(net, parts, var)[:3]
            assert abs(lnl - ol) <= tol * abs(ol), (protein, prec.name, var, lnl, ol)
            gt = 1e-7 if prec == Precision.FP64 else 5e-4
            assert np.allclose(g, og, rtol=gt, atol=gt * max(1.0, np.abs(og).max())), (protein, prec.name, var)
            assert np.allclose(gp, ogp, rtol=gt, atol=gt * max(1.0, np.abs(ogp).max())), (protein, prec.name, var)
            ann.close()
