"""GPU engine vs the oracle / golden vectors (FP64 plain form).

Mirrors the reference's likelihood tests (test/src/LikelihoodTest.cpp):
repeated evaluation is bit-identical, incremental == full recomputation,
and every lnL/gradient is compared with the CPU oracle at 1e-9 relative
(the parity bar of the plain FP64 form).
"""
import math
import sys

import numpy as np
import pytest

from fixtures import SAMPLE_PAIRS, dead_retic_sample, fixed_model, load_golden, load_sample
from netrax_amd import AnnotatedNetwork, LikelihoodVariant, NetraxOptions, computeLoglikelihood
from netrax_amd.simulate import make_workload, workload_small_1k, workload_30taxa_10k
from oracle import oracle as O

pytestmark = pytest.mark.gpu

RTOL = 1e-9  # plain FP64 form vs the FP64 oracle
GOLD = load_golden()


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def engine(net, parts, variant=0, **kw):
    opts = NetraxOptions(likelihood_variant=LikelihoodVariant(variant), **kw)
    return AnnotatedNetwork(net, parts, opts)


def _golden_model():
    m = GOLD["model"]
    from netrax_amd.simulate import Model
    return Model(4, np.array(m["subst_rates"]), np.array(m["freqs"]), m["alpha"], m["ncat"])


@pytest.mark.parametrize("name", sorted(GOLD["cases"]))
def test_golden_vectors(name):
    say("golden", name)
    case = GOLD["cases"][name]
    model = _golden_model()
    net, parts = dead_retic_sample(model) if name == "dead_retic" else load_sample(name, model)
    for var, key in ((0, "average"), (1, "best")):
        ann = engine(net, parts, var)
        got = computeLoglikelihood(ann)
        want = case[f"lnl_{key}"]
        assert abs(got - want) <= RTOL * abs(want), (name, key, got, want)
        if key == "average" and "grad_average" in case:
            _, g, _ = ann.gradient()
            assert np.allclose(g, case["grad_average"], rtol=1e-8, atol=1e-9)
        ann.close()


@pytest.mark.parametrize("name", sorted(SAMPLE_PAIRS))
def test_samples_match_oracle(name):
    say("sample", name)
    net, parts = load_sample(name, fixed_model())
    for var in (0, 1):
        ann = engine(net, parts, var)
        got = computeLoglikelihood(ann, incremental=0)
        want = O.loglikelihood(net, parts, var)[0]
        assert abs(got - want) <= RTOL * abs(want), (name, var, got, want)
        lnl, g, gp = ann.gradient()
        _, og, ogp = O.gradient(net, parts, var)
        assert np.allclose(g, og, rtol=1e-7, atol=1e-8 * max(1.0, np.abs(og).max())), (name, var)
        assert np.allclose(gp, ogp, rtol=1e-7, atol=1e-8), (name, var)
        ann.close()


@pytest.mark.parametrize("ncat", [1, 2, 8])
def test_gradients_other_category_counts_match_oracle(ncat):
    """Every rate-category count the kernels are built for (4 is covered
    above), on the three-reticulation sample."""
    say("three_reticulations, categories", ncat)
    net, parts = load_sample("three_reticulations", fixed_model(ncat=ncat))
    for var in (0, 1):
        ann = engine(net, parts, var)
        got = computeLoglikelihood(ann, incremental=0)
        want = O.loglikelihood(net, parts, var)[0]
        assert abs(got - want) <= RTOL * abs(want), (ncat, var, got, want)
        _, g, gp = ann.gradient()
        _, og, ogp = O.gradient(net, parts, var)
        assert np.allclose(g, og, rtol=1e-7, atol=1e-8 * max(1.0, np.abs(og).max())), (ncat, var)
        assert np.allclose(gp, ogp, rtol=1e-7, atol=1e-8), (ncat, var)
        ann.close()


def test_celine_zero_branch_lengths_match_oracle():
    say("celine with clamped zero lengths")
    net, parts = load_sample("celine", fixed_model(), with_lengths=False)
    ann = engine(net, parts)
    got = computeLoglikelihood(ann)
    want = O.loglikelihood(net, parts)[0]
    assert abs(got - want) <= RTOL * abs(want)


def test_tree_logls_and_classes_match_oracle():
    net, parts = load_sample("three_reticulations", fixed_model())
    ann = engine(net, parts)
    computeLoglikelihood(ann)
    tl, lp = ann.tree_logls()
    _, _, otl, olp = O.loglikelihood(net, parts, tree_logls=True)
    assert tl.shape == otl.shape
    assert np.allclose(lp, olp, rtol=0, atol=1e-14)
    assert np.allclose(tl, otl, rtol=1e-11)


@pytest.mark.parametrize("seed", [1, 2])
def test_config_8taxa_1retic_1k_sites(seed):
    say("8 taxa / 1 reticulation / 1k sites, seed", seed)
    w = workload_small_1k(seed)
    for var in (0, 1):
        ann = engine(w.network, w.partitions, var)
        got = computeLoglikelihood(ann)
        want = O.loglikelihood(w.network, w.partitions, var)[0]
        assert abs(got - want) <= RTOL * abs(want)
        lnl, g, gp = ann.gradient()
        _, og, ogp = O.gradient(w.network, w.partitions, var)
        assert np.allclose(g, og, rtol=1e-7, atol=1e-7)
        assert np.allclose(gp, ogp, rtol=1e-7, atol=1e-7)


def test_config_30taxa_3retic_four_partitions():
    say("30 taxa / 3 reticulations / 1k+2k+3k+4k sites")
    w = workload_30taxa_10k()
    for var in (0, 1):
        ann = engine(w.network, w.partitions, var)
        got = computeLoglikelihood(ann)
        per = np.array(ann.partition_loglh)
        want, wper = O.loglikelihood(w.network, w.partitions, var)
        assert abs(got - want) <= RTOL * abs(want)
        assert np.allclose(per, wper, rtol=RTOL)
        lnl, g, gp, pg = ann.gradient(per_partition=True)
        _, og, ogp, opg = O.gradient(w.network, w.partitions, var, per_partition=True)
        scale = max(1.0, np.abs(og).max())
        assert np.allclose(g, og, rtol=1e-7, atol=1e-9 * scale)
        assert np.allclose(pg, opg, rtol=1e-7, atol=1e-9 * scale)
        assert np.allclose(gp, ogp, rtol=1e-7, atol=1e-9 * scale)


def test_repeated_evaluation_is_bit_identical():
    """LikelihoodTest.cpp:237-239: repeated computeLoglikelihood(.,0,1)."""
    w = workload_30taxa_10k()
    ann = engine(w.network, w.partitions)
    a = computeLoglikelihood(ann, 0, 1)
    b = computeLoglikelihood(ann, 0, 1)
    c = computeLoglikelihood(ann, 1, 1)
    assert a == b == c
    g1 = ann.gradient()[1]
    g2 = ann.gradient()[1]
    assert np.array_equal(g1, g2)


def test_incremental_equals_full_recompute():
    """LikelihoodTest.cpp:257-279 incrementalTest."""
    w = workload_30taxa_10k()
    ann = engine(w.network, w.partitions)
    computeLoglikelihood(ann)
    rng = np.random.default_rng(3)
    for e in rng.choice(w.network.n_edges, 5, replace=False):
        ann.set_branch_length(int(e), float(w.network.brlen[e]) * 1.7 + 0.01)
        inc = computeLoglikelihood(ann, incremental=1)
        full = computeLoglikelihood(ann, incremental=0)
        assert inc == full
        want = O.loglikelihood(w.network, w.partitions)[0]
        assert abs(inc - want) <= RTOL * abs(want)


def test_reticulation_probability_update_and_threshold():
    """updateReticulationProb (LikelihoodTest.cpp:362-377) and the
    min_interesting_tree_logprob cut (NetraxOptions.hpp:108)."""
    net, parts = load_sample("two_reticulations", fixed_model())
    ann = engine(net, parts)
    computeLoglikelihood(ann)
    for probs in ([0.3, 0.8], [1e-4, 0.5], [0.5, 1e-7]):
        ann.set_reticulation_probs(np.array(probs))
        got = computeLoglikelihood(ann)
        want = O.loglikelihood(net, parts, probs=np.array(probs))[0]
        assert abs(got - want) <= RTOL * abs(want), (probs, got, want)


def test_gradient_matches_central_differences_50taxa():
    """Full-size headline shape (50 taxa, 4 reticulations, 50k sites):
    gradient vs central differences (h = 1e-6) on a few branches and on
    every reticulation probability."""
    say("50 taxa / 4 reticulations / 50k sites FD")
    w = make_workload("50taxa_4retic_50k", 50, 4, [20000, 30000], seed=3)
    ann = engine(w.network, w.partitions)
    lnl, g, gp = ann.gradient()
    h = 1e-6
    rng = np.random.default_rng(0)
    for e in rng.choice(w.network.n_edges, 6, replace=False):
        b0 = w.network.brlen.copy()
        b = b0.copy()
        b[e] += h
        ann.set_branch_lengths(b)
        up = computeLoglikelihood(ann)
        b[e] -= 2 * h
        ann.set_branch_lengths(b)
        dn = computeLoglikelihood(ann)
        ann.set_branch_lengths(b0)
        fd = (up - dn) / (2 * h)
        assert abs(fd - g[e]) < 2e-4 * max(1.0, abs(g[e])), (e, fd, g[e])
    p0 = w.network.retic_prob.copy()
    for r in range(w.network.n_retic):
        p = p0.copy()
        p[r] += h
        ann.set_reticulation_probs(p)
        up = computeLoglikelihood(ann)
        p[r] -= 2 * h
        ann.set_reticulation_probs(p)
        dn = computeLoglikelihood(ann)
        ann.set_reticulation_probs(p0)
        fd = (up - dn) / (2 * h)
        assert abs(fd - gp[r]) < 2e-4 * max(1.0, abs(gp[r])), (r, fd, gp[r])


def test_average_bounds_best():
    w = workload_30taxa_10k()
    a = computeLoglikelihood(engine(w.network, w.partitions, 0))
    b = computeLoglikelihood(engine(w.network, w.partitions, 1))
    assert a >= b


# ---------------------------------------------------------------- FP32 form
# Batched form: FP32 CLVs with power-of-two per-site exponents, FP64 sums.
FP32_RTOL = 1e-6  # |dlnL| <= 1e-6 relative to the FP64 result


def engine32(net, parts, variant=0):
    from netrax_amd import Precision
    return engine(net, parts, variant, precision=Precision.FP32)


@pytest.mark.parametrize("name", sorted(SAMPLE_PAIRS))
def test_fp32_samples_match_oracle(name):
    say("fp32 sample", name)
    net, parts = load_sample(name, fixed_model())
    for var in (0, 1):
        ann = engine32(net, parts, var)
        got = computeLoglikelihood(ann)
        want = O.loglikelihood(net, parts, var)[0]
        assert abs(got - want) <= FP32_RTOL * abs(want), (name, var, got, want)
        lnl, g, gp = ann.gradient()
        _, og, ogp = O.gradient(net, parts, var)
        scale = max(1.0, np.abs(og).max())
        assert np.allclose(g, og, rtol=1e-4, atol=1e-4 * scale), (name, var, np.abs(g - og).max())
        assert np.allclose(gp, ogp, rtol=1e-4, atol=1e-4 * max([1.0, *np.abs(ogp)]))
        ann.close()


@pytest.mark.parametrize("wl", ["8taxa", "30taxa", "50taxa"])
def test_fp32_configs_match_fp64(wl):
    say("fp32 vs fp64", wl)
    from netrax_amd.simulate import workload_50taxa_50k
    w = {"8taxa": workload_small_1k, "30taxa": workload_30taxa_10k,
         "50taxa": workload_50taxa_50k}[wl]()
    for var in (0, 1):
        a64 = engine(w.network, w.partitions, var)
        a32 = engine32(w.network, w.partitions, var)
        l64, g64, p64 = a64.gradient()
        l32, g32, p32 = a32.gradient()
        assert abs(l32 - l64) <= FP32_RTOL * abs(l64), (var, l32, l64)
        scale = max(1.0, np.abs(g64).max())
        assert np.abs(g32 - g64).max() <= 1e-4 * scale, (var, np.abs(g32 - g64).max(), scale)
        assert np.abs(p32 - p64).max(initial=0.0) <= 1e-4 * max([1.0, *np.abs(p64)])
        a64.close()
        a32.close()


def test_fp32_deep_underflow_matches_closed_form():
    """Saturated branches on 200 tips: site likelihoods ~1e-120 need the
    per-site exponents; lnL = sum of log stationary frequencies."""
    from netrax_amd.network import random_network
    from netrax_amd.simulate import Model, Partition
    rng = np.random.default_rng(11)
    net = random_network(200, 0, rng)
    net.brlen[:] = 100.0
    m = Model(4, np.ones(6), np.array([0.1, 0.2, 0.3, 0.4]), 1.0, 1)
    states = rng.integers(4, size=(200, 100))
    part = Partition(m, (1 << states).astype(np.uint32), np.ones(100, np.uint32))
    want = float(np.log(m.freqs)[states].sum())
    for eng in (engine, engine32):
        got = computeLoglikelihood(eng(net, [part]))
        assert abs(got - want) <= 1e-6 * abs(want), (got, want)


# ------------------------------------------------------ branch derivatives
# computePartitionSumtables + computeLoglikelihoodDerivatives
# (LikelihoodDerivatives.cpp:30-344): reference normalisation (mode 0) and
# the true derivative (mode 1), AVERAGE and BEST, vs the oracle.

@pytest.mark.parametrize("name", ["small", "two_reticulations", "three_reticulations",
                                  "reticulation_in_reticulation", "celine_smaller_1", "tree"])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_brlen_derivatives_match_oracle(name, prec):
    say("brlen derivatives", name, prec)
    from netrax_amd import Precision
    net, parts = load_sample(name, fixed_model())
    tol = 1e-7 if prec == "fp64" else 2e-4
    for var in (0, 1):
        ann = engine(net, parts, var, precision=Precision.FP64 if prec == "fp64" else Precision.FP32)
        for e in range(net.n_edges):
            for mode in (0, 1):
                d = ann.brlen_derivatives(e, mode=mode)
                o1, o2, op1, op2 = O.brlen_derivatives(net, parts, e, var, mode=mode)
                assert abs(d.logl_prime - o1) <= tol * max(1.0, abs(o1)), (name, var, e, mode, d.logl_prime, o1)
                assert abs(d.logl_prime_prime - o2) <= tol * max(1.0, abs(o2)), (name, var, e, mode, d.logl_prime_prime, o2)
        ann.close()


def test_brlen_derivatives_at_moved_length_and_reuse():
    """Newton-Raphson evaluates the same sumtables at several lengths."""
    w = workload_30taxa_10k()
    ann = engine(w.network, w.partitions)
    computeLoglikelihood(ann)
    rng = np.random.default_rng(5)
    for e in rng.choice(w.network.n_edges, 4, replace=False):
        b0 = float(w.network.brlen[e])
        for b in (b0, b0 * 0.5 + 1e-4, b0 * 2.0 + 0.01):
            for mode in (0, 1):
                d = ann.brlen_derivatives(int(e), brlen=b, mode=mode)
                bl = w.network.brlen.copy()
                bl[e] = b
                o1, o2, op1, op2 = O.brlen_derivatives(w.network, w.partitions, int(e), 0, mode=mode,
                                                      brlen=bl)
                assert abs(d.logl_prime - o1) <= 1e-7 * max(1.0, abs(o1)), (e, b, mode)
                assert abs(d.logl_prime_prime - o2) <= 1e-7 * max(1.0, abs(o2)), (e, b, mode)
                assert np.allclose(d.partition_logl_prime, op1, rtol=1e-7, atol=1e-7)
        # moving only this branch keeps the sumtables valid
        ann.set_branch_length(int(e), b0 * 1.3)
        d = ann.brlen_derivatives(int(e), mode=1)
        bl = w.network.brlen.copy()
        o1, _, _, _ = O.brlen_derivatives(w.network, w.partitions, int(e), 0, mode=1, brlen=bl)
        assert abs(d.logl_prime - o1) <= 1e-7 * max(1.0, abs(o1))


def test_true_derivative_matches_gradient_50taxa():
    w = make_workload("50taxa_4retic_50k", 50, 4, [20000, 30000], seed=3)
    ann = engine(w.network, w.partitions)
    _, g, _ = ann.gradient()
    for e in (0, 7, 33, w.network.n_edges - 1):
        d = ann.brlen_derivatives(e, mode=1)
        assert abs(d.logl_prime - g[e]) <= 1e-7 * max(1.0, abs(g[e])), (e, d.logl_prime, g[e])


def test_branch_length_sweep_improves_and_matches_oracle_driven_sweep():
    """optimize_branches (BranchLengthOptimization.cpp:422-477) on perturbed
    lengths: lnL never decreases, and the lengths equal those of the same
    sweep driven by the oracle's derivatives."""
    from netrax_amd.optimization import newton_maximise, optimize_branches
    net, parts = load_sample("two_reticulations", fixed_model())
    rng = np.random.default_rng(2)
    net.brlen = net.brlen * rng.uniform(0.3, 3.0, net.n_edges)
    start = net.brlen.copy()
    ann = engine(net, parts)
    before = computeLoglikelihood(ann)
    after = optimize_branches(ann)
    assert after >= before - 1e-9
    assert abs(after - O.loglikelihood(net, parts, brlen=net.brlen)[0]) <= RTOL * abs(after)
    # the same sweep with oracle derivatives and oracle acceptance checks
    b = start.copy()
    for e in range(net.n_edges):
        def deriv(x):
            bb = b.copy()
            bb[e] = x
            d1, d2, _, _ = O.brlen_derivatives(net, parts, e, O.AVERAGE, mode=0, brlen=bb)
            return d1, d2
        old = O.loglikelihood(net, parts, brlen=b)[0]
        x = newton_maximise(deriv, float(b[e]), 1e-6, 100.0, 1e-7, 32)
        bb = b.copy()
        bb[e] = x
        if O.loglikelihood(net, parts, brlen=bb)[0] >= old:
            b = bb
    assert np.allclose(net.brlen, b, rtol=1e-6, atol=1e-9), (net.brlen, b)


def test_cpp_mirror_example_matches_oracle():
    """C++ program over include/netrax_likelihood.hpp vs the oracle."""
    import subprocess
    from netrax_amd import build as nb
    from netrax_amd.simulate import Model
    res = subprocess.run([nb.EXAMPLE], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    rows = [list(map(float, l.split())) for l in res.stdout.strip().splitlines()]
    net, parts = load_sample("tiny", Model(4, np.array([1., 2, 1, 1, 2, 1]),
                                           np.array([.1, .2, .3, .4]), 0.7, 4))
    net.brlen = np.array([2.0, 1.0, 2.0, 2.0, 2.0, 0.1, 0.2])
    net.retic_prob = np.array([0.3])
    for var, (lnl, d1, d2) in zip((0, 1), rows):
        want = O.loglikelihood(net, parts, var)[0]
        assert abs(lnl - want) <= RTOL * abs(want)
        o1, o2, _, _ = O.brlen_derivatives(net, parts, 1, var, mode=0)
        assert abs(d1 - o1) <= 1e-7 * max(1, abs(o1)) and abs(d2 - o2) <= 1e-7 * max(1, abs(o2))


# ------------------------------------------------------------- protein data
def test_protein_fp64_matches_oracle():
    """20-state reversible model (Gamma(1,1) exchangeabilities, Dirichlet
    frequencies), Gamma4: 40 taxa / 2 reticulations on a 1,500-site sample
    of the protein workload's shape."""
    say("protein 40 taxa / 2 reticulations")
    w = make_workload("40taxa_2retic_aa", 40, 2, [1500], seed=4, protein=True)
    for var in (0, 1):
        ann = engine(w.network, w.partitions, var)
        got = computeLoglikelihood(ann)
        want = O.loglikelihood(w.network, w.partitions, var)[0]
        assert abs(got - want) <= RTOL * abs(want), (var, got, want)
        lnl, g, gp = ann.gradient()
        _, og, ogp = O.gradient(w.network, w.partitions, var)
        scale = max(1.0, np.abs(og).max())
        assert np.allclose(g, og, rtol=1e-7, atol=1e-9 * scale)
        assert np.allclose(gp, ogp, rtol=1e-7, atol=1e-9 * scale)
        for e in (0, 5, w.network.n_edges - 1):
            d = ann.brlen_derivatives(e, mode=0)
            o1, o2, _, _ = O.brlen_derivatives(w.network, w.partitions, e, var, mode=0)
            assert abs(d.logl_prime - o1) <= 1e-7 * max(1.0, abs(o1))
        ann.close()


def test_protein_fp32_matrix_cores_match_oracle():
    """20-state FP32 form (CLV x P on v_mfma_f32_32x32x2f32): two uneven
    partitions (tiles padded to 32/64 sites), AVERAGE and BEST, lnL within
    1e-6 relative and the full gradient vs the FP64 oracle."""
    say("protein fp32 40 taxa / 2 reticulations, partitions 700 + 333")
    w = make_workload("40taxa_2retic_aa", 40, 2, [700, 333], seed=4, protein=True)
    for var in (0, 1):
        ann = engine32(w.network, w.partitions, var)
        got = computeLoglikelihood(ann)
        want = O.loglikelihood(w.network, w.partitions, var)[0]
        assert abs(got - want) <= FP32_RTOL * abs(want), (var, got, want)
        lnl, g, gp = ann.gradient()
        _, og, ogp = O.gradient(w.network, w.partitions, var)
        scale = max(1.0, np.abs(og).max())
        assert np.abs(g - og).max() <= 1e-4 * scale, (var, np.abs(g - og).max(), scale)
        assert np.abs(gp - ogp).max(initial=0.0) <= 1e-4 * max([1.0, *np.abs(ogp)])
        ann.close()


def test_protein_fp32_full_size_matches_fp64():
    """The protein workload at full size (40 taxa, 2 reticulations, 20,000
    sites): FP32 matrix-core form vs the FP64 form (itself pinned to the
    oracle above)."""
    from netrax_amd.simulate import workload_protein_20k
    say("protein fp32 vs fp64 at 20k sites")
    w = workload_protein_20k()
    for var in (0, 1):
        a64 = engine(w.network, w.partitions, var)
        a32 = engine32(w.network, w.partitions, var)
        l64, g64, p64 = a64.gradient()
        l32, g32, p32 = a32.gradient()
        assert abs(l32 - l64) <= FP32_RTOL * abs(l64), (var, l32, l64)
        scale = max(1.0, np.abs(g64).max())
        assert np.abs(g32 - g64).max() <= 1e-4 * scale, (var, np.abs(g32 - g64).max(), scale)
        assert np.abs(p32 - p64).max(initial=0.0) <= 1e-4 * max([1.0, *np.abs(p64)])
        a64.close()
        a32.close()


def test_protein_fp32_deep_underflow_matches_closed_form():
    """Saturated branches on 150 tips with 20 states and 4 categories
    (equal exchangeabilities, alpha = 50 so every category rate is near 1):
    the site likelihoods underflow FP32 many times over; lnL = sum log pi."""
    from netrax_amd.network import random_network
    from netrax_amd.simulate import Model, Partition
    rng = np.random.default_rng(12)
    net = random_network(150, 0, rng)
    net.brlen[:] = 100.0
    freqs = rng.dirichlet(np.ones(20))
    m = Model(20, np.ones(190), freqs, 50.0, 4)
    states = rng.integers(20, size=(150, 90))
    part = Partition(m, (1 << states).astype(np.uint32), np.ones(90, np.uint32))
    want = float(np.log(freqs)[states].sum())
    got = computeLoglikelihood(engine32(net, [part]))
    assert abs(got - want) <= 1e-6 * abs(want), (got, want)


# ---------------------------------------------------------- network batches
def test_network_batch_matches_single_engines_and_oracle():
    """One launch per kernel for a batch of candidate networks (different
    sizes, reticulation counts and sites) == each network scored alone."""
    from netrax_amd import NetworkBatch, Precision
    from netrax_amd.simulate import candidate_batch
    cands = candidate_batch(3, 3, seed=6, sites=(300, 3000))
    anns = [engine(n, p, 0, precision=Precision.FP32) for n, p in cands]
    batch = NetworkBatch(anns)
    for var in (0, 1):
        lnl, dbs, dps, ms = batch.evaluate(var)
        assert ms > 0
        for i, (n, p) in enumerate(cands):
            single = engine(n, p, var, precision=Precision.FP32)
            l1, g1, p1 = single.gradient()
            assert abs(lnl[i] - l1) <= 1e-12 * abs(l1), (i, lnl[i], l1)
            assert np.allclose(dbs[i], g1, rtol=1e-9, atol=1e-9)
            assert np.allclose(dps[i], p1, rtol=1e-9, atol=1e-9)
            want = O.loglikelihood(n, p, var)[0]
            assert abs(lnl[i] - want) <= FP32_RTOL * abs(want)
            single.close()
    # parameters set between calls are picked up
    n0, p0 = cands[0]
    b = anns[0].network.brlen * 1.5
    anns[0].set_branch_lengths(b)
    lnl2, _, _, _ = batch.evaluate(0, gradient=False)
    want = O.loglikelihood(n0, p0, 0, brlen=b)[0]
    assert abs(lnl2[0] - want) <= FP32_RTOL * abs(want)
    batch.close()
