"""Incremental evaluation of the FP32 form: after a branch-length change
only the CLVs above the branch are recomputed, through a descriptor of the
invalid ops (the incrementalTest of LikelihoodTest.cpp:257-279, for the
batched form).  lnL and gradient must equal a full recompute bit for bit,
with and without captured graphs, for 4- and 20-state data."""
import sys

import numpy as np
import pytest

from netrax_amd import AnnotatedNetwork, NetraxOptions, Precision
from netrax_amd.simulate import make_workload, workload_30taxa_10k
from oracle import oracle as O

pytestmark = pytest.mark.gpu
FP32_RTOL = 2e-6


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _check(w, use_graph, n_moves=6):
    net, parts = w.network, w.partitions
    inc = AnnotatedNetwork(net.copy(), parts, NetraxOptions(precision=Precision.FP32, use_graph=use_graph))
    full = AnnotatedNetwork(net.copy(), parts, NetraxOptions(precision=Precision.FP32, use_graph=use_graph))
    inc.gradient()
    rng = np.random.default_rng(5)
    b = net.brlen.copy()
    for it, e in enumerate(rng.choice(net.n_edges, n_moves, replace=False)):
        b[e] = b[e] * rng.uniform(0.5, 1.8) + 1e-3
        inc.set_branch_length(int(e), float(b[e]))
        full.set_branch_lengths(b)
        if it % 2:
            li = inc.loglikelihood(incremental=1)
            lf = full.loglikelihood(incremental=0)
            assert li == lf, (it, li, lf)
        else:
            gi = inc.gradient()
            full.loglikelihood(incremental=0)
            gf = full.gradient()
            assert gi[0] == gf[0] and np.array_equal(gi[1], gf[1]) and np.array_equal(gi[2], gf[2]), it
    if net.n_retic:
        # a probability change leaves every CLV valid: no forward launch
        p = np.clip(net.retic_prob * 0.9, 0.05, 0.95)
        inc.set_reticulation_probs(p)
        full.set_reticulation_probs(p)
        li = inc.loglikelihood(incremental=1)
        assert li == full.loglikelihood(incremental=0)
        want = O.loglikelihood(net, parts, O.AVERAGE, brlen=b, probs=p)[0]
        assert abs(li - want) <= FP32_RTOL * abs(want), (li, want)
    inc.close()
    full.close()


@pytest.mark.parametrize("use_graph", [False, True])
def test_fp32_incremental_equals_full_recompute(use_graph):
    say("fp32 incremental, 30 taxa, graph", use_graph)
    _check(workload_30taxa_10k(), use_graph)


@pytest.mark.parametrize("use_graph", [False, True])
def test_fp32_protein_incremental_equals_full_recompute(use_graph):
    say("fp32 incremental, protein, graph", use_graph)
    _check(make_workload("inc_aa", 16, 2, [700], seed=41, protein=True), use_graph, n_moves=4)
