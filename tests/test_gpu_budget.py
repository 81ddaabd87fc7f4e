"""FP64 evaluation under a memory budget (nrx_set_memory_budget): the level
schedule cut into segments, boundary CLVs kept, each segment's other CLVs
recomputed before its reverse sweep.  lnL, branch gradient and d lnL/dp must
equal the resident-CLV evaluation bit for bit, for every budget that fits.
"""
import sys

import numpy as np
import pytest

from fixtures import fixed_model, load_sample
from netrax_amd import AnnotatedNetwork, LikelihoodVariant, NetraxOptions, Precision, _lib
from netrax_amd.simulate import make_workload, workload_30taxa_10k
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def eng64(net, parts, variant=0):
    return AnnotatedNetwork(net, parts, NetraxOptions(likelihood_variant=LikelihoodVariant(variant),
                                                      precision=Precision.FP64))


def _same(a, b):
    la, ga, pa = a.gradient()
    lb, gb, pb = b.gradient()
    assert la == lb, (la, lb)
    np.testing.assert_array_equal(ga, gb)
    np.testing.assert_array_equal(pa, pb)
    assert a.loglikelihood() == b.loglikelihood()
    return la


def _budget_sweep(net, parts, variant=0):
    ref = eng64(net, parts, variant)
    bud = eng64(net, parts, variant)
    plan0 = bud.memory_plan()
    assert plan0["segments"] == 1
    full = plan0["resident_rows"] * plan0["row_bytes"]
    tried = 0
    for frac in (0.9, 0.7, 0.5, 0.4, 0.3, 0.2):
        try:
            bud.set_memory_budget(int(full * frac))
        except _lib.NrxError as ex:
            assert ex.code == _lib.NRX_ERR_ARG, (ex.code, str(ex))
            # the previous plan stays in force and still evaluates
            _same(ref, bud)
            continue
        plan = bud.memory_plan()
        say("budget", frac, plan)
        assert plan["segments"] >= 2 and plan["rows"] * plan["row_bytes"] <= full * frac
        _same(ref, bud)
        tried += 1
    return ref, bud, tried


def test_budget_gradient_is_bit_identical_30taxa():
    w = workload_30taxa_10k()
    ref, bud, tried = _budget_sweep(w.network, w.partitions)
    assert tried >= 2
    # parameter changes under the budget: lengths (incremental path) and
    # reticulation probabilities
    rng = np.random.default_rng(4)
    b = ref.branch_lengths() * rng.uniform(0.7, 1.4, w.network.n_edges)
    ref.set_branch_lengths(b)
    bud.set_branch_lengths(b)
    _same(ref, bud)
    p = np.clip(rng.uniform(0.1, 0.9, w.network.n_retic), 0.05, 0.95)
    ref.set_reticulation_probs(p)
    bud.set_reticulation_probs(p)
    lnl = _same(ref, bud)
    ol = O.loglikelihood(w.network, w.partitions, O.AVERAGE, brlen=b, probs=p)[0]
    assert abs(lnl - ol) <= 1e-9 * abs(ol)
    # calls that need every CLV refuse while the budget is active ...
    with pytest.raises(_lib.NrxError) as ex:
        bud.brlen_derivatives(3)
    assert ex.value.code == _lib.NRX_ERR_UNSUPPORTED
    with pytest.raises(_lib.NrxError) as ex:
        bud.model_gradient()
    assert ex.value.code == _lib.NRX_ERR_UNSUPPORTED
    # ... and work again without it
    bud.set_memory_budget(0)
    assert bud.memory_plan()["segments"] == 1
    _same(ref, bud)
    d_ref = ref.brlen_derivatives(3)
    d_bud = bud.brlen_derivatives(3)
    assert d_ref.logl_prime == d_bud.logl_prime and d_ref.logl_prime_prime == d_bud.logl_prime_prime
    ref.close()
    bud.close()


@pytest.mark.parametrize("variant", [0, 1])
def test_budget_gradient_is_bit_identical_reticulated_workload(variant):
    w = make_workload("bud", 24, 4, [3000, 2000], seed=31)
    ref, bud, tried = _budget_sweep(w.network, w.partitions, variant)
    assert tried >= 1
    ref.close()
    bud.close()


def test_budget_on_sample_and_too_small_budget():
    net, parts = load_sample("three_reticulations", fixed_model(2))
    ref, bud, _tried = _budget_sweep(net, parts)
    plan = bud.memory_plan()
    with pytest.raises(_lib.NrxError) as ex:
        bud.set_memory_budget(1)
    assert ex.value.code == _lib.NRX_ERR_ARG
    assert bud.memory_plan() == plan
    _same(ref, bud)
    ref.close()
    bud.close()


def test_budget_refused_for_fp32():
    net, parts = load_sample("tiny")
    a = AnnotatedNetwork(net, parts, NetraxOptions(precision=Precision.FP32))
    with pytest.raises(_lib.NrxError) as ex:
        a.set_memory_budget(1 << 20)
    assert ex.value.code == _lib.NRX_ERR_UNSUPPORTED
    a.set_memory_budget(0)
    a.close()


def test_budget_protein_is_bit_identical():
    """20 states: the reverse sweep runs per run of ops without a common
    child; under a budget the remapped op list keeps those runs."""
    w = make_workload("bud_aa", 18, 2, [600], seed=33, protein=True)
    ref, bud, tried = _budget_sweep(w.network, w.partitions)
    assert tried >= 1
    ref.close()
    bud.close()
