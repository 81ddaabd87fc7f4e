"""Analytic gradient of the network lnL with respect to the substitution
model (exchangeabilities, frequencies, Gamma shape, category rates) on the
GPU, pinned by what the mathematics fixes:

* central finite differences of the oracle's lnL (the numerical route the
  reference's model optimisation takes, ModelOptimization.cpp:25-97) and of
  the engine's own FP64 lnL, parameter by parameter;
* scale invariance of the normalised rate matrix: scaling every
  exchangeability by one factor leaves lnL unchanged, so
  sum_k rho_k dlnL/drho_k = 0 (fails if the normalisation term is dropped);
* scaling every category rate equals scaling every branch length:
  sum_c r_c dlnL/dr_c = sum_e t_e dlnL/dt_e against the branch gradient.
"""
import dataclasses
import sys

import numpy as np
import pytest

from fixtures import fixed_model, load_sample
from netrax_amd import AnnotatedNetwork, LikelihoodVariant, NetraxOptions, Precision
from netrax_amd.simulate import make_workload
from oracle import oracle as O

pytestmark = pytest.mark.gpu

FD_RTOL = 2e-6   # analytic vs central difference, relative to the largest entry


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def eng64(net, parts, variant=0):
    return AnnotatedNetwork(net, parts, NetraxOptions(likelihood_variant=LikelihoodVariant(variant),
                                                      precision=Precision.FP64))


def _perturbed(model, field, idx, h):
    m = dataclasses.replace(model, subst_rates=np.array(model.subst_rates, float),
                            freqs=np.array(model.freqs, float))
    if field == "alpha":
        m.alpha = model.alpha + h
    else:
        getattr(m, field)[idx] += h
    return m


def _fd_oracle(net, parts, variant, p, field, idx, h):
    vals = []
    for sgn in (1, -1):
        ps = list(parts)
        ps[p] = dataclasses.replace(parts[p], model=_perturbed(parts[p].model, field, idx, sgn * h))
        vals.append(O.loglikelihood(net, ps, variant)[0])
    return (vals[0] - vals[1]) / (2 * h)


def _fd_engine(ann, parts, p, field, idx, h):
    vals = []
    for sgn in (1, -1):
        ann.set_model(p, _perturbed(parts[p].model, field, idx, sgn * h))
        vals.append(ann.loglikelihood())
    ann.set_model(p, parts[p].model)
    return (vals[0] - vals[1]) / (2 * h)


def _check_case(net, parts, variant, label):
    ann = eng64(net, parts, variant)
    lnl, grads = ann.model_gradient()
    ol = O.loglikelihood(net, parts, variant)[0]
    assert abs(lnl - ol) <= 1e-10 * abs(ol), (label, lnl, ol)
    for p, part in enumerate(parts):
        m = part.model
        g = grads[p]
        say(label, "partition", p, "subst", np.round(g["subst"], 4), "freqs", np.round(g["freqs"], 4),
            "alpha", g["alpha"])
        fd_s = np.array([_fd_oracle(net, parts, variant, p, "subst_rates", k, 1e-5 * m.subst_rates[k])
                         for k in range(len(m.subst_rates))])
        fd_f = np.array([_fd_oracle(net, parts, variant, p, "freqs", k, 1e-5 * m.freqs[k])
                         for k in range(len(m.freqs))])
        fd_a = _fd_oracle(net, parts, variant, p, "alpha", 0, 1e-5 * m.alpha)
        scale = max(1.0, np.abs(fd_s).max(), np.abs(fd_f).max(), abs(fd_a))
        assert np.abs(g["subst"] - fd_s).max() <= FD_RTOL * scale, (label, g["subst"], fd_s)
        assert np.abs(g["freqs"] - fd_f).max() <= FD_RTOL * scale, (label, g["freqs"], fd_f)
        assert abs(g["alpha"] - fd_a) <= FD_RTOL * scale, (label, g["alpha"], fd_a)
        # the engine's own lnL agrees with its gradient too
        fe = _fd_engine(ann, parts, p, "subst_rates", 0, 1e-5 * m.subst_rates[0])
        assert abs(g["subst"][0] - fe) <= FD_RTOL * scale, (label, g["subst"][0], fe)
        # scale invariance of the normalised rate matrix
        assert abs(np.dot(m.subst_rates, g["subst"])) <= 1e-8 * scale, (label, np.dot(m.subst_rates, g["subst"]))
    # rates vs branch lengths: sum_c r_c dJ/dr_c summed over partitions
    l2, db, _dp = ann.gradient()
    rates_term = sum(float(np.dot(O.gamma_rates(part.model.alpha, part.model.ncat), grads[p]["rates"]))
                     for p, part in enumerate(parts))
    br_term = float(np.dot(ann.branch_lengths(), db))
    assert abs(rates_term - br_term) <= 1e-8 * max(1.0, abs(br_term)), (label, rates_term, br_term)
    ann.close()


def test_model_gradient_two_reticulations_average_and_best():
    net, parts = load_sample("two_reticulations", fixed_model(3))
    for variant in (O.AVERAGE, O.BEST):
        say("two_reticulations variant", variant)
        _check_case(net, parts, variant, "two_reticulations/%d" % variant)


def test_model_gradient_multi_partition_workload():
    w = make_workload("mg", 16, 3, [700, 400], seed=11)
    say("16 taxa / 3 reticulations / 2 partitions")
    _check_case(w.network, w.partitions, O.AVERAGE, "16taxa")


def test_model_gradient_category_rates_and_single_category():
    net, parts = load_sample("celine_nonzero_branches", fixed_model(5))
    ann = eng64(net, parts)
    rates = np.array([0.3, 0.7, 1.1, 1.9])
    ann.set_category_rates(0, rates)
    lnl, grads = ann.model_gradient()
    assert np.isnan(grads[0]["alpha"])
    for c in range(4):
        h = 1e-6
        vals = []
        for sgn in (1, -1):
            r = rates.copy()
            r[c] += sgn * h
            ann.set_category_rates(0, r)
            vals.append(ann.loglikelihood())
        fd = (vals[0] - vals[1]) / (2 * h)
        assert abs(grads[0]["rates"][c] - fd) <= FD_RTOL * max(1.0, np.abs(grads[0]["rates"]).max()), (c, fd)
    ann.close()
    # one category: no alpha dependence
    m1 = dataclasses.replace(fixed_model(5, ncat=1))
    net, parts = load_sample("celine_nonzero_branches", m1)
    ann = eng64(net, parts)
    lnl, grads = ann.model_gradient()
    assert grads[0]["alpha"] == 0.0
    fd = _fd_oracle(net, parts, O.AVERAGE, 0, "freqs", 2, 1e-5 * m1.freqs[2])
    assert abs(grads[0]["freqs"][2] - fd) <= FD_RTOL * max(1.0, abs(fd))
    ann.close()


def test_model_gradient_rejects_fp32():
    net, parts = load_sample("tiny")
    ann = AnnotatedNetwork(net, parts, NetraxOptions(precision=Precision.FP32))
    with pytest.raises(Exception):
        ann.model_gradient()
    ann.close()
