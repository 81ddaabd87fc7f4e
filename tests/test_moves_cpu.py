"""Seeded candidate moves (arc insertion / arc removal / rSPR) (CPU)."""
import numpy as np
import pytest

from netrax_amd.moves import arc_insertion, arc_removal, rspr_tail
from netrax_amd.network import random_network
from netrax_amd.simulate import candidate_batch
from oracle import oracle as O


@pytest.mark.parametrize("seed", range(6))
def test_moves_give_valid_networks(seed):
    rng = np.random.default_rng(seed)
    net = random_network(int(rng.integers(6, 25)), int(rng.integers(1, 5)), rng)
    ins, rem, spr = arc_insertion(net, rng), arc_removal(net, rng), rspr_tail(net, rng)
    for cand, dr in ((ins, 1), (rem, -1), (spr, 0)):
        if cand is None:
            continue
        cand.validate()
        assert sorted(cand.tip_labels) == sorted(net.tip_labels)
        assert cand.n_retic == net.n_retic + dr
        for r in range(cand.n_retic):
            assert 0.0 < cand.retic_prob[r] < 1.0


def test_candidate_batch_is_seeded_and_scores():
    a = candidate_batch(2, 3, seed=4, sites=(100, 200))
    b = candidate_batch(2, 3, seed=4, sites=(100, 200))
    assert len(a) == 6
    for (n1, p1), (n2, p2) in zip(a, b):
        assert n1.to_newick() == n2.to_newick()
        assert np.array_equal(p1[0].masks, p2[0].masks)
        assert np.isfinite(O.loglikelihood(n1, p1)[0])
