"""Writes tests/golden/golden_vectors.json.

The reference ships no golden likelihood values (its tests only check
invariants, test/src/LikelihoodTest.cpp:204-255), so the vectors here come
from tests/bruteforce.py: a brute-force sum over all internal-state
assignments of every displayed tree, with scipy's matrix exponential and
Gamma quantiles -- an implementation that shares no code with the oracle or
the engine.  Inputs are the reference's sample networks and alignments
(test/sample_networks/, copied as data) under a fixed GTR+G4 model.

Run from the repository root:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import bruteforce as BF  # noqa: E402
from fixtures import dead_retic_sample, fixed_model, load_sample  # noqa: E402

# fixtures whose displayed trees are small enough for the brute force;
# the gradient is only brute-forced where it stays cheap
CASES = [
    ("tiny", True), ("small", True), ("tree", True), ("tree_rooted", True),
    ("reticulation_in_reticulation", True), ("two_reticulations", False),
    ("clv_averaging", False), ("dead_retic", True),
]


def main():
    out = {"model": None, "cases": {}}
    model = fixed_model()
    out["model"] = {"subst_rates": list(model.subst_rates), "freqs": list(model.freqs),
                    "alpha": model.alpha, "ncat": model.ncat}
    for name, with_grad in CASES:
        t0 = time.time()
        net, parts = dead_retic_sample(model) if name == "dead_retic" else load_sample(name, model)
        rec = {"newick": net.to_newick(), "brlen": list(net.brlen),
               "retic_prob": list(net.retic_prob), "n_classes": len(BF.classes(net))}
        for var, key in ((0, "average"), (1, "best")):
            tot, per = BF.network_logl(net, parts, var)
            rec[f"lnl_{key}"] = tot
        if with_grad:
            rec["grad_average"] = list(BF.network_gradient_brlen(net, parts, 0))
        rec["edge_parent"] = [int(x) for x in net.edge_parent]
        out["cases"][name] = rec
        print(f"{name}: {rec['lnl_average']:.12f} {rec['lnl_best']:.12f} ({time.time()-t0:.1f}s)")
    with open(os.path.join(HERE, "golden_vectors.json"), "w") as fh:
        json.dump(out, fh, indent=1)


if __name__ == "__main__":
    main()
