"""Shared test inputs: the reference's sample networks and alignments
(copied as data into tests/golden/sample_networks/) and fixed models."""
from __future__ import annotations

import json
import os

import numpy as np

from netrax_amd.alignment import read_fasta
from netrax_amd.network import parse_newick
from netrax_amd.simulate import Model, Partition, dna_masks

HERE = os.path.dirname(os.path.abspath(__file__))
SAMPLES = os.path.join(HERE, "golden", "sample_networks")
GOLDEN = os.path.join(HERE, "golden", "golden_vectors.json")

# (network file, alignment file) pairs used by LikelihoodTest.cpp:291-360
SAMPLE_PAIRS = {
    "tiny": ("tiny.nw", "tiny_fake_alignment.txt"),
    "small": ("small.nw", "small_fake_alignment.txt"),
    "tree": ("tree.nw", "small_fake_alignment.txt"),
    "tree_rooted": ("tree_rooted.nw", "small_fake_alignment.txt"),
    "clv_averaging": ("clv_averaging.nw", "5_taxa_fake_alignment.txt"),
    "two_reticulations": ("two_reticulations.nw", "5_taxa_fake_alignment.txt"),
    "three_reticulations": ("three_reticulations.nw", "7_taxa_fake_alignment.txt"),
    "interleaved_reticulations": ("interleaved_reticulations.nw", "5_taxa_fake_alignment.txt"),
    "reticulation_in_reticulation": ("reticulation_in_reticulation.nw", "small_fake_alignment.txt"),
    "celine": ("celine.nw", "celine_fake_alignment.txt"),
    "celine_nonzero_branches": ("celine_nonzero_branches.nw", "celine_fake_alignment.txt"),
    "celine_smaller_1": ("celine_smaller_1.nw", "celine_fake_alignment_smaller.txt"),
}

# A reticulation whose only child is another reticulation: when the child
# picks its other parent the first one is dead and its choice no longer
# matters, so two configurations share one displayed tree.
DEAD_RETIC_NEWICK = "((A:0.2,((B:0.1)#H2:0.15::0.4)#H1:0.1::0.7):0.1,(#H1:0.3::0.3,(#H2:0.2::0.6,C:0.25):0.1):0.2);"


def read_alignment(fname: str) -> dict:
    """One of the reference's sample alignments (FASTA), by file name."""
    return read_fasta(os.path.join(SAMPLES, fname))


def fixed_model(seed: int = 7, ncat: int = 4) -> Model:
    rng = np.random.default_rng(seed)
    return Model(4, rng.dirichlet(np.ones(6)) * 6, rng.dirichlet(np.ones(4) * 3),
                 float(rng.uniform(0.4, 1.5)), ncat)


def default_brlens(net):
    """Deterministic lengths for fixtures that carry none (all clamped)."""
    if np.all(net.brlen <= 1e-6):
        net.brlen = np.linspace(0.05, 0.3, net.n_edges)
    return net


def load_sample(name: str, model: Model = None, with_lengths: bool = True):
    nw, al = SAMPLE_PAIRS[name]
    with open(os.path.join(SAMPLES, nw)) as fh:
        net = parse_newick(fh.read().strip().split("\n")[0])
    if with_lengths:
        default_brlens(net)
    aln = read_alignment(al)
    masks = np.stack([dna_masks(aln[l]) for l in net.tip_labels])
    part = Partition(model or fixed_model(), masks, np.ones(masks.shape[1], np.uint32))
    return net, [part]


def dead_retic_sample(model: Model = None):
    net = parse_newick(DEAD_RETIC_NEWICK)
    aln = read_alignment("tiny_fake_alignment.txt")
    masks = np.stack([dna_masks(aln[l]) for l in net.tip_labels])
    return net, [Partition(model or fixed_model(), masks, np.ones(masks.shape[1], np.uint32))]


def load_golden() -> dict:
    with open(GOLDEN) as fh:
        return json.load(fh)
