"""Seeded synthetic inputs: substitution models, alignments and the
benchmark workloads (network / alignment / partition shapes).

This module only produces inputs.  Likelihood arithmetic lives in the HIP
engine; the numpy/scipy helpers here (matrix exponential, discrete Gamma)
are used to *simulate* sequences, never to score them.
"""
from __future__ import annotations

import dataclasses
from typing import List, Optional

import numpy as np

from .network import Network, active_edges, random_network, topo_order

DNA_CHARS = "ACGT"
# IUPAC nucleotide codes as state masks (bit0=A, bit1=C, bit2=G, bit3=T);
# gaps and unknowns are all-ones, as in libpll's pll_map_nt.
_DNA_MASK = {
    "A": 1, "C": 2, "G": 4, "T": 8, "U": 8,
    "R": 5, "Y": 10, "S": 6, "W": 9, "K": 12, "M": 3,
    "B": 14, "D": 13, "H": 11, "V": 7,
    "N": 15, "O": 15, "X": 15, "-": 15, "?": 15, ".": 15,
}
AA_CHARS = "ARNDCQEGHILKMFPSTWYV"


def dna_masks(seq: str) -> np.ndarray:
    return np.array([_DNA_MASK[ch.upper()] for ch in seq], np.uint32)


def aa_masks(seq: str) -> np.ndarray:
    full = (1 << 20) - 1
    out = []
    for ch in seq.upper():
        if ch in AA_CHARS:
            out.append(1 << AA_CHARS.index(ch))
        elif ch == "B":
            out.append((1 << AA_CHARS.index("N")) | (1 << AA_CHARS.index("D")))
        elif ch == "Z":
            out.append((1 << AA_CHARS.index("Q")) | (1 << AA_CHARS.index("E")))
        else:
            out.append(full)
    return np.array(out, np.uint32)


@dataclasses.dataclass
class Model:
    """Reversible substitution model with discrete Gamma rate categories.

    ``subst_rates`` are the S(S-1)/2 exchangeabilities in row-major upper
    triangle order (AC, AG, AT, CG, CT, GT for DNA), ``freqs`` the
    stationary frequencies, ``alpha`` the Gamma shape, ``ncat`` the number
    of equiprobable mean-rate categories.
    """

    states: int
    subst_rates: np.ndarray
    freqs: np.ndarray
    alpha: float
    ncat: int = 4

    def rate_matrix(self) -> np.ndarray:
        S = self.states
        Q = np.zeros((S, S))
        k = 0
        for i in range(S):
            for j in range(i + 1, S):
                Q[i, j] = self.subst_rates[k] * self.freqs[j]
                Q[j, i] = self.subst_rates[k] * self.freqs[i]
                k += 1
        np.fill_diagonal(Q, -Q.sum(axis=1))
        mu = -np.dot(self.freqs, np.diag(Q))
        return Q / mu


def gamma_mean_rates(alpha: float, ncat: int) -> np.ndarray:
    """Yang (1994) mean rates of ``ncat`` equiprobable Gamma categories
    (used for simulation only)."""
    from scipy.special import gammaincinv, gammainc

    if ncat == 1:
        return np.ones(1)
    k = np.arange(1, ncat)
    cuts = gammaincinv(alpha, k / ncat) / alpha  # quantiles of Gamma(alpha, 1/alpha)
    cdf1 = np.concatenate([[0.0], gammainc(alpha + 1, cuts * alpha), [1.0]])
    rates = np.diff(cdf1) * ncat
    return rates


def random_dna_model(rng: np.random.Generator, ncat: int = 4) -> Model:
    return Model(4, rng.dirichlet(np.ones(6)), rng.dirichlet(np.ones(4)),
                 float(rng.uniform(0.3, 2.0)), ncat)


def random_protein_model(rng: np.random.Generator, ncat: int = 4) -> Model:
    return Model(20, rng.gamma(1.0, 1.0, 190), rng.dirichlet(np.ones(20)),
                 float(rng.uniform(0.3, 2.0)), ncat)


def simulate_states(net: Network, model: Model, n_sites: int, rng: np.random.Generator) -> np.ndarray:
    """Evolve ``n_sites`` sites down the network; each site follows a
    displayed tree drawn with the reticulation probabilities and a Gamma
    category drawn uniformly.  Returns int states [n_tips, n_sites]."""
    from scipy.linalg import expm

    S = model.states
    Q = model.rate_matrix()
    rates = gamma_mean_rates(model.alpha, model.ncat)
    configs = np.zeros(n_sites, np.int64)
    for r in range(net.n_retic):
        second = rng.uniform(size=n_sites) >= net.retic_prob[r]
        configs |= second.astype(np.int64) << r
    cats = rng.integers(model.ncat, size=n_sites)
    order = topo_order(net)
    parent_edge = {}
    out = np.zeros((net.n_tips, n_sites), np.int64)
    for cfg in np.unique(configs):
        act = active_edges(net, int(cfg))
        parent_edge.clear()
        for e in range(net.n_edges):
            if act[e]:
                parent_edge[int(net.edge_child[e])] = e
        for k in range(model.ncat):
            idx = np.nonzero((configs == cfg) & (cats == k))[0]
            if len(idx) == 0:
                continue
            state = {}
            state[net.root] = rng.choice(S, size=len(idx), p=model.freqs)
            for v in order:
                if v == net.root:
                    continue
                e = parent_edge[v]
                P = expm(Q * rates[k] * net.brlen[e])
                cum = np.cumsum(P, axis=1)
                cum[:, -1] = 1.0
                ps = state[int(net.edge_parent[e])]
                u = rng.uniform(size=len(idx))
                state[v] = (u[:, None] > cum[ps]).sum(axis=1)
            for t in range(net.n_tips):
                out[t, idx] = state[t]
    return out


def compress_patterns(masks: np.ndarray):
    """Unique site columns with integer weights (RaxmlWrapper.cpp:330 turns
    pattern compression on).  ``masks`` is [n_tips, n_sites]; patterns in
    lexicographic column order, which fixes the seeded workloads."""
    from .alignment import compress_patterns as _compress

    return _compress(masks, order="sorted")


@dataclasses.dataclass
class Partition:
    model: Model
    masks: np.ndarray  # uint32 [n_tips, n_patterns] state masks
    weights: np.ndarray  # uint32 [n_patterns]

    @property
    def n_sites(self) -> int:
        return int(self.masks.shape[1])

    @property
    def states(self) -> int:
        return self.model.states


@dataclasses.dataclass
class Workload:
    name: str
    network: Network
    partitions: List[Partition]


def states_to_masks(states: np.ndarray) -> np.ndarray:
    return (np.uint32(1) << states.astype(np.uint32)).astype(np.uint32)


def make_workload(name: str, n_taxa: int, n_retic: int, part_sites: List[int], seed: int,
                  protein: bool = False, shared_model: bool = False, compress: bool = True) -> Workload:
    """Seeded network + per-partition model + simulated alignment."""
    rng = np.random.default_rng(seed)
    net = random_network(n_taxa, n_retic, rng)
    parts = []
    shared = None
    for ns in part_sites:
        if shared_model and shared is not None:
            model = shared
        else:
            model = random_protein_model(rng) if protein else random_dna_model(rng)
            shared = model
        st = simulate_states(net, model, ns, rng)
        masks = states_to_masks(st)
        if compress:
            masks, weights = compress_patterns(masks)
        else:
            weights = np.ones(masks.shape[1], np.uint32)
        parts.append(Partition(model, masks, weights))
    return Workload(name, net, parts)


# The five workloads of the benchmark plan, by name.
def workload_small_1k(seed: int = 1) -> Workload:
    """8 taxa, 1 reticulation, 1,000 DNA sites, one GTR+G4 partition."""
    return make_workload("8taxa_1retic_1k", 8, 1, [1000], seed)


def workload_30taxa_10k(seed: int = 2) -> Workload:
    """30 taxa, 3 reticulations, 10,000 DNA sites in 1k/2k/3k/4k partitions."""
    return make_workload("30taxa_3retic_10k", 30, 3, [1000, 2000, 3000, 4000], seed)


def workload_50taxa_50k(seed: int = 3) -> Workload:
    """50 taxa, 4 reticulations, 50,000 DNA sites in 20k/30k partitions."""
    return make_workload("50taxa_4retic_50k", 50, 4, [20000, 30000], seed)


def workload_protein_20k(seed: int = 4) -> Workload:
    """40 taxa, 2 reticulations, 20,000 amino-acid sites, one partition."""
    return make_workload("40taxa_2retic_20k_aa", 40, 2, [20000], seed, protein=True)


def candidate_batch(n_bases: int = 8, per_base: int = 32, seed: int = 5,
                    taxa=(20, 60), retic=(0, 5), sites=(2000, 50000)):
    """Candidate networks of search steps: ``n_bases`` seeded base networks
    (taxa uniform in ``taxa``, reticulations uniform in ``retic``, sites
    log-uniform in ``sites``, GTR+G4, alignment simulated on the base), each
    followed by ``per_base`` candidates from one seeded move (arc insertion,
    arc removal or rSPR).  Returns a list of (Network, partitions)."""
    from .moves import random_move

    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_bases):
        nt = int(rng.integers(taxa[0], taxa[1] + 1))
        nr = int(rng.integers(retic[0], retic[1] + 1))
        ns = int(round(float(np.exp(rng.uniform(np.log(sites[0]), np.log(sites[1]))))))
        base = random_network(nt, nr, rng)
        model = random_dna_model(rng)
        masks, weights = compress_patterns(states_to_masks(simulate_states(base, model, ns, rng)))
        parts = [Partition(model, masks, weights)]
        for _ in range(per_base):
            cand = random_move(base, rng, max_retic=retic[1])
            # the alignment rows follow the candidate's tip order
            order = [base.tip_labels.index(l) for l in cand.tip_labels]
            out.append((cand, [Partition(model, np.ascontiguousarray(masks[order]), weights)]))
    return out
