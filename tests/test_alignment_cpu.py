"""Alignment input (netrax_amd/alignment.py): FASTA / PHYLIP reading,
RAxML-style partition files, model strings and pattern compression.  CPU
only; likelihood invariance is checked with the oracle."""
import os

import numpy as np
import pytest

from fixtures import SAMPLE_PAIRS, SAMPLES, fixed_model, load_sample
from netrax_amd.alignment import (compress_patterns, load_partitions, parse_model, read_alignment,
                                  read_fasta, read_partition_file, read_phylip)
from netrax_amd.simulate import Partition
from oracle import oracle as O


def _raw_fasta(path):
    """Second, line-by-line reading of a sample FASTA file."""
    labels, seqs = [], []
    with open(path) as fh:
        for line in fh:
            line = line.rstrip("\n")
            if line.startswith(">"):
                labels.append(line[1:])
                seqs.append("")
            elif line.strip():
                seqs[-1] += line.strip()
    return dict(zip(labels, seqs))


@pytest.mark.parametrize("fname", sorted({a for _, a in SAMPLE_PAIRS.values()}))
def test_reads_every_sample_alignment(fname):
    aln = read_fasta(os.path.join(SAMPLES, fname))
    assert aln == _raw_fasta(os.path.join(SAMPLES, fname))
    assert read_alignment(os.path.join(SAMPLES, fname)) == aln


def test_phylip_round_trip():
    aln = read_fasta(os.path.join(SAMPLES, "5_taxa_fake_alignment.txt"))
    n = len(next(iter(aln.values())))
    text = f"{len(aln)} {n}\n" + "".join(f"{k}  {v[:7]}\n{v[7:]}\n" for k, v in aln.items())
    assert read_phylip(text) == aln
    assert read_alignment(text) == aln


def test_fasta_errors():
    with pytest.raises(ValueError):
        read_fasta(">A\nACGT\n>B\nACG\n")
    with pytest.raises(ValueError):
        read_fasta(">A\nACGT\n>A\nACGT\n")
    with pytest.raises(ValueError):
        read_fasta("ACGT\n")


def test_partition_file_ranges_and_strides():
    specs = read_partition_file("DNA, a = 1-4, 9\nGTR+G8, b=5-8\nJC, c = 10-15/2\n"
                                "GTR+G, d = 11-15\\2 # comment\n", n_sites=15)
    assert [s.name for s in specs] == ["a", "b", "c", "d"]
    assert list(specs[0].columns) == [0, 1, 2, 3, 8]
    assert list(specs[1].columns) == [4, 5, 6, 7]
    assert list(specs[2].columns) == [9, 11, 13]
    assert list(specs[3].columns) == [10, 12, 14]
    assert specs[1].model.ncat == 8 and specs[3].model.ncat == 4 and specs[0].model.ncat == 1
    assert specs[2].model.equal_rates and specs[2].model.equal_freqs


@pytest.mark.parametrize("text", ["DNA, a = 1-10\nDNA, b = 10-15\n",   # overlap
                                  "DNA, a = 1-10\n",                    # not covered
                                  "DNA, a = 1-16\n",                    # beyond the alignment
                                  "DNA a = 1-15\n",                     # no comma
                                  "WAG, a = 1-15\n",                    # unsupported model
                                  "DNA, a = 5-1\n"])                    # empty range
def test_partition_file_errors(text):
    with pytest.raises(ValueError):
        read_partition_file(text, n_sites=15)


def test_model_strings():
    m = parse_model("GTR+G4")
    assert (m.states, m.ncat, m.equal_rates) == (4, 4, False)
    assert parse_model("DNA").ncat == 1
    p = parse_model("PROT+G")
    assert (p.states, p.ncat, p.equal_rates) == (20, 4, True)
    with pytest.raises(ValueError):
        parse_model("GTR+I")


def test_compression_keeps_first_appearance_order():
    m = np.array([[1, 2, 1, 4, 2, 1], [8, 8, 8, 1, 8, 8]], np.uint32)
    cols, w = compress_patterns(m)
    assert cols.tolist() == [[1, 2, 4], [8, 8, 1]]
    assert w.tolist() == [3, 2, 1]
    scols, sw = compress_patterns(m, order="sorted")
    assert sorted(map(tuple, cols.T.tolist())) == list(map(tuple, scols.T.tolist()))
    assert int(sw.sum()) == 6


@pytest.mark.parametrize("name", ["two_reticulations", "three_reticulations", "celine_smaller_1"])
def test_partitioned_input_reproduces_fixture_columns_and_lnl(name, tmp_path):
    """A sample alignment split by a partition file gives the fixture's
    state masks column for column; compressing patterns leaves every
    partition's lnL unchanged."""
    net, parts = load_sample(name, fixed_model())
    full = parts[0].masks
    n = full.shape[1]
    pf = tmp_path / "parts.txt"
    pf.write_text(f"GTR+G4, first = 1-{n // 2}\nGTR+G4, second = {n // 2 + 1}-{n}\n")
    msa = os.path.join(SAMPLES, SAMPLE_PAIRS[name][1])
    raw = load_partitions(msa, net.tip_labels, str(pf), compress=False)
    assert np.array_equal(raw[0].masks, full[:, :n // 2])
    assert np.array_equal(raw[1].masks, full[:, n // 2:])
    packed = load_partitions(msa, net.tip_labels, str(pf), compress=True)
    for r, p in zip(raw, packed):
        assert int(p.weights.sum()) == r.masks.shape[1]
        assert p.masks.shape[1] <= r.masks.shape[1]
        m = fixed_model()
        a = O.loglikelihood(net, [Partition(m, r.masks, r.weights)])[0]
        b = O.loglikelihood(net, [Partition(m, p.masks, p.weights)])[0]
        assert abs(a - b) <= 1e-11 * abs(a), (a, b)
    # the models start from empirical frequencies of each partition
    assert all(abs(p.model.freqs.sum() - 1.0) < 1e-12 for p in packed)


def test_missing_taxon_is_reported():
    net, _ = load_sample("two_reticulations", fixed_model())
    msa = os.path.join(SAMPLES, "small_fake_alignment.txt")
    with pytest.raises(ValueError):
        load_partitions(msa, net.tip_labels)
