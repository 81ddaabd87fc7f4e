"""Alignment input for the engine: FASTA / relaxed PHYLIP reading, RAxML-style
partition files, model strings and site-pattern compression.

In the reference these inputs reach the likelihood path through raxml-ng
(the MSA file and ``model_file`` options, src/NetraxOptions.hpp:136-137, read
when the treeinfo is built, src/RaxmlWrapper.cpp:539-669, with pattern
compression on, RaxmlWrapper.cpp:330).  raxml-ng itself is not in the
reference tree, so the formats follow its published documentation and the
reference's own writers:

* FASTA: ``>label`` lines followed by sequence lines (the reference's
  fixtures, test/sample_networks/*_fake_alignment.txt);
* partition file: one partition per line, ``MODEL, NAME = RANGES`` with
  RANGES a comma-separated list of ``a-b``, ``a`` or ``a-b/k`` (every k-th
  site; ``\\k`` is accepted too), 1-based and inclusive, as the reference's
  dataset builder writes them (``DNA, tree_0=1-500``,
  experiments/src/dataset_builder.py:90);
* model string: ``DNA``/``GTR``/``JC``/``F81`` (4 states) or ``PROT``/``AA``
  (20 states, equal exchangeabilities) with optional ``+G`` / ``+Gn``
  (n discrete Gamma categories; ``+G`` means 4).  Empirical protein
  matrices are not provided.
* pattern compression: identical site columns (as state masks) merged into
  one pattern whose weight is the number of columns, in order of first
  appearance.  The exact pattern order of raxml-ng is not pinned; the
  likelihood does not depend on it.

Parsing is host work (no GPU involved); the results feed
``likelihood.AnnotatedNetwork`` through ``simulate.Partition``.
"""
from __future__ import annotations

import dataclasses
import os
import re
from typing import Dict, List, Optional, Sequence, Tuple, Union

import numpy as np

from .simulate import Model, Partition, aa_masks, dna_masks


# ------------------------------------------------------------------ MSA files

def _text(src: Union[str, os.PathLike]) -> str:
    if isinstance(src, (str, os.PathLike)) and os.path.exists(src):
        with open(src) as fh:
            return fh.read()
    return str(src)


def read_fasta(src) -> Dict[str, str]:
    """Label -> sequence, in file order.  `src` is a path or the text."""
    aln: Dict[str, str] = {}
    label: Optional[str] = None
    for raw in _text(src).splitlines():
        line = raw.strip()
        if not line or line.startswith(";"):
            continue
        if line.startswith(">"):
            label = line[1:].strip().split()[0] if line[1:].strip() else ""
            if not label:
                raise ValueError("FASTA record without a label")
            if label in aln:
                raise ValueError(f"duplicate sequence label {label!r}")
            aln[label] = ""
        else:
            if label is None:
                raise ValueError("sequence data before the first FASTA header")
            aln[label] += "".join(line.split())
    _check_lengths(aln)
    return aln


def read_phylip(src) -> Dict[str, str]:
    """Relaxed sequential PHYLIP: a ``n_taxa n_sites`` header, then one
    ``label sequence`` record per taxon (the sequence may continue on
    following lines until n_sites characters are read)."""
    lines = [l for l in _text(src).splitlines() if l.strip()]
    if not lines:
        raise ValueError("empty PHYLIP file")
    head = lines[0].split()
    if len(head) < 2:
        raise ValueError("PHYLIP header must give the number of taxa and sites")
    n_taxa, n_sites = int(head[0]), int(head[1])
    aln: Dict[str, str] = {}
    i = 1
    for _ in range(n_taxa):
        if i >= len(lines):
            raise ValueError("PHYLIP file ends before all taxa are read")
        parts = lines[i].split()
        label, seq = parts[0], "".join(parts[1:])
        i += 1
        while len(seq) < n_sites and i < len(lines):
            seq += "".join(lines[i].split())
            i += 1
        if label in aln:
            raise ValueError(f"duplicate sequence label {label!r}")
        aln[label] = seq
    _check_lengths(aln)
    if any(len(s) != n_sites for s in aln.values()):
        raise ValueError("PHYLIP sequences do not match the header's site count")
    return aln


def read_alignment(path) -> Dict[str, str]:
    """FASTA or relaxed PHYLIP, by the first non-blank character."""
    text = _text(path)
    first = next((l.strip() for l in text.splitlines() if l.strip()), "")
    return read_fasta(text) if first.startswith(">") else read_phylip(text)


def _check_lengths(aln: Dict[str, str]) -> None:
    if not aln:
        raise ValueError("alignment holds no sequences")
    lens = {len(s) for s in aln.values()}
    if len(lens) != 1:
        raise ValueError("alignment sequences differ in length")
    if 0 in lens:
        raise ValueError("alignment sequences are empty")


# ------------------------------------------------------------ model strings

@dataclasses.dataclass
class ModelSpec:
    """What a model string fixes: the state count, the substitution-rate
    family and the number of Gamma rate categories."""

    name: str
    states: int
    ncat: int
    equal_rates: bool  # exchangeabilities fixed to 1 (JC / F81 / PROT)
    equal_freqs: bool  # frequencies fixed to uniform (JC)

    def initial_model(self, freqs: Optional[np.ndarray] = None, alpha: float = 1.0) -> Model:
        """A starting model: exchangeabilities 1, the given (empirical)
        frequencies unless the model fixes them, Gamma shape `alpha`."""
        S = self.states
        fr = np.full(S, 1.0 / S) if (freqs is None or self.equal_freqs) else np.asarray(freqs, float)
        return Model(S, np.ones(S * (S - 1) // 2), fr, float(alpha), self.ncat)


_DNA_NAMES = {"DNA": (False, False), "GTR": (False, False), "JC": (True, True), "JC69": (True, True),
              "F81": (True, False)}
_AA_NAMES = {"PROT": (True, False), "AA": (True, False), "PROTGTR": (False, False)}


def parse_model(s: str) -> ModelSpec:
    toks = [t.strip() for t in s.strip().split("+") if t.strip()]
    if not toks:
        raise ValueError("empty model string")
    base = toks[0].upper()
    if base in _DNA_NAMES:
        states, (er, ef) = 4, _DNA_NAMES[base]
    elif base in _AA_NAMES:
        states, (er, ef) = 20, _AA_NAMES[base]
    else:
        raise ValueError(f"unsupported model {toks[0]!r} (DNA, GTR, JC, F81, PROT, AA, PROTGTR)")
    ncat = 1
    for t in toks[1:]:
        u = t.upper()
        m = re.fullmatch(r"G(\d*)", u)
        if m:
            ncat = int(m.group(1)) if m.group(1) else 4
            if ncat not in (1, 2, 4, 8):
                raise ValueError("Gamma categories must be 1, 2, 4 or 8")
        elif u in ("FO", "FC", "FE"):
            if u == "FE":
                ef = True
        else:
            raise ValueError(f"unsupported model modifier +{t}")
    return ModelSpec(s.strip(), states, ncat, er, ef)


# ---------------------------------------------------------- partition files

@dataclasses.dataclass
class PartitionSpec:
    model: ModelSpec
    name: str
    columns: np.ndarray  # 0-based alignment columns, in file order


def _parse_ranges(text: str, n_sites: Optional[int]) -> np.ndarray:
    cols: List[int] = []
    for tok in text.split(","):
        tok = tok.strip()
        if not tok:
            continue
        m = re.fullmatch(r"(\d+)(?:\s*-\s*(\d+))?(?:\s*[/\\]\s*(\d+))?", tok)
        if not m:
            raise ValueError(f"bad site range {tok!r}")
        a = int(m.group(1))
        b = int(m.group(2)) if m.group(2) else a
        k = int(m.group(3)) if m.group(3) else 1
        if a < 1 or b < a or k < 1:
            raise ValueError(f"bad site range {tok!r}")
        if n_sites is not None and b > n_sites:
            raise ValueError(f"site range {tok!r} exceeds the alignment's {n_sites} sites")
        cols.extend(range(a - 1, b, k))
    if not cols:
        raise ValueError("partition without sites")
    return np.asarray(cols, np.int64)


def read_partition_file(src, n_sites: Optional[int] = None) -> List[PartitionSpec]:
    """RAxML-style partition file (path or text).  Every column may belong
    to at most one partition; with `n_sites` given every column must be
    covered."""
    out: List[PartitionSpec] = []
    for ln, raw in enumerate(_text(src).splitlines(), 1):
        line = raw.split("#", 1)[0].strip()
        if not line:
            continue
        if "," not in line or "=" not in line:
            raise ValueError(f"line {ln}: expected 'MODEL, NAME = RANGES'")
        model_s, rest = line.split(",", 1)
        name, ranges = rest.split("=", 1)
        name = name.strip()
        if not name:
            raise ValueError(f"line {ln}: partition without a name")
        out.append(PartitionSpec(parse_model(model_s), name, _parse_ranges(ranges, n_sites)))
    if not out:
        raise ValueError("partition file defines no partition")
    allc = np.concatenate([p.columns for p in out])
    if len(np.unique(allc)) != len(allc):
        raise ValueError("a site belongs to more than one partition")
    if n_sites is not None and len(allc) != n_sites:
        raise ValueError(f"partitions cover {len(allc)} of {n_sites} sites")
    names = [p.name for p in out]
    if len(set(names)) != len(names):
        raise ValueError("duplicate partition name")
    return out


# ------------------------------------------------------ pattern compression

def compress_patterns(masks: np.ndarray, order: str = "first") -> Tuple[np.ndarray, np.ndarray]:
    """Merges identical columns of ``masks`` ([n_tips, n_sites] state masks)
    into patterns with integer weights.  order = "first": patterns in order
    of first appearance; "sorted": lexicographic column order."""
    masks = np.asarray(masks)
    if masks.ndim != 2 or masks.shape[1] == 0:
        raise ValueError("masks must be [n_tips, n_sites] with at least one site")
    cols, first, inv, counts = np.unique(masks, axis=1, return_index=True, return_inverse=True,
                                         return_counts=True)
    if order == "sorted":
        return np.ascontiguousarray(cols.astype(np.uint32)), counts.astype(np.uint32)
    if order != "first":
        raise ValueError("order must be 'first' or 'sorted'")
    perm = np.argsort(first, kind="stable")
    return np.ascontiguousarray(cols[:, perm].astype(np.uint32)), counts[perm].astype(np.uint32)


def expand_weights(weights: np.ndarray) -> int:
    """Number of alignment columns the patterns stand for."""
    return int(np.asarray(weights, np.int64).sum())


# ----------------------------------------------------------- engine inputs

def _masks(seq: str, states: int) -> np.ndarray:
    return dna_masks(seq) if states == 4 else aa_masks(seq)


def empirical_freqs(masks: np.ndarray, weights: np.ndarray, states: int) -> np.ndarray:
    """Base frequencies counted over unambiguous characters (each pattern
    counted with its weight); uniform where a state is never seen alone."""
    m = np.asarray(masks, np.uint64)
    w = np.asarray(weights, float)
    counts = np.zeros(states)
    for i in range(states):
        counts[i] = float(((m == (1 << i)) * w[None, :]).sum())
    counts += 1.0  # pseudo-count: no zero frequency
    return counts / counts.sum()


def load_partitions(msa, tip_labels: Sequence[str], partition_file=None, model: str = "GTR+G4",
                    compress: bool = True, alpha: float = 1.0) -> List[Partition]:
    """Engine partitions for a network with tips `tip_labels` (the
    network's tip order, NetworkIO's index layout): reads the MSA (path,
    text, or a label -> sequence dict), splits it by the partition file
    (or one partition with `model`), compresses patterns and starts each
    partition's model from empirical frequencies."""
    aln = msa if isinstance(msa, dict) else read_alignment(msa)
    missing = [l for l in tip_labels if l not in aln]
    if missing:
        raise ValueError(f"taxa missing from the alignment: {missing}")
    n_sites = len(next(iter(aln.values())))
    specs = (read_partition_file(partition_file, n_sites) if partition_file is not None
             else [PartitionSpec(parse_model(model), "p0", np.arange(n_sites))])
    out = []
    for sp in specs:
        S = sp.model.states
        rows = [_masks("".join(aln[l][c] for c in sp.columns), S) for l in tip_labels]
        masks = np.stack(rows).astype(np.uint32)
        if compress:
            masks, weights = compress_patterns(masks)
        else:
            weights = np.ones(masks.shape[1], np.uint32)
        fr = empirical_freqs(masks, weights, S)
        out.append(Partition(sp.model.initial_model(fr, alpha), masks, weights))
    return out
