"""Network model, Newick reader and writer, and generators (CPU)."""
import dataclasses
import re

import numpy as np
import pytest

from fixtures import SAMPLE_PAIRS, load_sample
from netrax_amd import _lib
from netrax_amd.network import active_edges, is_ancestor, parse_newick, random_network
from netrax_amd.simulate import compress_patterns, make_workload


def test_tiny_index_layout_follows_reference_conversion():
    # ((A:2,(B:1)X#H1:0::0.3)Q:2,(C:2,X#H1:0::0.7)R:2);  NetworkIO.cpp:114-185:
    # tips A,B,C = 0,1,2; inner tree nodes in collectNodes order (R, Q), root last;
    # the reticulation X follows; its parent edges are 5 (first, from Q) and 6.
    net, _ = load_sample("tiny", with_lengths=False)
    assert net.tip_labels == ["A", "B", "C"]
    assert (net.n_tips, net.n_inner_tree, net.n_retic, net.root) == (3, 3, 1, 5)
    assert list(net.edge_child) == [0, 1, 2, 3, 4, 6, 6]
    assert net.retic_first_edge[0] == 5 and net.retic_second_edge[0] == 6
    assert net.retic_prob[0] == pytest.approx(0.3)
    # Q (first parent) holds A; lengths of the reticulation edges clamp to 1e-6
    q = net.edge_parent[5]
    assert 0 in [int(net.edge_child[e]) for e in range(net.n_edges) if net.edge_parent[e] == q]
    assert net.brlen[5] == pytest.approx(1e-6) and net.brlen[6] == pytest.approx(1e-6)


@pytest.mark.parametrize("name", sorted(SAMPLE_PAIRS))
def test_samples_parse(name):
    net, parts = load_sample(name)
    net.validate()
    assert parts[0].masks.shape[0] == net.n_tips


def test_unary_root_and_trifurcation():
    net, _ = load_sample("celine", with_lengths=False)
    assert (net.edge_parent == net.root).sum() == 1  # extra parentheses: unary root
    tri, _ = load_sample("tree", with_lengths=False)
    assert (tri.edge_parent == tri.root).sum() == 2  # split into two children
    assert tri.n_inner_tree == 3


def test_probability_defaults_and_validation():
    net = parse_newick("((A,(B)#H1),(#H1,C));")
    assert net.retic_prob[0] == pytest.approx(0.5)
    with pytest.raises(ValueError):
        parse_newick("((A,(B)#H1:1::0.9),(#H1:1::0.9,C));")
    with pytest.raises(ValueError):
        parse_newick("((A,B),C)")


@pytest.mark.parametrize("n,r", [(8, 1), (30, 3), (50, 4), (60, 5)])
def test_generator_shapes(n, r):
    net = random_network(n, r, np.random.default_rng(n * 10 + r))
    net.validate()
    assert net.n_tips == n and net.n_retic == r
    assert net.n_edges == 2 * n - 2 + 3 * r
    assert np.all(net.brlen >= 1e-6)
    assert np.all((net.retic_prob >= 0.2) & (net.retic_prob <= 0.8))
    for k in range(r):
        p1 = net.edge_parent[net.retic_first_edge[k]]
        p2 = net.edge_parent[net.retic_second_edge[k]]
        assert not is_ancestor(net, p1, p2) or not is_ancestor(net, p2, p1)


def test_active_edges_bit_order():
    net, _ = load_sample("two_reticulations")
    act = active_edges(net, 0b01)  # reticulation 0 takes its second parent
    assert not act[net.retic_first_edge[0]] and act[net.retic_second_edge[0]]
    assert act[net.retic_first_edge[1]] and not act[net.retic_second_edge[1]]


def test_workload_is_seeded_and_compressed():
    a = make_workload("a", 12, 2, [500, 700], seed=9)
    b = make_workload("b", 12, 2, [500, 700], seed=9)
    assert np.array_equal(a.partitions[1].masks, b.partitions[1].masks)
    assert [int(p.weights.sum()) for p in a.partitions] == [500, 700]
    cols, w = compress_patterns(a.partitions[0].masks)
    assert cols.shape == a.partitions[0].masks.shape


def test_library_exports_every_header_symbol():
    import os
    hdr = open(os.path.join(os.path.dirname(__file__), "..", "include", "netrax_gpu.h")).read()
    declared = sorted(set(re.findall(r"\b(nrx_[a-z_]+)\s*\(", hdr)))
    assert sorted(_lib.EXPORTED) == declared
    lib = _lib.load()
    for sym in declared:
        assert hasattr(lib, sym), sym


def test_abi_rejects_null_arguments_without_touching_the_device():
    import ctypes as C
    lib = _lib.load()
    h = C.c_void_p()
    rc = lib.nrx_create(None, None, 0, None, C.byref(h))
    assert rc == _lib.NRX_ERR_ARG
    assert b"null" in lib.nrx_last_error()
    opt = _lib.Options()
    lib.nrx_default_options(C.byref(opt))
    assert opt.precision == _lib.NRX_FP64 and opt.brlen_min == pytest.approx(1e-6)


def test_cpp_mirror_reports_engine_errors():
    """The C++ mirror (include/netrax_likelihood.hpp) turns a failing engine
    call into std::runtime_error with the engine's message; without a GPU the
    example program reports the HIP failure and exits 2."""
    import os
    import subprocess
    from netrax_amd import build as nb
    if not os.path.exists(nb.EXAMPLE):
        nb.build()
    env = dict(os.environ, HIP_VISIBLE_DEVICES="-1")
    res = subprocess.run([nb.EXAMPLE], capture_output=True, text=True, env=env, timeout=120)
    assert res.returncode == 2
    assert "error:" in res.stderr


@pytest.mark.parametrize("name", sorted(SAMPLE_PAIRS))
def test_newick_writer_round_trip_keeps_the_likelihood(name):
    """to_newick (extended Newick, `:len::prob` on reticulation edges) read
    back by the parser gives the same network: same taxa, edge and
    reticulation counts, and the same oracle lnL on the sample alignment
    (the node numbering may differ)."""
    from oracle import oracle as O

    net, parts = load_sample(name)
    back = parse_newick(net.to_newick())
    assert sorted(back.tip_labels) == sorted(net.tip_labels)
    assert (back.n_edges, back.n_retic) == (net.n_edges, net.n_retic)
    order = [net.tip_labels.index(l) for l in back.tip_labels]
    parts_b = [dataclasses.replace(p, masks=np.ascontiguousarray(p.masks[order])) for p in parts]
    for variant in (O.AVERAGE, O.BEST):
        l1 = O.loglikelihood(net, parts, variant)[0]
        l2 = O.loglikelihood(back, parts_b, variant)[0]
        assert abs(l1 - l2) <= 1e-10 * abs(l1), (name, variant, l1, l2)
