"""Allocation-free branch derivatives (nrx_brlen_derivatives_noalloc) against
nrx_brlen_derivatives, both called through the C ABI: two engines in a row,
the second with a much smaller input, both variants and derivative modes,
FP64 and FP32 CLVs; plus the entry point's state and size refusals."""
import ctypes as C

import numpy as np
import pytest

from fixtures import fixed_model, load_sample
from netrax_amd import AnnotatedNetwork, NetraxOptions, Precision, _lib
from netrax_amd.network import random_network
from netrax_amd.simulate import Partition, random_dna_model, workload_30taxa_10k

pytestmark = pytest.mark.gpu


def _call(fn, ann, variant, edge, mode, brlen):
    P = len(ann.partitions)
    d1, d2 = C.c_double(), C.c_double()
    p1, p2 = np.zeros(P), np.zeros(P)
    pd = C.POINTER(C.c_double)
    rc = fn(ann.handle, variant, edge, mode, brlen, C.byref(d1), C.byref(d2), p1.ctypes.data_as(pd),
            p2.ctypes.data_as(pd))
    return rc, d1.value, d2.value, p1, p2


def _close(a, b, rtol=1e-12):
    return abs(a - b) <= rtol * max(1.0, abs(b))


@pytest.mark.parametrize("prec", [Precision.FP64, Precision.FP32])
def test_noalloc_matches_brlen_derivatives_on_two_engines(prec):
    lib = _lib.load()
    big = workload_30taxa_10k()
    small_net, small_parts = load_sample("two_reticulations", fixed_model())
    rng = np.random.default_rng(3)
    for net, parts in ((big.network, big.partitions), (small_net, small_parts)):
        print("engine", net.n_tips, "taxa", sum(p.n_sites for p in parts), "patterns", prec.name, flush=True)
        ann = AnnotatedNetwork(net, parts, NetraxOptions(precision=prec))
        edges = rng.choice(net.n_edges, min(6, net.n_edges), replace=False)
        for variant in (0, 1):
            ann.loglikelihood(variant=variant)
            for e in edges:
                for mode in (0, 1):
                    b = float(net.brlen[e]) * 1.3
                    rn = _call(lib.nrx_brlen_derivatives_noalloc, ann, variant, int(e), mode, b)
                    rr = _call(lib.nrx_brlen_derivatives, ann, variant, int(e), mode, b)
                    assert rn[0] == _lib.NRX_OK, lib.nrx_last_error()
                    assert rr[0] == _lib.NRX_OK, lib.nrx_last_error()
                    assert _close(rn[1], rr[1]) and _close(rn[2], rr[2]), (variant, e, mode, rn[1:3], rr[1:3])
                    assert np.allclose(rn[3], rr[3], rtol=1e-12, atol=1e-12)
                    assert np.allclose(rn[4], rr[4], rtol=1e-12, atol=1e-12)
        # parameters changed since the last evaluation: refused, then fine again
        b = net.brlen.copy()
        b[edges[0]] *= 1.1
        ann.set_branch_lengths(b)
        rc = _call(lib.nrx_brlen_derivatives_noalloc, ann, 0, int(edges[0]), 0, float(b[edges[0]]))[0]
        assert rc == _lib.NRX_ERR_STATE
        assert b"nrx_loglikelihood" in lib.nrx_last_error()
        ann.loglikelihood(variant=0)
        rn = _call(lib.nrx_brlen_derivatives_noalloc, ann, 0, int(edges[0]), 0, float(b[edges[0]]))
        rr = _call(lib.nrx_brlen_derivatives, ann, 0, int(edges[0]), 0, float(b[edges[0]]))
        assert rn[0] == rr[0] == _lib.NRX_OK
        assert _close(rn[1], rr[1]) and _close(rn[2], rr[2])
        ann.close()


def test_noalloc_refuses_derivative_scratch_beyond_its_limit():
    """An edge in 32 displayed trees at 600k sites needs more than the fixed
    16 MiB of scratch: refused before any device work."""
    lib = _lib.load()
    rng = np.random.default_rng(1)
    net = random_network(12, 5, rng)  # 32 distinct displayed trees
    n = 600_000
    masks = (np.uint32(1) << rng.integers(0, 4, size=(net.n_tips, n)).astype(np.uint32)).astype(np.uint32)
    part = Partition(random_dna_model(rng), masks, np.ones(n, np.uint32))
    ann = AnnotatedNetwork(net, [part], NetraxOptions(precision=Precision.FP32))
    tip_edge = int(np.nonzero(np.asarray(net.edge_child) == 0)[0][0])
    rc = _call(lib.nrx_brlen_derivatives_noalloc, ann, 0, tip_edge, 0, 0.1)[0]
    assert rc == _lib.NRX_ERR_ARG
    assert b"no-alloc limit" in lib.nrx_last_error()
    ann.close()
