"""ctypes binding of the CPU FP64 oracle (``nrx_oracle.c``).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker.  The product package never
imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from typing import Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libnrx_oracle.so")

AVERAGE = 0
BEST = 1
MIN_TREE_LOGPROB = float(np.log(1e-6))  # NetraxOptions.hpp:108


class _Net(C.Structure):
    _fields_ = [("n_nodes", C.c_int32), ("n_tips", C.c_int32), ("root", C.c_int32),
                ("n_edges", C.c_int32), ("n_retic", C.c_int32),
                ("edge_parent", C.POINTER(C.c_int32)), ("edge_child", C.POINTER(C.c_int32)),
                ("retic_first_edge", C.POINTER(C.c_int32)),
                ("retic_second_edge", C.POINTER(C.c_int32))]


class _Part(C.Structure):
    _fields_ = [("n_sites", C.c_int32), ("states", C.c_int32), ("ncat", C.c_int32),
                ("tip_masks", C.POINTER(C.c_uint32)), ("weights", C.POINTER(C.c_uint32)),
                ("subst_rates", C.POINTER(C.c_double)), ("freqs", C.POINTER(C.c_double)),
                ("alpha", C.c_double)]


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        _lib.orc_last_error.restype = C.c_char_p
    return _lib


def _ptr(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


class _Args:
    """Keeps numpy buffers alive while ctypes structs point at them."""

    def __init__(self, net, parts):
        self.keep = []
        ep = np.ascontiguousarray(net.edge_parent, np.int32)
        ec = np.ascontiguousarray(net.edge_child, np.int32)
        rf = np.ascontiguousarray(net.retic_first_edge, np.int32)
        rs = np.ascontiguousarray(net.retic_second_edge, np.int32)
        if len(rf) == 0:
            rf = np.zeros(1, np.int32)
            rs = np.zeros(1, np.int32)
        self.keep += [ep, ec, rf, rs]
        self.net = _Net(net.n_nodes, net.n_tips, net.root, net.n_edges, net.n_retic,
                        _ptr(ep, C.c_int32), _ptr(ec, C.c_int32), _ptr(rf, C.c_int32),
                        _ptr(rs, C.c_int32))
        arr = (_Part * len(parts))()
        for i, p in enumerate(parts):
            m = np.ascontiguousarray(p.masks, np.uint32)
            w = np.ascontiguousarray(p.weights, np.uint32)
            sr = np.ascontiguousarray(p.model.subst_rates, np.float64)
            fr = np.ascontiguousarray(p.model.freqs, np.float64)
            self.keep += [m, w, sr, fr]
            arr[i] = _Part(m.shape[1], p.model.states, p.model.ncat, _ptr(m, C.c_uint32),
                           _ptr(w, C.c_uint32), _ptr(sr, C.c_double), _ptr(fr, C.c_double),
                           float(p.model.alpha))
        self.parts = arr
        self.n_parts = len(parts)


def _check(rc):
    if rc != 0:
        raise RuntimeError("oracle: " + lib().orc_last_error().decode())


def set_threads(n: int) -> None:
    lib().orc_set_threads(C.c_int(n))


def gamma_rates(alpha: float, ncat: int) -> np.ndarray:
    out = np.zeros(ncat)
    lib().orc_gamma_rates(C.c_double(alpha), C.c_int(ncat), _ptr(out, C.c_double))
    return out


def pmatrix(model, rate: float, t: float):
    S = model.states
    P, dP, d2P = np.zeros((S, S)), np.zeros((S, S)), np.zeros((S, S))
    sr = np.ascontiguousarray(model.subst_rates, np.float64)
    fr = np.ascontiguousarray(model.freqs, np.float64)
    lib().orc_pmatrix(C.c_int(S), _ptr(sr, C.c_double), _ptr(fr, C.c_double), C.c_double(rate),
                      C.c_double(t), _ptr(P, C.c_double), _ptr(dP, C.c_double), _ptr(d2P, C.c_double))
    return P, dP, d2P


def _common(net, brlen, probs):
    brlen = np.ascontiguousarray(net.brlen if brlen is None else brlen, np.float64)
    probs = np.ascontiguousarray(net.retic_prob if probs is None else probs, np.float64)
    if len(probs) == 0:
        probs = np.zeros(1)
    return brlen, probs


def num_classes(net, probs=None) -> int:
    a = _Args(net, [])
    _, probs = _common(net, None, probs)
    return lib().orc_num_classes(C.byref(a.net), _ptr(probs, C.c_double))


def num_variants(net, probs=None) -> int:
    a = _Args(net, [])
    _, probs = _common(net, None, probs)
    return lib().orc_num_variants(C.byref(a.net), _ptr(probs, C.c_double))


def loglikelihood(net, parts: Sequence, variant: int = AVERAGE, brlen=None, probs=None,
                  min_tree_logprob: float = MIN_TREE_LOGPROB, tree_logls: bool = False):
    """Returns (total, per_partition[, tree_logl [P][classes], class_logprob])."""
    a = _Args(net, parts)
    brlen, probs = _common(net, brlen, probs)
    total = C.c_double()
    per = np.zeros(a.n_parts)
    nc = num_classes(net, probs)
    tl = np.zeros((a.n_parts, nc))
    lp = np.zeros(nc)
    _check(lib().orc_loglikelihood(C.byref(a.net), _ptr(brlen, C.c_double), _ptr(probs, C.c_double),
                                   a.parts, C.c_int(a.n_parts), C.c_int(variant),
                                   C.c_double(min_tree_logprob), C.byref(total), _ptr(per, C.c_double),
                                   _ptr(tl, C.c_double) if tree_logls else None,
                                   _ptr(lp, C.c_double)))
    if tree_logls:
        return total.value, per, tl, lp
    return total.value, per


def gradient(net, parts: Sequence, variant: int = AVERAGE, brlen=None, probs=None,
             min_tree_logprob: float = MIN_TREE_LOGPROB, per_partition: bool = False):
    """Returns (lnL, dlnL/dbrlen [E], dlnL/dprob [R][, per-partition dbrlen [P][E]])."""
    a = _Args(net, parts)
    brlen, probs = _common(net, brlen, probs)
    total = C.c_double()
    db = np.zeros(net.n_edges)
    dp = np.zeros(max(net.n_retic, 1))
    pdb = np.zeros((a.n_parts, net.n_edges))
    _check(lib().orc_gradient(C.byref(a.net), _ptr(brlen, C.c_double), _ptr(probs, C.c_double),
                              a.parts, C.c_int(a.n_parts), C.c_int(variant),
                              C.c_double(min_tree_logprob), C.byref(total), _ptr(db, C.c_double),
                              _ptr(dp, C.c_double), _ptr(pdb, C.c_double)))
    if per_partition:
        return total.value, db, dp[:net.n_retic], pdb
    return total.value, db, dp[:net.n_retic]


def brlen_derivatives(net, parts: Sequence, edge: int, variant: int = AVERAGE, mode: int = 0,
                      brlen=None, probs=None, min_tree_logprob: float = MIN_TREE_LOGPROB):
    """First/second lnL derivative for one branch.  mode 0 mirrors the
    reference's tree normalisation, mode 1 is the true derivative.
    Returns (d1, d2, per_partition_d1, per_partition_d2)."""
    a = _Args(net, parts)
    brlen, probs = _common(net, brlen, probs)
    d1, d2 = C.c_double(), C.c_double()
    p1, p2 = np.zeros(a.n_parts), np.zeros(a.n_parts)
    _check(lib().orc_brlen_derivatives(C.byref(a.net), _ptr(brlen, C.c_double), _ptr(probs, C.c_double),
                                       a.parts, C.c_int(a.n_parts), C.c_int(variant),
                                       C.c_double(min_tree_logprob), C.c_int(edge), C.c_int(mode),
                                       C.byref(d1), C.byref(d2), _ptr(p1, C.c_double),
                                       _ptr(p2, C.c_double)))
    return d1.value, d2.value, p1, p2
