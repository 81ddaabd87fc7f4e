"""Host-side mirror of NetRAX's likelihood interface over the C ABI.

Names, argument meaning and error behaviour follow the reference:

* ``computeLoglikelihood(ann_network, incremental=1, update_pmatrices=1)``
  (src/likelihood/LikelihoodComputation.hpp:17-18) returns the network lnL,
  fills ``ann_network.partition_loglh`` and raises ``RuntimeError`` with
  the reference's message on a -infinity result
  (ImprovedLoglikelihood.cpp:637-640).
* ``computePartitionSumtables`` + ``computeLoglikelihoodDerivatives``
  (src/likelihood/LikelihoodDerivatives.hpp:83-88) return a
  ``LoglDerivatives`` with total and per-partition first/second derivatives
  of lnL in one branch length.
* ``LikelihoodVariant`` (src/likelihood/LikelihoodVariant.hpp:3-7).

Every call runs on the GPU engine; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import enum
import math
from typing import List, Optional, Sequence

import numpy as np

from . import _lib
from .network import BRLEN_MAX, BRLEN_MIN, Network
from .simulate import Partition


class LikelihoodVariant(enum.IntEnum):
    AVERAGE_DISPLAYED_TREES = 0
    BEST_DISPLAYED_TREE = 1
    SARAH_PSEUDO = 2  # not provided by the engine


class Precision(enum.IntEnum):
    FP64 = _lib.NRX_FP64
    FP32 = _lib.NRX_FP32


@dataclasses.dataclass
class LoglDerivatives:
    """src/likelihood/LikelihoodDerivatives.hpp:76-81"""

    logl_prime: float
    logl_prime_prime: float
    partition_logl_prime: List[float]
    partition_logl_prime_prime: List[float]


@dataclasses.dataclass
class NetraxOptions:
    """The options of NetraxOptions.hpp this path reads."""

    likelihood_variant: LikelihoodVariant = LikelihoodVariant.AVERAGE_DISPLAYED_TREES
    min_interesting_tree_logprob: float = math.log(1e-6)  # NetraxOptions.hpp:108
    brlen_min: float = BRLEN_MIN
    brlen_max: float = BRLEN_MAX
    brprob_min: float = 1e-6        # NetraxOptions.hpp:102
    brprob_max: float = 1.0 - 1e-6  # NetraxOptions.hpp:103
    prob_tolerance: float = 1e-7    # width of the final bracket of a probability
    precision: Precision = Precision.FP64
    device: int = 0
    use_graph: bool = False
    timing: bool = False


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class AnnotatedNetwork:
    """A network with its partitions and the engine that scores it.

    Plays the role of the reference's AnnotatedNetwork
    (src/graph/AnnotatedNetwork.hpp:42-89) for the likelihood path: it owns
    the device state (CLVs, scalers, P matrices) through the engine handle.
    """

    def __init__(self, network: Network, partitions: Sequence[Partition],
                 options: Optional[NetraxOptions] = None):
        self.lib = _lib.load()
        self.network = network
        self.partitions = list(partitions)
        self.options = options or NetraxOptions()
        self.partition_loglh = [0.0] * len(self.partitions)
        self.cached_logl = None
        self._keep = []
        ep = np.ascontiguousarray(network.edge_parent, np.int32)
        ec = np.ascontiguousarray(network.edge_child, np.int32)
        rf = np.ascontiguousarray(network.retic_first_edge, np.int32)
        rs = np.ascontiguousarray(network.retic_second_edge, np.int32)
        if len(rf) == 0:
            rf, rs = np.zeros(1, np.int32), np.zeros(1, np.int32)
        self._keep += [ep, ec, rf, rs]
        ip = C.POINTER(C.c_int32)
        nd = _lib.NetworkDesc(network.n_nodes, network.n_tips, network.root, network.n_edges,
                              network.n_retic, ep.ctypes.data_as(ip), ec.ctypes.data_as(ip),
                              rf.ctypes.data_as(ip), rs.ctypes.data_as(ip))
        pds = (_lib.PartitionDesc * len(self.partitions))()
        for i, p in enumerate(self.partitions):
            m = np.ascontiguousarray(p.masks, np.uint32)
            w = np.ascontiguousarray(p.weights, np.uint32)
            if m.shape[0] != network.n_tips:
                raise ValueError("alignment rows must match the network tips")
            self._keep += [m, w]
            up = C.POINTER(C.c_uint32)
            pds[i] = _lib.PartitionDesc(m.shape[1], p.model.states, p.model.ncat,
                                        m.ctypes.data_as(up), w.ctypes.data_as(up))
        opt = _lib.Options()
        self.lib.nrx_default_options(C.byref(opt))
        o = self.options
        opt.precision = int(o.precision)
        opt.device = o.device
        opt.min_tree_logprob = o.min_interesting_tree_logprob
        opt.brlen_min = o.brlen_min
        opt.brlen_max = o.brlen_max
        opt.use_graph = int(o.use_graph)
        opt.timing = int(o.timing)
        h = C.c_void_p()
        _lib.check(self.lib.nrx_create(C.byref(nd), pds, len(self.partitions), C.byref(opt),
                                       C.byref(h)))
        self.handle = h
        for i, p in enumerate(self.partitions):
            self.set_model(i, p.model)
        self.set_branch_lengths(network.brlen)
        if network.n_retic:
            self.set_reticulation_probs(network.retic_prob)

    # -- parameters --------------------------------------------------------
    def set_model(self, i: int, model) -> None:
        sr = np.ascontiguousarray(model.subst_rates, np.float64)
        fr = np.ascontiguousarray(model.freqs, np.float64)
        _lib.check(self.lib.nrx_set_model(self.handle, i, _dptr(sr), _dptr(fr), float(model.alpha)))
        self.cached_logl = None

    def set_branch_lengths(self, brlen) -> None:
        b = np.ascontiguousarray(brlen, np.float64)
        if b.shape != (self.network.n_edges,):
            raise ValueError("one branch length per edge expected")
        _lib.check(self.lib.nrx_set_branch_lengths(self.handle, _dptr(b)))
        self.network.brlen = np.clip(b, self.options.brlen_min, self.options.brlen_max)

    def set_branch_length(self, edge: int, length: float) -> None:
        b = self.network.brlen.copy()
        b[edge] = length
        self.set_branch_lengths(b)

    def set_reticulation_probs(self, probs) -> None:
        p = np.ascontiguousarray(probs, np.float64)
        if p.shape != (self.network.n_retic,):
            raise ValueError("one probability per reticulation expected")
        _lib.check(self.lib.nrx_set_reticulation_probs(self.handle, _dptr(p)))
        self.network.retic_prob = p.copy()

    # -- evaluation --------------------------------------------------------
    def loglikelihood(self, incremental: int = 1, update_pmatrices: int = 1,
                      variant: Optional[int] = None) -> float:
        v = int(self.options.likelihood_variant if variant is None else variant)
        if v == LikelihoodVariant.SARAH_PSEUDO:
            raise NotImplementedError("SARAH_PSEUDO is not provided by the engine")
        tot = C.c_double()
        per = np.zeros(len(self.partitions))
        rc = self.lib.nrx_loglikelihood(self.handle, v, int(incremental), int(update_pmatrices),
                                        C.byref(tot), _dptr(per))
        if rc == _lib.NRX_ERR_NEGINF:
            raise RuntimeError("Invalid network likelihood: negative infinity \n")
        _lib.check(rc)
        self.partition_loglh = list(per)
        self.cached_logl = tot.value
        return tot.value

    def gradient(self, variant: Optional[int] = None, per_partition: bool = False):
        """(lnL, dlnL/dbrlen [E], dlnL/dprob [R][, per-partition dbrlen])."""
        v = int(self.options.likelihood_variant if variant is None else variant)
        tot = C.c_double()
        db = np.zeros(self.network.n_edges)
        dp = np.zeros(max(self.network.n_retic, 1))
        pdb = np.zeros((len(self.partitions), self.network.n_edges))
        rc = self.lib.nrx_gradient(self.handle, v, C.byref(tot), _dptr(db), _dptr(dp), _dptr(pdb))
        if rc == _lib.NRX_ERR_NEGINF:
            raise RuntimeError("Invalid network likelihood: negative infinity \n")
        _lib.check(rc)
        self.cached_logl = tot.value
        if per_partition:
            return tot.value, db, dp[:self.network.n_retic], pdb
        return tot.value, db, dp[:self.network.n_retic]

    def brlen_derivatives(self, edge: int, brlen: Optional[float] = None, mode: int = 0,
                          variant: Optional[int] = None) -> LoglDerivatives:
        v = int(self.options.likelihood_variant if variant is None else variant)
        b = float(self.network.brlen[edge] if brlen is None else brlen)
        d1, d2 = C.c_double(), C.c_double()
        p1 = np.zeros(len(self.partitions))
        p2 = np.zeros(len(self.partitions))
        _lib.check(self.lib.nrx_brlen_derivatives(self.handle, v, int(edge), int(mode), b,
                                                  C.byref(d1), C.byref(d2), _dptr(p1), _dptr(p2)))
        return LoglDerivatives(d1.value, d2.value, list(p1), list(p2))

    def model_gradient(self, variant: Optional[int] = None):
        """(lnL, [per partition: dict(subst=, freqs=, alpha=, rates=)]) --
        dlnL with respect to each partition's exchangeabilities,
        frequencies, Gamma shape and category rates (nrx_model_gradient)."""
        v = int(self.options.likelihood_variant if variant is None else variant)
        P = len(self.partitions)
        S = self.partitions[0].model.states
        Cn = self.partitions[0].model.ncat
        tot = C.c_double()
        ds = np.zeros((P, S * (S - 1) // 2))
        df = np.zeros((P, S))
        da = np.zeros(P)
        dr = np.zeros((P, Cn))
        rc = self.lib.nrx_model_gradient(self.handle, v, C.byref(tot), _dptr(ds), _dptr(df), _dptr(da), _dptr(dr))
        if rc == _lib.NRX_ERR_NEGINF:
            raise RuntimeError("Invalid network likelihood: negative infinity \n")
        _lib.check(rc)
        self.cached_logl = tot.value
        return tot.value, [dict(subst=ds[p], freqs=df[p], alpha=da[p], rates=dr[p]) for p in range(P)]

    def set_category_rates(self, i: int, rates) -> None:
        r = np.ascontiguousarray(rates, np.float64)
        _lib.check(self.lib.nrx_set_category_rates(self.handle, i, _dptr(r)))
        self.cached_logl = None

    def set_memory_budget(self, nbytes: int) -> None:
        """Bytes for the CLV, adjoint and scaler rows of the FP64 form (0:
        all resident); see nrx_set_memory_budget."""
        _lib.check(self.lib.nrx_set_memory_budget(self.handle, int(nbytes)))
        self.cached_logl = None

    def memory_plan(self) -> dict:
        mp = _lib.MemoryPlan()
        _lib.check(self.lib.nrx_memory_plan(self.handle, C.byref(mp)))
        return dict(segments=mp.segments, rows=mp.rows, boundary_rows=mp.boundary_rows,
                    resident_rows=mp.resident_rows, row_bytes=mp.row_bytes)

    def branch_lengths(self) -> np.ndarray:
        """The engine's (clamped) branch lengths."""
        b = np.zeros(self.network.n_edges)
        _lib.check(self.lib.nrx_get_branch_lengths(self.handle, _dptr(b)))
        return b

    def optimize_branches(self, edges=None, max_iters: int = 32, mode: int = _lib.NRX_DERIV_REFERENCE,
                          variant: Optional[int] = None) -> float:
        """Newton-Raphson sweep over `edges` (all by default) with the
        iteration on the device (nrx_optimize_branches); returns the final
        lnL and updates ``network.brlen``."""
        v = int(self.options.likelihood_variant if variant is None else variant)
        ed = np.ascontiguousarray(range(self.network.n_edges) if edges is None else list(edges), np.int32)
        out = C.c_double()
        rc = self.lib.nrx_optimize_branches(self.handle, v, int(mode), ed.ctypes.data_as(C.POINTER(C.c_int32)),
                                            len(ed), int(max_iters), C.byref(out))
        if rc == _lib.NRX_ERR_NEGINF:
            raise RuntimeError("Invalid network likelihood: negative infinity \n")
        if rc == _lib.NRX_ERR_STATE:
            raise RuntimeError(self.lib.nrx_last_error().decode())
        _lib.check(rc)
        self.network.brlen = self.branch_lengths()
        self.cached_logl = out.value
        return out.value

    def prob_derivatives(self, variant: Optional[int] = None):
        """(lnL, d lnL/dp [R], d2 lnL/dp2 [R]) with the CLVs reused."""
        v = int(self.options.likelihood_variant if variant is None else variant)
        tot = C.c_double()
        R = max(self.network.n_retic, 1)
        d1, d2 = np.zeros(R), np.zeros(R)
        rc = self.lib.nrx_prob_derivatives(self.handle, v, C.byref(tot), _dptr(d1), _dptr(d2))
        if rc == _lib.NRX_ERR_NEGINF:
            raise RuntimeError("Invalid network likelihood: negative infinity \n")
        _lib.check(rc)
        self.cached_logl = tot.value
        return tot.value, d1[:self.network.n_retic], d2[:self.network.n_retic]

    def reticulation_probs(self) -> np.ndarray:
        p = np.zeros(max(self.network.n_retic, 1))
        _lib.check(self.lib.nrx_get_reticulation_probs(self.handle, _dptr(p)))
        return p[:self.network.n_retic]

    def optimize_reticulations(self, max_iters: int = 10, variant: Optional[int] = None) -> float:
        """Reticulation probabilities by Newton on the exact derivative
        (nrx_optimize_reticulations); returns the final lnL and updates
        ``network.retic_prob``."""
        v = int(self.options.likelihood_variant if variant is None else variant)
        o = self.options
        out = C.c_double()
        rc = self.lib.nrx_optimize_reticulations(self.handle, v, int(max_iters), float(o.brprob_min),
                                                 float(o.brprob_max), float(o.prob_tolerance), C.byref(out))
        if rc == _lib.NRX_ERR_NEGINF:
            raise RuntimeError("Invalid network likelihood: negative infinity \n")
        _lib.check(rc)
        if self.network.n_retic:
            self.network.retic_prob = self.reticulation_probs()
        self.cached_logl = out.value
        return out.value

    def tree_logls(self):
        """([P][classes] per-tree lnL of the last pass, [classes] log P(t))."""
        st = self.stats()
        tl = np.zeros((len(self.partitions), st.n_classes))
        lp = np.zeros(st.n_classes)
        _lib.check(self.lib.nrx_tree_logls(self.handle, _dptr(tl), _dptr(lp)))
        return tl, lp

    def stats(self) -> _lib.Stats:
        s = _lib.Stats()
        _lib.check(self.lib.nrx_get_stats(self.handle, C.byref(s)))
        return s

    def last_timing(self) -> _lib.Timing:
        t = _lib.Timing()
        _lib.check(self.lib.nrx_last_timing(self.handle, C.byref(t)))
        return t

    def autotune(self, variant: Optional[int] = None, reps: int = 3):
        """FP32 form: times every kernel shape (ops per group, sites per
        workgroup, lanes per row in the reverse sweep), keeps the fastest;
        returns (best, [all candidates]) as (group_width, tile_sites,
        row_lanes, ms) tuples."""
        v = int(self.options.likelihood_variant if variant is None else variant)
        best = _lib.Tuning()
        res = (_lib.Tuning * 64)()
        n = C.c_int32()
        _lib.check(self.lib.nrx_autotune(self.handle, v, int(reps), C.byref(best), res, 64, C.byref(n)))
        return ((best.group_width, best.tile_sites, best.row_lanes, best.ms),
                [(r.group_width, r.tile_sites, r.row_lanes, r.ms) for r in res[:n.value]])

    def set_tuning(self, group_width: int, tile_sites: int, row_lanes: int = 1) -> None:
        _lib.check(self.lib.nrx_set_tuning(self.handle, int(group_width), int(tile_sites), int(row_lanes)))

    def tuning(self):
        """(group_width, tile_sites, row_lanes) of the FP32 form's kernels."""
        t = _lib.Tuning()
        _lib.check(self.lib.nrx_get_tuning(self.handle, C.byref(t)))
        return t.group_width, t.tile_sites, t.row_lanes

    def close(self) -> None:
        if getattr(self, "handle", None):
            self.lib.nrx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- reference names

def computeLoglikelihood(ann_network: AnnotatedNetwork, incremental: int = 1,
                         update_pmatrices: int = 1) -> float:
    """src/likelihood/LikelihoodComputation.cpp:18-34"""
    return ann_network.loglikelihood(incremental, update_pmatrices)


class SumtableHandle:
    """Stands for the sumtables of one branch: the engine keeps them on the
    device; this records which branch they belong to."""

    def __init__(self, pmatrix_index: int):
        self.pmatrix_index = pmatrix_index


def computePartitionSumtables(ann_network: AnnotatedNetwork, pmatrix_index: int) -> SumtableHandle:
    """src/likelihood/LikelihoodDerivatives.cpp:291-344"""
    if not 0 <= pmatrix_index < ann_network.network.n_edges:
        raise IndexError("pmatrix_index out of range")
    return SumtableHandle(pmatrix_index)


def computeLoglikelihoodDerivatives(ann_network: AnnotatedNetwork, sumtables: SumtableHandle,
                                    pmatrix_index: int, brlen: Optional[float] = None,
                                    mode: int = _lib.NRX_DERIV_REFERENCE) -> LoglDerivatives:
    """src/likelihood/LikelihoodDerivatives.cpp:190-232"""
    if sumtables.pmatrix_index != pmatrix_index:
        raise ValueError("sumtables belong to another branch")
    return ann_network.brlen_derivatives(pmatrix_index, brlen, mode)


class NetworkBatch:
    """Scores several FP32-form networks with one launch per kernel
    (the candidate loop of a search step, src/search/Filtering.cpp:236-270,
    scores them one at a time)."""

    def __init__(self, networks: Sequence[AnnotatedNetwork]):
        self.lib = _lib.load()
        self.networks = list(networks)
        if len({id(a) for a in self.networks}) != len(self.networks):
            raise ValueError("a network appears twice in the batch")
        self._check_open()
        arr = (C.c_void_p * len(self.networks))(*[a.handle.value for a in self.networks])
        h = C.c_void_p()
        _lib.check(self.lib.nrx_batch_create(arr, len(self.networks), C.byref(h)))
        self.handle = h

    def _check_open(self) -> None:
        for i, a in enumerate(self.networks):
            if not getattr(a, "handle", None):
                raise RuntimeError(f"network {i} of the batch is closed")

    def evaluate(self, variant: int = 0, gradient: bool = True):
        """(lnL [n], per-network dlnL/dbrlen list, per-network dlnL/dprob
        list, device ms); the lists are empty when gradient is False."""
        if not getattr(self, "handle", None):
            raise RuntimeError("the batch is closed")
        self._check_open()
        n = len(self.networks)
        lnl = np.zeros(n)
        E = [a.network.n_edges for a in self.networks]
        R = [a.network.n_retic for a in self.networks]
        db = np.zeros(max(1, sum(E)))
        dp = np.zeros(max(1, sum(R)))
        ms = C.c_double()
        rc = self.lib.nrx_batch_evaluate(self.handle, int(variant), _dptr(lnl),
                                         _dptr(db) if gradient else None,
                                         _dptr(dp) if gradient else None, C.byref(ms))
        if rc == _lib.NRX_ERR_NEGINF:
            raise RuntimeError(self.lib.nrx_last_error().decode())
        _lib.check(rc)
        dbs, dps = [], []
        if gradient:
            eo = ro = 0
            for e, r in zip(E, R):
                dbs.append(db[eo:eo + e].copy())
                dps.append(dp[ro:ro + r].copy())
                eo += e
                ro += r
        return lnl, dbs, dps, ms.value

    def optimize(self, edges, retics, brlen_iters: int = 1, prob_iters: int = 1, variant: int = 0,
                 mode: int = _lib.NRX_DERIV_REFERENCE, n_threads: int = 0):
        """Per-candidate branch and reticulation optimisation, then one
        batched scoring (nrx_batch_optimize).  ``edges[i]`` / ``retics[i]``
        list candidate i's branches and reticulations.  Returns (lnL [n],
        status [n]) and refreshes each network's lengths and probabilities."""
        if not getattr(self, "handle", None):
            raise RuntimeError("the batch is closed")
        self._check_open()
        n = len(self.networks)
        if len(edges) != n or len(retics) != n:
            raise ValueError("one edge list and one reticulation list per candidate expected")
        eb = np.zeros(n + 1, np.int32)
        rb = np.zeros(n + 1, np.int32)
        eb[1:] = np.cumsum([len(x) for x in edges])
        rb[1:] = np.cumsum([len(x) for x in retics])
        ev = np.ascontiguousarray(np.concatenate([np.asarray(x, np.int32) for x in edges] + [np.zeros(1, np.int32)]))
        rv = np.ascontiguousarray(np.concatenate([np.asarray(x, np.int32) for x in retics] + [np.zeros(1, np.int32)]))
        o = self.networks[0].options
        lnl = np.zeros(n)
        st = np.zeros(n, np.int32)
        ip = C.POINTER(C.c_int32)
        rc = self.lib.nrx_batch_optimize(self.handle, int(variant), int(mode), eb.ctypes.data_as(ip),
                                         ev.ctypes.data_as(ip), int(brlen_iters), rb.ctypes.data_as(ip),
                                         rv.ctypes.data_as(ip), int(prob_iters), float(o.brprob_min),
                                         float(o.brprob_max), float(o.prob_tolerance), int(n_threads),
                                         _dptr(lnl), st.ctypes.data_as(ip))
        if rc == _lib.NRX_ERR_NEGINF:
            raise RuntimeError(self.lib.nrx_last_error().decode())
        _lib.check(rc)
        for a in self.networks:
            a.network.brlen = a.branch_lengths()
            if a.network.n_retic:
                a.network.retic_prob = a.reticulation_probs()
            a.cached_logl = None
        return lnl, st

    def close(self) -> None:
        if getattr(self, "handle", None):
            self.lib.nrx_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
