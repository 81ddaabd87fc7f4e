"""ctypes binding of ``libnetrax_gpu.so`` (the C ABI in include/netrax_gpu.h)."""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnetrax_gpu.so")

NRX_OK = 0
NRX_ERR_ARG = 1
NRX_ERR_NEGINF = 2
NRX_ERR_DEVICE = 3
NRX_ERR_UNSUPPORTED = 4
NRX_ERR_STATE = 5
NRX_FP64 = 0
NRX_FP32 = 1
NRX_DERIV_REFERENCE = 0
NRX_DERIV_TRUE = 1

# every symbol include/netrax_gpu.h declares
EXPORTED = [
    "nrx_default_options", "nrx_last_error", "nrx_create", "nrx_destroy", "nrx_set_model",
    "nrx_set_category_rates", "nrx_set_branch_lengths", "nrx_set_reticulation_probs",
    "nrx_loglikelihood", "nrx_gradient", "nrx_brlen_derivatives", "nrx_brlen_derivatives_noalloc",
    "nrx_tree_logls", "nrx_optimize_branches", "nrx_get_branch_lengths", "nrx_model_gradient",
    "nrx_prob_derivatives", "nrx_optimize_reticulations", "nrx_get_reticulation_probs", "nrx_batch_optimize",
    "nrx_set_memory_budget", "nrx_memory_plan",
    "nrx_get_stats", "nrx_last_timing", "nrx_batch_create", "nrx_batch_destroy",
    "nrx_batch_evaluate", "nrx_autotune", "nrx_set_tuning", "nrx_get_tuning",
]


class NetworkDesc(C.Structure):
    _fields_ = [("n_nodes", C.c_int32), ("n_tips", C.c_int32), ("root", C.c_int32),
                ("n_edges", C.c_int32), ("n_retic", C.c_int32),
                ("edge_parent", C.POINTER(C.c_int32)), ("edge_child", C.POINTER(C.c_int32)),
                ("retic_first_edge", C.POINTER(C.c_int32)),
                ("retic_second_edge", C.POINTER(C.c_int32))]


class PartitionDesc(C.Structure):
    _fields_ = [("n_sites", C.c_int32), ("states", C.c_int32), ("rate_cats", C.c_int32),
                ("tip_masks", C.POINTER(C.c_uint32)), ("weights", C.POINTER(C.c_uint32))]


class Options(C.Structure):
    _fields_ = [("precision", C.c_int32), ("device", C.c_int32),
                ("min_tree_logprob", C.c_double), ("brlen_min", C.c_double),
                ("brlen_max", C.c_double), ("use_graph", C.c_int32), ("timing", C.c_int32)]


class Stats(C.Structure):
    _fields_ = [("n_variants", C.c_int32), ("n_ops_needed", C.c_int32), ("n_levels", C.c_int32),
                ("n_classes", C.c_int32), ("n_root_clvs", C.c_int32), ("n_uses", C.c_int32),
                ("n_inner_uses", C.c_int32), ("n_lds_handoffs", C.c_int32), ("n_sites_padded", C.c_int64), ("fwd_op_bytes", C.c_double),
                ("root_bytes", C.c_double), ("adj_bytes", C.c_double), ("fwd_bytes", C.c_double),
                ("grad_bytes", C.c_double), ("clv_bytes", C.c_double)]


class Timing(C.Structure):
    _fields_ = [("total_ms", C.c_double), ("fwd_ms", C.c_double), ("fwd_launches", C.c_int32),
                ("adj_launches", C.c_int32), ("adj_ms", C.c_double)]


class Tuning(C.Structure):
    _fields_ = [("group_width", C.c_int32), ("tile_sites", C.c_int32), ("row_lanes", C.c_int32),
                ("ms", C.c_double)]


class MemoryPlan(C.Structure):
    _fields_ = [("segments", C.c_int32), ("rows", C.c_int32), ("boundary_rows", C.c_int32),
                ("resident_rows", C.c_int32), ("row_bytes", C.c_double)]


_lib = None


def load(build_if_missing: bool = True):
    """Loads the engine library; builds it first when no build exists."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        if not build_if_missing:
            raise RuntimeError(f"{LIB_PATH} is missing: run python -m netrax_amd.build")
        from . import build as _build
        _build.build()
    lib = C.CDLL(LIB_PATH)
    pd = C.POINTER(C.c_double)
    lib.nrx_last_error.restype = C.c_char_p
    lib.nrx_default_options.argtypes = [C.POINTER(Options)]
    lib.nrx_default_options.restype = None
    lib.nrx_create.argtypes = [C.POINTER(NetworkDesc), C.POINTER(PartitionDesc), C.c_int32,
                               C.POINTER(Options), C.POINTER(C.c_void_p)]
    lib.nrx_destroy.argtypes = [C.c_void_p]
    lib.nrx_set_model.argtypes = [C.c_void_p, C.c_int32, pd, pd, C.c_double]
    lib.nrx_set_category_rates.argtypes = [C.c_void_p, C.c_int32, pd]
    lib.nrx_set_branch_lengths.argtypes = [C.c_void_p, pd]
    lib.nrx_set_reticulation_probs.argtypes = [C.c_void_p, pd]
    lib.nrx_loglikelihood.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, pd, pd]
    lib.nrx_gradient.argtypes = [C.c_void_p, C.c_int32, pd, pd, pd, pd]
    lib.nrx_brlen_derivatives.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_double,
                                          pd, pd, pd, pd]
    lib.nrx_brlen_derivatives_noalloc.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_double,
                                                  pd, pd, pd, pd]
    lib.nrx_tree_logls.argtypes = [C.c_void_p, pd, pd]
    lib.nrx_optimize_branches.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.POINTER(C.c_int32), C.c_int32,
                                          C.c_int32, pd]
    lib.nrx_get_branch_lengths.argtypes = [C.c_void_p, pd]
    lib.nrx_model_gradient.argtypes = [C.c_void_p, C.c_int32, pd, pd, pd, pd, pd]
    lib.nrx_prob_derivatives.argtypes = [C.c_void_p, C.c_int32, pd, pd, pd]
    lib.nrx_optimize_reticulations.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_double, C.c_double,
                                               C.c_double, pd]
    lib.nrx_get_reticulation_probs.argtypes = [C.c_void_p, pd]
    lib.nrx_set_memory_budget.argtypes = [C.c_void_p, C.c_uint64]
    lib.nrx_memory_plan.argtypes = [C.c_void_p, C.POINTER(MemoryPlan)]
    pi = C.POINTER(C.c_int32)
    lib.nrx_batch_optimize.argtypes = [C.c_void_p, C.c_int32, C.c_int32, pi, pi, C.c_int32, pi, pi, C.c_int32,
                                       C.c_double, C.c_double, C.c_double, C.c_int32, pd, pi]
    lib.nrx_get_stats.argtypes = [C.c_void_p, C.POINTER(Stats)]
    lib.nrx_last_timing.argtypes = [C.c_void_p, C.POINTER(Timing)]
    lib.nrx_batch_create.argtypes = [C.POINTER(C.c_void_p), C.c_int32, C.POINTER(C.c_void_p)]
    lib.nrx_batch_destroy.argtypes = [C.c_void_p]
    lib.nrx_batch_evaluate.argtypes = [C.c_void_p, C.c_int32, pd, pd, pd, pd]
    lib.nrx_autotune.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.POINTER(Tuning),
                                 C.POINTER(Tuning), C.c_int32, C.POINTER(C.c_int32)]
    lib.nrx_set_tuning.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32]
    lib.nrx_get_tuning.argtypes = [C.c_void_p, C.POINTER(Tuning)]
    _lib = lib
    return lib


class NrxError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(msg)
        self.code = code


def check(rc: int) -> None:
    if rc != NRX_OK:
        raise NrxError(rc, load().nrx_last_error().decode())
