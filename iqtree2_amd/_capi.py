"""ctypes bindings of libplk.so (include/plk.h and include/plk_host.h).

The library is the product: when it is missing or fails to load this module
raises ImportError; there is no fallback implementation.
"""
import atexit
import ctypes
import os

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libplk.so")

c_int, c_double, c_size_t, c_void_p, c_char_p = (
    ctypes.c_int, ctypes.c_double, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_char_p)
P = ctypes.POINTER

PLK_OK = 0
STATUS_NAMES = {0: "PLK_OK", 1: "PLK_ERR_ARG", 2: "PLK_ERR_NOMEM", 3: "PLK_ERR_DEVICE",
                4: "PLK_ERR_NONFINITE", 5: "PLK_ERR_STATE", 6: "PLK_ERR_UNSUPPORTED",
                7: "PLK_ERR_NEWTON"}
PREC_F64, PREC_F32 = 0, 1
DNA, PROTEIN, CODON = 0, 1, 2


class PlkChild(ctypes.Structure):
    _fields_ = [("is_tip", c_int), ("id", c_int), ("length", c_double)]


class PlkOp(ctypes.Structure):
    _fields_ = [("dad_slot", c_int), ("nchildren", c_int), ("child", PlkChild * 2)]


class PlkEnd(ctypes.Structure):
    _fields_ = [("is_tip", c_int), ("id", c_int)]


class PlkConfig(ctypes.Structure):
    _fields_ = [("device", c_int), ("stream", c_void_p), ("hbm_budget", c_size_t)]


class PlkInstanceDesc(ctypes.Structure):
    _fields_ = [("ntaxa", c_int), ("nstates", c_int), ("ncat", c_int), ("npatterns", c_int),
                ("nslots", c_int), ("state_unknown", c_int), ("precision", c_int),
                ("safe_numeric", c_int)]


class PlkBatchItem(ctypes.Structure):
    _fields_ = [("instance", c_int), ("ops", c_void_p), ("nops", c_int), ("node", PlkEnd),
                ("dad", PlkEnd), ("length", c_double)]


class PlkKernelStat(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 32), ("launches", c_int), ("total_ms", c_double),
                ("bytes", c_double), ("flops", c_double)]


class PlkhWorkloadSpec(ctypes.Structure):
    _fields_ = [("seq_type", c_int), ("ntaxa", c_int), ("nsite", c_int), ("ncat", c_int),
                ("mean_len", c_double), ("alpha", c_double), ("p_inv", c_double),
                ("gap_frac", c_double), ("seed", ctypes.c_uint64)]


class PlksConfig(ctypes.Structure):
    _fields_ = [("max_inflight", c_int), ("hbm_budget", c_size_t)]


class PlksRequest(ctypes.Structure):
    _fields_ = [("desc", PlkInstanceDesc), ("states", c_void_p), ("freq", c_void_p),
                ("invar", c_void_p), ("eval", c_void_p), ("evec", c_void_p),
                ("inv_evec", c_void_p), ("tip_eigen", c_void_p), ("cat_rates", c_void_p),
                ("cat_props", c_void_p), ("ops", c_void_p), ("nops", c_int), ("node", PlkEnd),
                ("dad", PlkEnd), ("length", c_double), ("kind", c_int), ("min_len", c_double),
                ("max_len", c_double), ("max_nr", c_int), ("weights", c_void_p),
                ("nboot", c_int)]


class PlksResult(ctypes.Structure):
    _fields_ = [("status", c_int), ("lnl", c_double), ("length", c_double), ("n_derv", c_int),
                ("nboot", c_int), ("submit_s", c_double), ("admit_s", c_double),
                ("finish_s", c_double)]


class PlksStats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_long), ("traversal_batches", ctypes.c_long),
                ("derv_batches", ctypes.c_long), ("completed", ctypes.c_long),
                ("peak_inflight", c_int), ("peak_bytes", c_size_t), ("admit_s", c_double),
                ("traverse_s", c_double), ("derv_s", c_double), ("retire_s", c_double)]


PLKS_LNL, PLKS_NEWTON, PLKS_RELL = 0, 1, 2


# numpy view of an op list with the memory layout of plk_op
OP_DTYPE = np.dtype([("dad_slot", "<i4"), ("nchildren", "<i4"),
                     ("c0_tip", "<i4"), ("c0_id", "<i4"), ("c0_len", "<f8"),
                     ("c1_tip", "<i4"), ("c1_id", "<i4"), ("c1_len", "<f8")])
assert OP_DTYPE.itemsize == ctypes.sizeof(PlkOp) == 40

_DP = P(c_double)
_U8 = P(ctypes.c_uint8)
_I32 = P(ctypes.c_int32)
_F32 = P(ctypes.c_float)
BRENT_FN = ctypes.CFUNCTYPE(c_double, c_double, c_void_p)   # plkh_minimize_one_dimen's f(x, arg)

# name -> (restype, argtypes)
_PROTOS = {
    # plk.h
    "plk_create": (c_int, [P(PlkConfig), P(c_void_p)]),
    "plk_destroy": (c_int, [c_void_p]),
    "plk_last_error": (c_char_p, []),
    "plk_memory": (c_int, [c_void_p, P(c_size_t), P(c_size_t)]),
    "plk_stream": (c_void_p, [c_void_p]),
    "plk_synchronize": (c_int, [c_void_p]),
    "plk_instance_footprint": (c_size_t, [P(PlkInstanceDesc)]),
    "plk_instance_create": (c_int, [c_void_p, P(PlkInstanceDesc), P(c_int)]),
    "plk_instance_destroy": (c_int, [c_void_p, c_int]),
    "plk_set_patterns": (c_int, [c_void_p, c_int, _U8, _DP, _DP]),
    "plk_set_model": (c_int, [c_void_p, c_int, _DP, _DP, _DP, _DP, _DP, _DP]),
    "plk_update_partials": (c_int, [c_void_p, c_int, c_void_p, c_int]),
    "plk_likelihood_branch": (c_int, [c_void_p, c_int, c_void_p, c_int, PlkEnd, PlkEnd, c_double,
                                      _DP, _DP]),
    "plk_likelihood_buffer": (c_int, [c_void_p, c_int, c_void_p, c_int, PlkEnd, PlkEnd]),
    "plk_likelihood_derv": (c_int, [c_void_p, c_int, c_double, _DP, _DP, _DP]),
    "plk_likelihood_from_buffer": (c_int, [c_void_p, c_int, c_double, _DP, _DP]),
    "plk_optimize_branch": (c_int, [c_void_p, c_int, c_void_p, c_int, PlkEnd, PlkEnd, c_double,
                                    c_double, c_double, c_int, _DP, _DP, P(c_int)]),
    "plk_rell": (c_int, [c_void_p, c_int, _F32, c_int, _DP]),
    "plk_likelihood_branch_batch": (c_int, [c_void_p, P(PlkBatchItem), c_int, _DP]),
    "plk_likelihood_derv_batch": (c_int, [c_void_p, _I32, _DP, c_int, _DP, _DP, _DP]),
    "plk_likelihood_derv_batch_noalloc": (c_int, [c_void_p, _I32, _DP, c_int, _DP, _DP, _DP]),
    "plk_copy_bandwidth": (c_int, [c_void_p, c_size_t, c_int, _DP, _DP]),
    "plk_set_asc": (c_int, [c_void_p, c_int, c_int]),
    "plk_parsimony_branch": (c_int, [c_void_p, c_int, c_void_p, c_int, PlkEnd, PlkEnd, _DP, _DP]),
    "plkh_tree_plan_full": (c_int, [c_void_p, c_int, c_int, c_void_p, c_int, P(c_int), P(PlkEnd),
                                    P(PlkEnd)]),
    "plkh_tree_parsimony_branch": (c_int, [c_void_p, c_int, c_int, _DP, _DP]),
    "plk_batch_evaluate": (c_int, [c_void_p, P(PlkBatchItem), _I32, c_int, _DP]),
    "plk_optimize_sweep": (c_int, [c_void_p, c_int, c_void_p, c_int, c_void_p, c_int, _DP, c_int,
                                   c_double, c_double, c_int, _I32, _DP]),
    "plk_read_partial": (c_int, [c_void_p, c_int, c_int, _DP, _I32]),
    "plk_profile_start": (c_int, [c_void_p]),
    "plk_profile_stop": (c_int, [c_void_p, P(PlkKernelStat), c_int, P(c_int)]),
    # plk_serve.h
    "plks_create": (c_int, [c_void_p, P(PlksConfig), P(c_void_p)]),
    "plks_destroy": (c_int, [c_void_p]),
    "plks_submit": (c_int, [c_void_p, P(PlksRequest), P(ctypes.c_uint64)]),
    "plks_wait": (c_int, [c_void_p, ctypes.c_uint64, c_double, P(PlksResult), _DP, P(c_int)]),
    "plks_get_stats": (c_int, [c_void_p, P(PlksStats)]),
    "plks_clock": (c_double, [c_void_p]),
    # plk_host.h
    "plkh_num_states": (c_int, [c_int]),
    "plkh_state_unknown": (c_int, [c_int]),
    "plkh_rate_matrix": (c_int, [c_int, _DP, _DP, _DP]),
    "plkh_decompose": (c_int, [c_int, _DP, _DP, _DP, _DP, _DP]),
    "plkh_gamma_rates": (c_int, [c_int, c_double, c_double, _DP, _DP]),
    "plk_set_cost_matrix": (c_int, [c_void_p, c_int, P(ctypes.c_uint32)]),
    "plk_parsimony_sankoff": (c_int, [c_void_p, c_int, c_void_p, c_int, PlkEnd, PlkEnd, _DP, _DP]),
    "plk_set_model_nonrev": (c_int, [c_void_p, c_int, c_int, c_int] + [_DP] * 8 + [c_int, _DP, c_int]),
    "plkh_uninformative_patterns": (c_int, [c_int, c_int, ctypes.c_long, _U8, ctypes.c_long,
                                            ctypes.POINTER(ctypes.c_long)]),
    "plkh_minimize_one_dimen": (c_int, [BRENT_FN, c_void_p, c_double, c_double, c_double, c_double, _DP, _DP,
                                        ctypes.POINTER(c_int)]),
    "plkh_tree_set_gamma_shape": (c_int, [c_void_p, c_double]),
    "plkh_tree_optimize_gamma_shape": (c_int, [c_void_p, c_double, _DP, _DP, ctypes.POINTER(c_int)]),
    "plkh_models_optimize_gamma_shape": (c_int, [P(c_void_p), c_int, c_double, _DP, _DP, _I32, _I32]),
    "plkh_tip_vectors": (c_int, [c_int, _DP, _DP]),
    "plkh_codon_exchangeabilities": (c_int, [c_double, c_double, _DP]),
    "plkh_patterns_build": (c_void_p, [c_int, c_int, c_int, c_char_p, P(c_int)]),
    "plkh_patterns_destroy": (None, [c_void_p]),
    "plkh_patterns_dims": (c_int, [c_void_p, P(c_int), P(c_int), P(c_int)]),
    "plkh_patterns_get": (c_int, [c_void_p, _U8, _I32, _I32, _I32]),
    "plkh_patterns_invar": (c_int, [c_void_p, _DP, c_double, _DP]),
    "plkh_workload_create": (c_void_p, [P(PlkhWorkloadSpec), P(c_int)]),
    "plkh_workload_create_on_tree": (c_void_p, [P(PlkhWorkloadSpec), c_int, _I32, _I32, _DP,
                                                P(c_int)]),
    "plkh_workload_destroy": (None, [c_void_p]),
    "plkh_workload_dims": (c_int, [c_void_p] + [P(c_int)] * 6),
    "plkh_workload_get": (c_int, [c_void_p, c_char_p, c_void_p]),
    "plkh_tree_create": (c_void_p, [c_void_p, c_void_p, c_int, c_int, P(c_int)]),
    "plkh_tree_destroy": (None, [c_void_p]),
    "plkh_tree_instance": (c_int, [c_void_p]),
    "plkh_tree_num_branches": (c_int, [c_void_p]),
    "plkh_tree_branches": (c_int, [c_void_p, _I32, _I32]),
    "plkh_tree_get_length": (c_int, [c_void_p, c_int, c_int, _DP]),
    "plkh_tree_set_length": (c_int, [c_void_p, c_int, c_int, c_double]),
    "plkh_tree_clear_all_partial": (c_int, [c_void_p]),
    "plkh_tree_reset": (c_int, [c_void_p]),
    "plkh_tree_plan": (c_int, [c_void_p, c_int, c_int, c_void_p, c_int, P(c_int), P(PlkEnd),
                               P(PlkEnd), _DP]),
    "plkh_tree_compute_likelihood": (c_int, [c_void_p, _DP, _DP]),
    "plkh_tree_likelihood_branch": (c_int, [c_void_p, c_int, c_int, _DP, _DP]),
    "plkh_tree_likelihood_derv": (c_int, [c_void_p, c_int, c_int, _DP, _DP]),
    "plkh_tree_optimize_branch": (c_int, [c_void_p, c_int, c_int, c_int, _DP, P(c_int)]),
    "plkh_tree_optimize_all_branches": (c_int, [c_void_p, c_int, c_double, c_int, _DP, P(c_int)]),
    "plkh_tree_set_branch_bounds": (c_int, [c_void_p, c_double, c_double]),
    "plkh_tree_set_device_newton": (c_int, [c_void_p, c_int]),
    "plkh_tree_pass_counts": (c_int, [c_void_p, _I32, c_int, P(c_int)]),
    "plkh_supertree_optimize_all_branches": (c_int, [P(c_void_p), c_int, c_int, c_double, c_int, _DP,
                                                     _I32, _I32]),
    "plkh_tree_neighbors": (c_int, [c_void_p, c_int, _I32, P(c_int)]),
    "plkh_tree_set_length_keep": (c_int, [c_void_p, c_int, c_int, c_double]),
    "plkh_tree_nni_begin": (c_int, [c_void_p, c_int, c_int]),
    "plkh_tree_nni_swap": (c_int, [c_void_p, c_int, c_int, c_int, c_int]),
    "plkh_tree_clear_partial": (c_int, [c_void_p, c_int, c_int]),
    "plkh_tree_nni_end": (c_int, [c_void_p]),
    "plkh_tree_inner_branches": (c_int, [c_void_p, _I32, _I32, c_int, P(c_int)]),
    "plkh_tree_nni_evaluate": (c_int, [c_void_p, c_int, c_int, c_int, _DP, _DP, _I32]),
    "plkh_tree_nni_evaluate_all": (c_int, [c_void_p, c_int, _DP, _DP, _I32, c_int]),
    "plk_optimize_sweep_ex": (c_int, [c_void_p, c_int, c_void_p, c_int, c_void_p, c_int, _DP, c_int,
                                      c_double, c_double, c_int, _I32, _DP, _DP, c_int, _DP]),
}

_lib = None

# Set when the interpreter starts shutting down: finalizers (__del__) then
# make no further library calls. Device objects must be closed explicitly
# before that (Context.close() closes what was created on it); a call into
# the HIP runtime from a finalizer at exit can run after the runtime, or a
# profiler attached to it, has been torn down.
_shutting_down = False


def _mark_shutdown():
    global _shutting_down
    _shutting_down = True


atexit.register(_mark_shutdown)


def shutting_down():
    return _shutting_down


def lib():
    """The loaded libplk.so with prototypes set (ImportError when absent)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"libplk.so not built at {LIB_PATH}; run __graft_entry__.build()")
        try:
            handle = ctypes.CDLL(LIB_PATH)
        except OSError as e:
            raise ImportError(f"failed to load {LIB_PATH}: {e}") from e
        for name, (res, args) in _PROTOS.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def exported_names():
    return list(_PROTOS)


class PlkError(RuntimeError):
    def __init__(self, status, where=""):
        msg = lib().plk_last_error().decode(errors="replace")
        self.status = status
        super().__init__(f"{where}: {STATUS_NAMES.get(status, status)}: {msg}")


def check(status, where=""):
    if status != PLK_OK:
        raise PlkError(status, where)
    return status


def dptr(a):
    return a.ctypes.data_as(_DP)


def u8ptr(a):
    return a.ctypes.data_as(_U8)


def i32ptr(a):
    return a.ctypes.data_as(_I32)


def f32ptr(a):
    return a.ctypes.data_as(_F32)
