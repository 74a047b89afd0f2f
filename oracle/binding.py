"""ctypes binding of oracle/liboracle.so (the FP64 CPU restatement).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker, never by the product.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

c_int, c_double, c_void_p = ctypes.c_int, ctypes.c_double, ctypes.c_void_p
_DP = ctypes.POINTER(c_double)


class OrcEnd(ctypes.Structure):
    _fields_ = [("is_tip", c_int), ("id", c_int)]


_PROTOS = {
    "orc_create": (c_void_p, [c_int] * 8),
    "orc_destroy": (None, [c_void_p]),
    "orc_set_threads": (None, [c_void_p, c_int]),
    "orc_set_patterns": (None, [c_void_p, ctypes.POINTER(ctypes.c_uint8), _DP, _DP]),
    "orc_set_model": (None, [c_void_p] + [_DP] * 6),
    "orc_update_partials": (None, [c_void_p, c_void_p, c_int]),
    "orc_likelihood_branch": (c_double, [c_void_p, c_void_p, c_int, OrcEnd, OrcEnd, c_double, _DP,
                                         ctypes.POINTER(c_int)]),
    "orc_likelihood_buffer": (None, [c_void_p, c_void_p, c_int, OrcEnd, OrcEnd]),
    "orc_likelihood_derv": (c_int, [c_void_p, c_double, _DP, _DP]),
    "orc_likelihood_from_buffer": (c_double, [c_void_p, c_double, _DP]),
    "orc_optimize_branch": (c_double, [c_void_p, c_void_p, c_int, OrcEnd, OrcEnd, c_double,
                                       c_double, c_double, c_int, _DP, ctypes.POINTER(c_int),
                                       ctypes.POINTER(c_int)]),
    "orc_read_partial": (None, [c_void_p, c_int, _DP, ctypes.POINTER(ctypes.c_int32)]),
    "orc_set_asc": (None, [c_void_p, c_int]),
    "orc_set_cost_matrix": (None, [c_void_p, ctypes.POINTER(ctypes.c_uint32), c_int]),
    "orc_parsimony_sankoff": (c_double, [c_void_p, c_void_p, c_int, OrcEnd, OrcEnd,
                                         ctypes.POINTER(ctypes.c_uint32), _DP]),
    "orc_set_model_nonrev": (c_int, [c_void_p, c_int, c_int] + [_DP] * 8 + [c_int, _DP, c_int]),
    "orc_parsimony_branch": (c_double, [c_void_p, c_void_p, c_int, OrcEnd, OrcEnd,
                                        ctypes.POINTER(ctypes.c_int32), _DP]),
    "orc_parsimony_fast": (c_double, [c_void_p, c_void_p, c_int, OrcEnd, OrcEnd, _DP]),
}

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        h = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _PROTOS.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _lib = h
    return _lib


def _end(e):
    return OrcEnd(int(e.is_tip), int(e.id))


def _ops(ops):
    if ops is None or len(ops) == 0:
        return None, 0
    ops = np.ascontiguousarray(ops)
    assert ops.dtype.itemsize == 40
    return ops.ctypes.data_as(c_void_p), len(ops)


class OracleTree:
    """Host restatement with the same call shape as plk.h's instance API."""

    def __init__(self, ntaxa, nstates, ncat, npatterns, nslots, state_unknown, safe_numeric=False,
                 nthreads=1):
        self._h = lib().orc_create(ntaxa, nstates, ncat, npatterns, nslots, state_unknown,
                                   int(bool(safe_numeric)), nthreads)
        self.P = npatterns
        self.nasc = 0
        self.block = nstates * ncat
        self.ncat = ncat
        self.safe = bool(safe_numeric)

    @classmethod
    def from_arrays(cls, ntaxa, nstates, ncat, state_unknown, states, freq, invar, model,
                    safe_numeric=False, nthreads=1):
        # one slot per internal node and the two NNI buffers, as the planner numbers them
        t = cls(ntaxa, nstates, ncat, len(freq), ntaxa, state_unknown, safe_numeric, nthreads)
        t.set_patterns(states, freq, invar)
        t.set_model(model["eval"], model["evec"], model["inv_evec"], model["tip"], model["rates"],
                    model["props"])
        return t

    @classmethod
    def from_workload(cls, w, safe_numeric=False, nthreads=1):
        return cls.from_arrays(w.ntaxa, w.nstates, w.ncat, w.state_unknown, w.get("states"),
                               w.get("freq"), w.get("invar"), w.model_arrays(), safe_numeric,
                               nthreads)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_destroy(self._h)
            self._h = None

    def set_asc(self, nasc):
        """+ASC (Lewis): the last nasc patterns are the unobserved constant ones;
        pattern lnL outputs then cover the observed patterns only."""
        lib().orc_set_asc(self._h, int(nasc))
        self.nasc = int(nasc)

    def set_model_nonrev(self, nm, tip, rooted=False):
        """State-basis (non-reversible / mixture) model: nm is a dict with
        ceval [nmix][n] complex, cevec and cinv_evec [nmix][n][n] complex, Q
        [nmix][n][n], pi [nmix][n], mixw [nmix], rates and props [nrates] and
        fused; tip [SU+1][n] the plain tip vectors."""
        def cplx(a):
            a = np.asarray(a, np.complex128)
            return np.ascontiguousarray(np.stack([a.real, a.imag], axis=-1), np.float64)
        arrs = [cplx(nm["ceval"]), cplx(nm["cevec"]), cplx(nm["cinv_evec"])]
        arrs += [np.ascontiguousarray(nm[k], np.float64) for k in ("Q", "pi", "mixw", "rates", "props")]
        tip = np.ascontiguousarray(tip, np.float64)
        nmix = len(nm["mixw"])
        st = lib().orc_set_model_nonrev(self._h, nmix, int(bool(nm["fused"])),
                                        *[a.ctypes.data_as(_DP) for a in arrs], len(nm["rates"]),
                                        tip.ctypes.data_as(_DP), int(bool(rooted)))
        if st != 0:
            raise ValueError("state-basis model does not match the instance's categories")

    def set_cost_matrix(self, cost, seq_type):
        """Sankoff cost matrix [n][n] (uint32); seq_type picks the tip ambiguity rules."""
        self._cost = np.ascontiguousarray(cost, np.uint32)
        lib().orc_set_cost_matrix(self._h, self._cost.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                  int(seq_type))

    def parsimony_sankoff(self, ops, node, dad, want_pattern=False):
        """(tree score, weighted cost on the branch[, per-pattern scores])."""
        p, n = _ops(ops)
        per = np.zeros(self.P, np.uint32) if want_pattern else None
        sub = c_double()
        s = lib().orc_parsimony_sankoff(self._h, p, n, _end(node), _end(dad),
                                        per.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)) if want_pattern
                                        else None, ctypes.byref(sub))
        return (s, sub.value, per) if want_pattern else (s, sub.value)

    def set_threads(self, n):
        lib().orc_set_threads(self._h, n)

    def set_patterns(self, states, freq, invar):
        self._states = np.ascontiguousarray(states, np.uint8)
        f = np.ascontiguousarray(freq, np.float64)
        v = np.ascontiguousarray(invar, np.float64)
        lib().orc_set_patterns(self._h, self._states.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                               f.ctypes.data_as(_DP), v.ctypes.data_as(_DP))

    def set_model(self, *arrs):
        a = [np.ascontiguousarray(x, np.float64) for x in arrs]
        lib().orc_set_model(self._h, *[x.ctypes.data_as(_DP) for x in a])

    def update_partials(self, ops):
        p, n = _ops(ops)
        lib().orc_update_partials(self._h, p, n)

    def likelihood_branch(self, ops, node, dad, length, want_pattern_lh=False):
        p, n = _ops(ops)
        plh = np.zeros(self.P - self.nasc)
        st = c_int()
        lnl = lib().orc_likelihood_branch(self._h, p, n, _end(node), _end(dad), float(length),
                                          plh.ctypes.data_as(_DP), ctypes.byref(st))
        return (lnl, plh) if want_pattern_lh else lnl

    def likelihood_buffer(self, ops, node, dad):
        p, n = _ops(ops)
        lib().orc_likelihood_buffer(self._h, p, n, _end(node), _end(dad))

    def likelihood_derv(self, length):
        df, ddf = c_double(), c_double()
        lib().orc_likelihood_derv(self._h, float(length), ctypes.byref(df), ctypes.byref(ddf))
        return df.value, ddf.value

    def likelihood_from_buffer(self, length, want_pattern_lh=False):
        plh = np.zeros(self.P - self.nasc)
        lnl = lib().orc_likelihood_from_buffer(self._h, float(length), plh.ctypes.data_as(_DP))
        return (lnl, plh) if want_pattern_lh else lnl

    def optimize_branch(self, ops, node, dad, length, min_len=1e-6, max_len=10.0, max_nr=100):
        p, n = _ops(ops)
        lnl, nd, st = c_double(), c_int(), c_int()
        x = lib().orc_optimize_branch(self._h, p, n, _end(node), _end(dad), float(length),
                                      float(min_len), float(max_len), int(max_nr),
                                      ctypes.byref(lnl), ctypes.byref(nd), ctypes.byref(st))
        return x, lnl.value, nd.value, st.value

    def parsimony_branch(self, ops, node, dad, want_pattern=False):
        """(score, branch_subst), with each pattern's changes appended when
        want_pattern."""
        p, n = _ops(ops)
        per = np.zeros(self.P, np.int32)
        sub = ctypes.c_double()
        score = lib().orc_parsimony_branch(self._h, p, n, _end(node), _end(dad),
                                           per.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                           ctypes.byref(sub))
        return (score, sub.value, per) if want_pattern else (score, sub.value)

    def parsimony_fast(self, ops, node, dad):
        """(score, branch_subst) by the bit-parallel form."""
        p, n = _ops(ops)
        sub = ctypes.c_double()
        score = lib().orc_parsimony_fast(self._h, p, n, _end(node), _end(dad), ctypes.byref(sub))
        return score, sub.value

    def read_partial(self, slot, partial=True):
        part = np.zeros((self.P, self.block)) if partial else None
        scale = np.zeros((self.P, self.ncat) if self.safe else (self.P,), np.int32)
        lib().orc_read_partial(self._h, int(slot),
                               part.ctypes.data_as(_DP) if partial else None,
                               scale.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
        return part, scale
