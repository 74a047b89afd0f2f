"""Python mirror of the continuous-batching request server (include/plk_serve.h).

A request is one likelihood evaluation a tree search, model selection or
bootstrap would make: the full tree lnL at a branch (PLKS_LNL), a Newton
optimisation of that branch (PLKS_NEWTON, optimizeOneBranch), or the lnL and
its RELL reweightings (PLKS_RELL). The server copies what it needs at submit
time, admits requests by their device footprint and runs the resident ones
in shared launches.
"""
import ctypes

import numpy as np

from . import _capi
from ._capi import (PLKS_LNL, PLKS_NEWTON, PLKS_RELL, PlkEnd, PlkInstanceDesc, PlksConfig,
                    PlksRequest, PlksResult, PlksStats, check, lib)
from .engine import OP_DTYPE, PREC_F64, PhyloTree

KINDS = {"lnl": PLKS_LNL, "newton": PLKS_NEWTON, "rell": PLKS_RELL}


class Request:
    """The inputs of one evaluation on branch `branch` of a workload's tree."""

    def __init__(self, workload, kind, branch=0, weights=None, precision=PREC_F64,
                 min_len=1e-6, max_len=10.0, max_nr=100, length_scale=1.0, safe_numeric=None):
        w = workload
        self.kind = KINDS[kind] if isinstance(kind, str) else int(kind)
        planner = PhyloTree(None, w)
        branches = planner.branches()
        a, b = branches[branch % len(branches)]
        ops, node, dad, length = planner.plan(a, b)
        self.ops = np.ascontiguousarray(ops, dtype=OP_DTYPE)
        self.node, self.dad = node, dad
        self.length = float(length) * length_scale
        # IQ-TREE runs every state count other than 4 and 20 with per-category
        # scaling (tree/phylotreesse.cpp:94-95)
        safe = int(w.nstates not in (4, 20) or bool(safe_numeric))
        self.desc = PlkInstanceDesc(w.ntaxa, w.nstates, w.ncat, w.npatterns, w.ntaxa - 2,
                                    w.state_unknown, precision, safe)
        self.states = np.ascontiguousarray(w.get("states"), np.uint8)
        self.freq = np.ascontiguousarray(w.get("freq"), np.float64)
        self.invar = np.ascontiguousarray(w.get("invar"), np.float64)
        m = w.model_arrays()
        self.model = [np.ascontiguousarray(m[k], np.float64)
                      for k in ("eval", "evec", "inv_evec", "tip", "rates", "props")]
        self.min_len, self.max_len, self.max_nr = min_len, max_len, max_nr
        self.weights = None
        self.nboot = 0
        if self.kind == PLKS_RELL:
            wts = np.atleast_2d(np.ascontiguousarray(weights, np.float32))
            assert wts.shape[1] == w.npatterns
            self.weights, self.nboot = wts, wts.shape[0]
        self.npatterns = w.npatterns
        self.nstates = w.nstates
        self.ntaxa = w.ntaxa

    def footprint(self):
        return lib().plk_instance_footprint(ctypes.byref(self.desc))

    def to_c(self):
        r = PlksRequest()
        r.desc = self.desc
        r.states = self.states.ctypes.data
        r.freq = self.freq.ctypes.data
        r.invar = self.invar.ctypes.data
        (r.eval, r.evec, r.inv_evec, r.tip_eigen, r.cat_rates,
         r.cat_props) = [a.ctypes.data for a in self.model]
        r.ops = self.ops.ctypes.data if len(self.ops) else None
        r.nops = len(self.ops)
        r.node = PlkEnd(self.node.is_tip, self.node.id)
        r.dad = PlkEnd(self.dad.is_tip, self.dad.id)
        r.length = self.length
        r.kind = self.kind
        r.min_len, r.max_len, r.max_nr = self.min_len, self.max_len, self.max_nr
        r.weights = self.weights.ctypes.data if self.weights is not None else None
        r.nboot = self.nboot
        return r


class Server:
    """plks_create / plks_submit / plks_wait on a Context."""

    def __init__(self, ctx, max_inflight=0, hbm_budget=0):
        self.ctx = ctx
        cfg = PlksConfig(int(max_inflight), int(hbm_budget))
        h = ctypes.c_void_p()
        check(lib().plks_create(ctx.handle, ctypes.byref(cfg), ctypes.byref(h)), "plks_create")
        self._h = h
        self._nboot = {}
        ctx.adopt(self)

    def submit(self, req):
        t = ctypes.c_uint64()
        c = req.to_c()
        check(lib().plks_submit(self._h, ctypes.byref(c), ctypes.byref(t)), "plks_submit")
        self._nboot[t.value] = req.nboot
        return t.value

    def wait(self, ticket, timeout=-1.0):
        """The request's result as a dict, or None when it is still pending
        after `timeout` seconds."""
        res = PlksResult()
        nb = self._nboot.get(ticket, 0)
        rell = np.zeros(max(nb, 1))
        done = ctypes.c_int(0)
        check(lib().plks_wait(self._h, ctypes.c_uint64(ticket), float(timeout), ctypes.byref(res),
                              _capi.dptr(rell), ctypes.byref(done)), "plks_wait")
        if not done.value:
            return None
        self._nboot.pop(ticket, None)
        return {"status": res.status, "lnl": res.lnl, "length": res.length,
                "n_derv": res.n_derv, "rell": rell[:res.nboot].copy() if res.nboot else None,
                "submit_s": res.submit_s, "admit_s": res.admit_s, "finish_s": res.finish_s}

    def clock(self):
        return lib().plks_clock(self._h)

    def stats(self):
        s = PlksStats()
        check(lib().plks_get_stats(self._h, ctypes.byref(s)), "plks_get_stats")
        return {f: getattr(s, f) for f, _ in PlksStats._fields_}

    def close(self):
        if getattr(self, "_h", None):
            lib().plks_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        if not _capi.shutting_down():
            self.close()
