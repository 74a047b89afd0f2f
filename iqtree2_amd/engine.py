"""Python host mirror of the likelihood engine.

Thin wrappers over libplk.so: `Context` / `Instance` expose the drop-in
C-ABI (include/plk.h) whose entry points replace the PhyloTree kernel-set
members computePartialLikelihood, computeLikelihoodBranch,
computeLikelihoodDerv and computeLikelihoodFromBuffer
(tree/phylotree.h:899-1001, 1340-1346); `PhyloTree` is the C++ driver that
plans traversals and runs Newton-Raphson branch optimisation the way
PhyloTree does (tree/phylotree.cpp:2594-2773, 5918-6028); `Workload` and
`Patterns` expose the host data formats (site patterns, model eigensystem).
"""
import ctypes
import weakref

import numpy as np

from . import _capi
from ._capi import (OP_DTYPE, PREC_F32, PREC_F64, PlkConfig, PlkEnd, PlkError,
                    PlkhWorkloadSpec, PlkInstanceDesc, check, dptr, f32ptr, i32ptr, lib,
                    shutting_down, u8ptr)

__all__ = ["Context", "Instance", "PhyloTree", "Workload", "Patterns", "end", "ops_array",
           "likelihood_branch_batch",
           "PREC_F32", "PREC_F64", "PlkError"]


def end(is_tip, id_):
    return PlkEnd(int(is_tip), int(id_))


def ops_array(n=0):
    return np.zeros(n, dtype=OP_DTYPE)


def _ops_ptr(ops):
    if ops is None or len(ops) == 0:
        return None, 0
    ops = np.ascontiguousarray(ops, dtype=OP_DTYPE)
    return ops.ctypes.data_as(ctypes.c_void_p), len(ops)


class Workload:
    """Seeded synthetic likelihood input (tree, model, patterns)."""

    def __init__(self, seq_type=0, ntaxa=16, nsite=1000, ncat=4, mean_len=0.1, alpha=0.5,
                 p_inv=0.0, gap_frac=0.0, seed=1, edges=None):
        spec = PlkhWorkloadSpec(seq_type, ntaxa, nsite, ncat, mean_len, alpha, p_inv, gap_frac,
                                seed)
        st = ctypes.c_int(0)
        L = lib()
        if edges is None:
            h = L.plkh_workload_create(ctypes.byref(spec), ctypes.byref(st))
        else:
            u, v, ln = (np.ascontiguousarray(edges[0], np.int32),
                        np.ascontiguousarray(edges[1], np.int32),
                        np.ascontiguousarray(edges[2], np.float64))
            h = L.plkh_workload_create_on_tree(ctypes.byref(spec), len(u), i32ptr(u), i32ptr(v),
                                               dptr(ln), ctypes.byref(st))
        if not h:
            raise PlkError(st.value, "plkh_workload_create")
        self._h = h
        self.seq_type = seq_type
        dims = [ctypes.c_int() for _ in range(6)]
        check(L.plkh_workload_dims(h, *[ctypes.byref(d) for d in dims]), "plkh_workload_dims")
        (self.ntaxa, self.nstates, self.ncat, self.npatterns, self.state_unknown,
         self.nsite) = [d.value for d in dims]

    def close(self):
        if getattr(self, "_h", None):
            lib().plkh_workload_destroy(self._h)
            self._h = None

    def __del__(self):
        if not shutting_down():
            self.close()

    @property
    def handle(self):
        return self._h

    def get(self, name):
        n, C, P, SU, N = self.nstates, self.ncat, self.npatterns, self.state_unknown, self.ntaxa
        shapes = {
            "states": ((N, P), np.uint8), "freq": ((P,), np.float64), "invar": ((P,), np.float64),
            "eval": ((n,), np.float64), "evec": ((n, n), np.float64),
            "inv_evec": ((n, n), np.float64), "Q": ((n, n), np.float64),
            "pi": ((n,), np.float64), "tip": ((SU + 1, n), np.float64),
            "rates": ((C,), np.float64), "props": ((C,), np.float64),
            "exch": ((n * (n - 1) // 2,), np.float64),
            "edges_u": ((2 * N - 3,), np.int32), "edges_v": ((2 * N - 3,), np.int32),
            "edges_len": ((2 * N - 3,), np.float64),
            "site_pattern": ((self.nsite,), np.int32), "const_char": ((P,), np.int32),
            "params": ((4,), np.float64),
        }
        shape, dt = shapes[name]
        out = np.zeros(shape, dtype=dt)
        check(lib().plkh_workload_get(self._h, name.encode(), out.ctypes.data_as(ctypes.c_void_p)),
              f"plkh_workload_get({name})")
        return out

    def model_arrays(self):
        return {k: self.get(k) for k in ("eval", "evec", "inv_evec", "tip", "rates", "props")}

    def edges(self):
        return self.get("edges_u"), self.get("edges_v"), self.get("edges_len")


class Patterns:
    """Site-pattern compression of an alignment (alignment/alignment.cpp:1725-1929)."""

    def __init__(self, seq_type, rows):
        rows = [r.encode() if isinstance(r, str) else bytes(r) for r in rows]
        nchar = len(rows[0])
        if any(len(r) != nchar for r in rows):
            raise ValueError("sequences differ in length")
        buf = b"".join(rows)
        st = ctypes.c_int(0)
        h = lib().plkh_patterns_build(seq_type, len(rows), nchar, buf, ctypes.byref(st))
        if not h:
            raise PlkError(st.value, "plkh_patterns_build")
        self._h = h
        self.seq_type = seq_type
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        lib().plkh_patterns_dims(h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
        self.nseq, self.nsite, self.npatterns = a.value, b.value, c.value
        self.states = np.zeros((self.nseq, self.npatterns), np.uint8)
        self.freq = np.zeros(self.npatterns, np.int32)
        self.const_char = np.zeros(self.npatterns, np.int32)
        self.site_pattern = np.zeros(self.nsite, np.int32)
        check(lib().plkh_patterns_get(h, u8ptr(self.states), i32ptr(self.freq),
                                      i32ptr(self.const_char), i32ptr(self.site_pattern)),
              "plkh_patterns_get")

    def invar(self, pi, p_inv):
        out = np.zeros(self.npatterns)
        pi = np.ascontiguousarray(pi, np.float64)
        check(lib().plkh_patterns_invar(self._h, dptr(pi), float(p_inv), dptr(out)),
              "plkh_patterns_invar")
        return out

    def __del__(self):
        if getattr(self, "_h", None) and not shutting_down():
            lib().plkh_patterns_destroy(self._h)
            self._h = None


class Context:
    """plk_context: one device, one stream, an HBM budget. close() first
    closes every instance, tree and server created on the context."""

    def __init__(self, device=0, stream=None, hbm_budget=0):
        cfg = PlkConfig(device, stream, hbm_budget)
        h = ctypes.c_void_p()
        self._children = weakref.WeakSet()
        check(lib().plk_create(ctypes.byref(cfg), ctypes.byref(h)), "plk_create")
        self._h = h

    def adopt(self, child):
        """Registers an object holding device state on this context."""
        self._children.add(child)

    @property
    def handle(self):
        return self._h

    @property
    def stream(self):
        return lib().plk_stream(self._h)

    def memory(self):
        u, b = ctypes.c_size_t(), ctypes.c_size_t()
        check(lib().plk_memory(self._h, ctypes.byref(u), ctypes.byref(b)), "plk_memory")
        return u.value, b.value

    def synchronize(self):
        check(lib().plk_synchronize(self._h), "plk_synchronize")

    def copy_bandwidth(self, nbytes=256 << 20, reps=10):
        """HBM bandwidth of a plain device copy of nbytes (read + written per
        launch): (GB/s, average ms per launch)."""
        g, ms = ctypes.c_double(), ctypes.c_double()
        check(lib().plk_copy_bandwidth(self._h, int(nbytes), int(reps), ctypes.byref(g),
                                       ctypes.byref(ms)), "plk_copy_bandwidth")
        return g.value, ms.value

    def profile_start(self):
        check(lib().plk_profile_start(self._h), "plk_profile_start")

    def profile_stop(self):
        """Per-kernel {launches, total_ms, bytes, flops} since profile_start."""
        stats = (_capi.PlkKernelStat * 32)()
        n = ctypes.c_int()
        check(lib().plk_profile_stop(self._h, stats, 32, ctypes.byref(n)), "plk_profile_stop")
        return {s.name.decode(): dict(launches=s.launches, total_ms=s.total_ms, bytes=s.bytes,
                                      flops=s.flops) for s in stats[:n.value]}

    def close(self):
        if getattr(self, "_h", None):
            for child in list(self._children):
                child.close()
            lib().plk_destroy(self._h)
            self._h = None

    def __del__(self):
        if not shutting_down():
            self.close()


class Instance:
    """One PhyloTree's device state (plk_instance_*)."""

    def __init__(self, ctx, ntaxa, nstates, ncat, npatterns, nslots, state_unknown,
                 precision=PREC_F64, safe_numeric=False):
        self.ctx = ctx
        self.desc = PlkInstanceDesc(ntaxa, nstates, ncat, npatterns, nslots, state_unknown,
                                    precision, int(bool(safe_numeric)))
        iid = ctypes.c_int(-1)
        check(lib().plk_instance_create(ctx.handle, ctypes.byref(self.desc), ctypes.byref(iid)),
              "plk_instance_create")
        self.id = iid.value
        ctx.adopt(self)
        self.P = npatterns
        self.block = nstates * ncat
        self.safe = bool(safe_numeric)
        self.ncat = ncat

    @staticmethod
    def footprint(ntaxa, nstates, ncat, npatterns, nslots, state_unknown, precision=PREC_F64,
                  safe_numeric=False):
        d = PlkInstanceDesc(ntaxa, nstates, ncat, npatterns, nslots, state_unknown, precision,
                            int(bool(safe_numeric)))
        return lib().plk_instance_footprint(ctypes.byref(d))

    @classmethod
    def from_workload(cls, ctx, w, precision=PREC_F64, safe_numeric=False):
        """Instance with the planner's slots: one per internal node (ntaxa - 2)
        and the two NNI buffers."""
        inst = cls(ctx, w.ntaxa, w.nstates, w.ncat, w.npatterns, w.ntaxa, w.state_unknown,
                   precision, safe_numeric)
        inst.set_patterns(w.get("states"), w.get("freq"), w.get("invar"))
        m = w.model_arrays()
        inst.set_model(m["eval"], m["evec"], m["inv_evec"], m["tip"], m["rates"], m["props"])
        return inst

    def close(self):
        if getattr(self, "id", -1) >= 0 and self.ctx.handle:
            lib().plk_instance_destroy(self.ctx.handle, self.id)
        self.id = -1

    def __del__(self):
        if not shutting_down():
            self.close()

    def parsimony_branch(self, ops, node, dad):
        """Fitch parsimony (plk_parsimony_branch) after ops: (the tree's
        frequency-weighted changes, the changes on branch (node, dad))."""
        p, n = _ops_ptr(ops)
        score, sub = ctypes.c_double(), ctypes.c_double()
        check(lib().plk_parsimony_branch(self.ctx.handle, self.id, p, n, node, dad, ctypes.byref(score),
                                         ctypes.byref(sub)),
              "plk_parsimony_branch")
        return score.value, sub.value

    def set_cost_matrix(self, cost):
        """Sankoff cost matrix [n][n] (state i to j, unsigned)."""
        self._cost = np.ascontiguousarray(cost, np.uint32)
        check(lib().plk_set_cost_matrix(self.ctx.handle, self.id,
                                        self._cost.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))),
              "plk_set_cost_matrix")

    def parsimony_sankoff(self, ops, node, dad):
        """computeParsimonyBranchSankoff: (tree score, weighted cost on the branch)."""
        p, n = _ops_ptr(ops)
        score, sub = ctypes.c_double(), ctypes.c_double()
        check(lib().plk_parsimony_sankoff(self.ctx.handle, self.id, p, n, node, dad, ctypes.byref(score),
                                          ctypes.byref(sub)), "plk_parsimony_sankoff")
        return score.value, sub.value

    def set_model_nonrev(self, nm, tip, rooted=False):
        """State-basis (non-reversible / mixture) model: nm holds ceval
        [nmix][n], cevec and cinv_evec [nmix][n][n] (complex), Q [nmix][n][n],
        pi [nmix][n], mixw [nmix], rates and props [nrates] and fused; tip
        [SU+1][n] the plain tip vectors. Op children and ends then carry the
        flags 2 (toward the root) and 4 (the root leaf) in is_tip."""
        def cplx(a):
            a = np.asarray(a, np.complex128)
            return np.ascontiguousarray(np.stack([a.real, a.imag], axis=-1), np.float64)
        arrs = [cplx(nm["ceval"]), cplx(nm["cevec"]), cplx(nm["cinv_evec"])]
        arrs += [np.ascontiguousarray(nm[k], np.float64) for k in ("Q", "pi", "mixw", "rates", "props")]
        t = np.ascontiguousarray(tip, np.float64)
        check(lib().plk_set_model_nonrev(self.ctx.handle, self.id, len(nm["mixw"]), int(bool(nm["fused"])),
                                         *[dptr(a) for a in arrs], len(nm["rates"]), dptr(t),
                                         int(bool(rooted))), "plk_set_model_nonrev")

    def set_asc(self, nasc):
        """Lewis's ascertainment bias correction: the last nasc patterns are the
        unobserved ones (weight 0, +I mass 0): the constant patterns for +ASC,
        uninformative_patterns() for +ASC_INF."""
        check(lib().plk_set_asc(self.ctx.handle, self.id, int(nasc)), "plk_set_asc")

    def set_patterns(self, states, freq, invar):
        s = np.ascontiguousarray(states, np.uint8)
        f = np.ascontiguousarray(freq, np.float64)
        v = np.ascontiguousarray(invar, np.float64)
        check(lib().plk_set_patterns(self.ctx.handle, self.id, u8ptr(s), dptr(f), dptr(v)),
              "plk_set_patterns")

    def set_model(self, eval_, evec, inv_evec, tip, rates, props):
        arrs = [np.ascontiguousarray(a, np.float64) for a in (eval_, evec, inv_evec, tip, rates,
                                                                props)]
        check(lib().plk_set_model(self.ctx.handle, self.id, *[dptr(a) for a in arrs]),
              "plk_set_model")

    def update_partials(self, ops):
        p, n = _ops_ptr(ops)
        check(lib().plk_update_partials(self.ctx.handle, self.id, p, n), "plk_update_partials")

    def likelihood_branch(self, ops, node, dad, length, want_pattern_lh=False):
        p, n = _ops_ptr(ops)
        lnl = ctypes.c_double()
        plh = np.zeros(self.P) if want_pattern_lh else None
        check(lib().plk_likelihood_branch(self.ctx.handle, self.id, p, n, node, dad, float(length),
                                          ctypes.byref(lnl), dptr(plh) if plh is not None else None),
              "plk_likelihood_branch")
        return (lnl.value, plh) if want_pattern_lh else lnl.value

    def likelihood_buffer(self, ops, node, dad):
        p, n = _ops_ptr(ops)
        check(lib().plk_likelihood_buffer(self.ctx.handle, self.id, p, n, node, dad),
              "plk_likelihood_buffer")

    def likelihood_derv(self, length):
        df, ddf, lnl = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        check(lib().plk_likelihood_derv(self.ctx.handle, self.id, float(length), ctypes.byref(df),
                                        ctypes.byref(ddf), ctypes.byref(lnl)), "plk_likelihood_derv")
        return df.value, ddf.value, lnl.value

    def likelihood_from_buffer(self, length, want_pattern_lh=False):
        lnl = ctypes.c_double()
        plh = np.zeros(self.P) if want_pattern_lh else None
        check(lib().plk_likelihood_from_buffer(self.ctx.handle, self.id, float(length),
                                               ctypes.byref(lnl),
                                               dptr(plh) if plh is not None else None),
              "plk_likelihood_from_buffer")
        return (lnl.value, plh) if want_pattern_lh else lnl.value

    def optimize_branch(self, ops, node, dad, length, min_len=1e-6, max_len=10.0, max_nr=100):
        p, n = _ops_ptr(ops)
        out_len, out_lnl, nd = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
        check(lib().plk_optimize_branch(self.ctx.handle, self.id, p, n, node, dad, float(length),
                                        float(min_len), float(max_len), int(max_nr),
                                        ctypes.byref(out_len), ctypes.byref(out_lnl),
                                        ctypes.byref(nd)), "plk_optimize_branch")
        return out_len.value, out_lnl.value, nd.value

    def rell(self, weights):
        w = np.ascontiguousarray(weights, np.float32)
        out = np.zeros(w.shape[0])
        check(lib().plk_rell(self.ctx.handle, self.id, f32ptr(w), w.shape[0], dptr(out)),
              "plk_rell")
        return out

    def read_partial(self, slot, partial=True):
        """(partial [P][block] FP64 or None, scale counters) of one slot."""
        part = np.zeros((self.P, self.block)) if partial else None
        scale = np.zeros((self.P, self.ncat) if self.safe else (self.P,), np.int32)
        check(lib().plk_read_partial(self.ctx.handle, self.id, int(slot),
                                     dptr(part) if partial else None, i32ptr(scale)),
              "plk_read_partial")
        return part, scale


class BatchPlan:
    """A batch of (instance, ops, node, dad, length) jobs marshalled once for
    repeated plk_batch_evaluate calls (the items reference the op arrays,
    which the plan keeps alive). modes: None (all branch lnL) or per-job
    0 (branch lnL) / 1 (theta buffer)."""

    def __init__(self, ctx, jobs, modes=None):
        self.ctx = ctx
        self.n = len(jobs)
        self.items = (_capi.PlkBatchItem * self.n)()
        self._keep = []
        for k, (inst, ops, node, dad, length) in enumerate(jobs):
            arr = np.ascontiguousarray(ops, dtype=OP_DTYPE) if ops is not None and len(ops) else None
            self._keep.append(arr)
            it = self.items[k]
            it.instance = inst.id
            it.ops = arr.ctypes.data if arr is not None else None
            it.nops = 0 if arr is None else len(arr)
            it.node, it.dad, it.length = node, dad, float(length)
        self.modes = None if modes is None else np.ascontiguousarray(modes, np.int32)
        self.out = np.zeros(self.n)

    def run(self):
        m = i32ptr(self.modes) if self.modes is not None else None
        check(lib().plk_batch_evaluate(self.ctx.handle, self.items, m, self.n, dptr(self.out)),
              "plk_batch_evaluate")
        return self.out.copy()


def likelihood_branch_batch(ctx, jobs):
    """jobs: list of (instance, ops, node, dad, length); one fused launch
    (plk_likelihood_branch_batch semantics). Returns the lnL of every job."""
    return BatchPlan(ctx, jobs).run()


def batch_evaluate(ctx, jobs, modes):
    """plk_batch_evaluate: jobs as likelihood_branch_batch, modes[i] 0 for the
    branch lnL, 1 for the theta buffer of the branch (lnl 0)."""
    return BatchPlan(ctx, jobs, modes).run()


def likelihood_derv_batch(ctx, pairs):
    """pairs: list of (instance, length) on current theta buffers; one launch
    per shape (plk_likelihood_derv_batch). Returns arrays df, ddf, lnl."""
    n = len(pairs)
    ids = np.array([inst.id for inst, _ in pairs], np.int32)
    lens = np.array([float(x) for _, x in pairs], np.float64)
    df, ddf, lnl = np.zeros(n), np.zeros(n), np.zeros(n)
    check(lib().plk_likelihood_derv_batch(ctx.handle, i32ptr(ids), dptr(lens), n, dptr(df),
                                          dptr(ddf), dptr(lnl)), "plk_likelihood_derv_batch")
    return df, ddf, lnl


def minimize_one_dimen(f, xmin, xguess, xmax, tol):
    """Optimization::minimizeOneDimen in the host library: (x, f(x), evaluations)."""
    cb = _capi.BRENT_FN(lambda x, _arg: float(f(x)))
    x, fx, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
    check(lib().plkh_minimize_one_dimen(cb, None, xmin, xguess, xmax, tol, ctypes.byref(x), ctypes.byref(fx),
                                        ctypes.byref(n)),
          "plkh_minimize_one_dimen")
    return x.value, fx.value, n.value


def models_optimize_gamma_shape(trees, eps=0.0):
    """optimize_gamma_shape of every tree advanced together (one batched
    traversal per round): (shapes, lnL values, evaluations, statuses)."""
    n = len(trees)
    hs = (ctypes.c_void_p * max(n, 1))(*[t._h for t in trees])
    a, lnl = np.zeros(n), np.zeros(n)
    ne, st = np.zeros(n, np.int32), np.zeros(n, np.int32)
    check(lib().plkh_models_optimize_gamma_shape(hs, n, eps, dptr(a), dptr(lnl), i32ptr(ne), i32ptr(st)),
          "plkh_models_optimize_gamma_shape")
    return a, lnl, ne, st


def uninformative_patterns(nstates, nseq, max_patterns=1 << 24):
    """ASC_INFORMATIVE's unobserved patterns: array [count][nseq] of states."""
    n = ctypes.c_long()
    check(lib().plkh_uninformative_patterns(nstates, nseq, max_patterns, None, 0, ctypes.byref(n)),
          "plkh_uninformative_patterns")
    out = np.zeros((n.value, nseq), np.uint8)
    check(lib().plkh_uninformative_patterns(nstates, nseq, max_patterns, u8ptr(out), n.value, ctypes.byref(n)),
          "plkh_uninformative_patterns")
    return out


def gamma_rates(ncat, alpha, p_inv=0.0):
    """Discrete-Gamma category rates and weights (host library)."""
    rates, props = np.zeros(ncat), np.zeros(ncat)
    check(lib().plkh_gamma_rates(ncat, alpha, p_inv, dptr(rates), dptr(props)), "plkh_gamma_rates")
    return rates, props


def supertree_optimize_all_branches(trees, iterations=1, tolerance=0.001, max_nr=100):
    """PhyloSuperTree::optimizeAllBranches (tree/phylosupertree.cpp:773-792):
    every tree's optimizeAllBranches(iterations, tolerance / min(n, 10),
    max_nr) with the host Newton loops advanced together (one batched
    traversal and one batched derivative launch per round). Returns the
    per-tree lnL and derivative-evaluation counts."""
    n = len(trees)
    hs = (ctypes.c_void_p * max(n, 1))(*[t._h for t in trees])
    lnl = np.zeros(n)
    nd = np.zeros(n, np.int32)
    st = np.zeros(n, np.int32)
    check(lib().plkh_supertree_optimize_all_branches(hs, n, iterations, tolerance, max_nr, dptr(lnl),
                                                     i32ptr(nd), i32ptr(st)),
          "plkh_supertree_optimize_all_branches")
    return lnl, nd


class PhyloTree:
    """PhyloTree-shaped C++ driver (traversal planning + Newton branch optimisation)."""

    def __init__(self, ctx, workload, precision=PREC_F64, safe_numeric=False):
        st = ctypes.c_int(0)
        h = lib().plkh_tree_create(ctx.handle if ctx is not None else None, workload.handle,
                                   precision, int(bool(safe_numeric)), ctypes.byref(st))
        if not h:
            raise PlkError(st.value, "plkh_tree_create")
        self._h = h
        self.ctx = ctx
        self.workload = workload
        self.P = workload.npatterns
        if ctx is not None:
            ctx.adopt(self)

    def close(self):
        if getattr(self, "_h", None):
            lib().plkh_tree_destroy(self._h)
            self._h = None

    def __del__(self):
        if not shutting_down():
            self.close()

    @property
    def instance_id(self):
        return lib().plkh_tree_instance(self._h)

    def branches(self):
        nb = lib().plkh_tree_num_branches(self._h)
        a, b = np.zeros(nb, np.int32), np.zeros(nb, np.int32)
        check(lib().plkh_tree_branches(self._h, i32ptr(a), i32ptr(b)), "plkh_tree_branches")
        return list(zip(a.tolist(), b.tolist()))

    def get_length(self, a, b):
        x = ctypes.c_double()
        check(lib().plkh_tree_get_length(self._h, a, b, ctypes.byref(x)), "plkh_tree_get_length")
        return x.value

    def set_length(self, a, b, length):
        check(lib().plkh_tree_set_length(self._h, a, b, float(length)), "plkh_tree_set_length")

    def clear_all_partial(self):
        check(lib().plkh_tree_clear_all_partial(self._h), "plkh_tree_clear_all_partial")

    def reset(self):
        check(lib().plkh_tree_reset(self._h), "plkh_tree_reset")

    def plan(self, a, b):
        cap = 2 * self.workload.ntaxa
        ops = ops_array(cap)
        n = ctypes.c_int()
        node, dad = PlkEnd(), PlkEnd()
        length = ctypes.c_double()
        check(lib().plkh_tree_plan(self._h, a, b, ops.ctypes.data_as(ctypes.c_void_p), cap,
                                   ctypes.byref(n), ctypes.byref(node), ctypes.byref(dad),
                                   ctypes.byref(length)), "plkh_tree_plan")
        return ops[:n.value].copy(), node, dad, length.value

    def plan_full(self, a, b):
        """The full post-order op list for branch (a, b), no partial marked
        (parsimony): (ops, node, dad)."""
        cap = 2 * self.workload.ntaxa
        ops = ops_array(cap)
        n = ctypes.c_int()
        node, dad = PlkEnd(), PlkEnd()
        check(lib().plkh_tree_plan_full(self._h, a, b, ops.ctypes.data_as(ctypes.c_void_p), cap,
                                        ctypes.byref(n), ctypes.byref(node), ctypes.byref(dad)),
              "plkh_tree_plan_full")
        return ops[:n.value].copy(), node, dad

    def parsimony_branch(self, a, b):
        """computeParsimonyBranchFast at branch (a, b) on the device, only the
        stale slots recomputed: (score, changes on the branch)."""
        score, sub = ctypes.c_double(), ctypes.c_double()
        check(lib().plkh_tree_parsimony_branch(self._h, a, b, ctypes.byref(score), ctypes.byref(sub)),
              "plkh_tree_parsimony_branch")
        return score.value, sub.value

    def compute_likelihood(self, want_pattern_lh=False):
        lnl = ctypes.c_double()
        plh = np.zeros(self.P) if want_pattern_lh else None
        check(lib().plkh_tree_compute_likelihood(self._h, ctypes.byref(lnl),
                                                 dptr(plh) if plh is not None else None),
              "plkh_tree_compute_likelihood")
        return (lnl.value, plh) if want_pattern_lh else lnl.value

    def likelihood_branch(self, a, b):
        lnl = ctypes.c_double()
        check(lib().plkh_tree_likelihood_branch(self._h, a, b, ctypes.byref(lnl), None),
              "plkh_tree_likelihood_branch")
        return lnl.value

    def likelihood_derv(self, a, b):
        df, ddf = ctypes.c_double(), ctypes.c_double()
        check(lib().plkh_tree_likelihood_derv(self._h, a, b, ctypes.byref(df), ctypes.byref(ddf)),
              "plkh_tree_likelihood_derv")
        return df.value, ddf.value

    def optimize_branch(self, a, b, max_nr=100):
        x, nd = ctypes.c_double(), ctypes.c_int()
        check(lib().plkh_tree_optimize_branch(self._h, a, b, max_nr, ctypes.byref(x),
                                              ctypes.byref(nd)), "plkh_tree_optimize_branch")
        return x.value, nd.value

    def set_gamma_shape(self, alpha):
        """RateGamma::setGammaShape: new category rates, every partial stale."""
        check(lib().plkh_tree_set_gamma_shape(self._h, alpha), "plkh_tree_set_gamma_shape")

    def optimize_gamma_shape(self, eps=0.0):
        """RateGamma::optimizeParameters on the device: (shape, lnL, evaluations)."""
        a, lnl, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
        check(lib().plkh_tree_optimize_gamma_shape(self._h, eps, ctypes.byref(a), ctypes.byref(lnl),
                                                   ctypes.byref(n)),
              "plkh_tree_optimize_gamma_shape")
        return a.value, lnl.value, n.value

    def optimize_all_branches(self, iterations=1, tolerance=0.001, max_nr=100):
        lnl, nd = ctypes.c_double(), ctypes.c_int()
        check(lib().plkh_tree_optimize_all_branches(self._h, iterations, tolerance, max_nr,
                                                    ctypes.byref(lnl), ctypes.byref(nd)),
              "plkh_tree_optimize_all_branches")
        return lnl.value, nd.value

    def pass_counts(self):
        """Derivative evaluations per branch of the last optimize_all_branches
        pass, in branches() order."""
        n = ctypes.c_int()
        check(lib().plkh_tree_pass_counts(self._h, None, 0, ctypes.byref(n)), "plkh_tree_pass_counts")
        out = np.zeros(n.value, np.int32)
        check(lib().plkh_tree_pass_counts(self._h, i32ptr(out), n.value, ctypes.byref(n)),
              "plkh_tree_pass_counts")
        return out

    def neighbors(self, node):
        nb, deg = np.zeros(3, np.int32), ctypes.c_int()
        check(lib().plkh_tree_neighbors(self._h, int(node), i32ptr(nb), ctypes.byref(deg)),
              "plkh_tree_neighbors")
        return nb[:deg.value].tolist()

    def set_length_keep(self, a, b, length):
        check(lib().plkh_tree_set_length_keep(self._h, a, b, float(length)),
              "plkh_tree_set_length_keep")

    def nni_begin(self, node1, node2):
        check(lib().plkh_tree_nni_begin(self._h, node1, node2), "plkh_tree_nni_begin")

    def nni_swap(self, node1, x, node2, y):
        check(lib().plkh_tree_nni_swap(self._h, node1, x, node2, y), "plkh_tree_nni_swap")

    def clear_partial(self, node, frm):
        check(lib().plkh_tree_clear_partial(self._h, node, frm), "plkh_tree_clear_partial")

    def nni_end(self):
        check(lib().plkh_tree_nni_end(self._h), "plkh_tree_nni_end")

    def inner_branches(self):
        n = ctypes.c_int()
        check(lib().plkh_tree_inner_branches(self._h, None, None, 0, ctypes.byref(n)),
              "plkh_tree_inner_branches")
        a, b = np.zeros(n.value, np.int32), np.zeros(n.value, np.int32)
        check(lib().plkh_tree_inner_branches(self._h, i32ptr(a), i32ptr(b), n.value,
                                             ctypes.byref(n)), "plkh_tree_inner_branches")
        return list(zip(a.tolist(), b.tolist()))

    def nni_evaluate(self, node1, node2, max_nr=10):
        """getBestNNIForBran's two moves: (lnl[2], lens[2][5], n_derv[2][5])."""
        lnl, lens, nd = np.zeros(2), np.zeros(10), np.zeros(10, np.int32)
        check(lib().plkh_tree_nni_evaluate(self._h, node1, node2, max_nr, dptr(lnl), dptr(lens),
                                           i32ptr(nd)), "plkh_tree_nni_evaluate")
        return lnl, lens.reshape(2, 5), nd.reshape(2, 5)

    def nni_evaluate_all(self, max_nr=10):
        """Every inner branch's NNI evaluation in one device sweep:
        (lnl[m][2], lens[m][2][5], n_derv[m][2][5])."""
        m = len(self.inner_branches())
        lnl, lens, nd = np.zeros(2 * m), np.zeros(10 * m), np.zeros(10 * m, np.int32)
        check(lib().plkh_tree_nni_evaluate_all(self._h, max_nr, dptr(lnl), dptr(lens), i32ptr(nd),
                                               m), "plkh_tree_nni_evaluate_all")
        return lnl.reshape(m, 2), lens.reshape(m, 2, 5), nd.reshape(m, 2, 5)

    def set_device_newton(self, on):
        check(lib().plkh_tree_set_device_newton(self._h, int(bool(on))),
              "plkh_tree_set_device_newton")

    def set_branch_bounds(self, min_len, max_len):
        check(lib().plkh_tree_set_branch_bounds(self._h, min_len, max_len),
              "plkh_tree_set_branch_bounds")
