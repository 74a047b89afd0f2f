"""GPU engine against the oracle where the behaviour is a rule (the oracle's
side is pinned in test_oracle_edge_cases.py):

- optimizeOneBranch's divergence guard (tree/phylotree.cpp:2616-2625) in
  the single-branch entry point, the device-resident Newton sweep and the
  request server, with a bracket whose upper end lies below many branches'
  optima, so that the guard both keeps and restores lengths;
- patterns whose likelihood is exactly zero (a state made impossible by a
  zero tip vector): the per-pattern fix-up of the branch lnL, derivatives
  reported as zero, Newton stopping in the normal kernel and continuing in
  the safe one (tree/phylokernelnew.h:2577-2579, 2629-2632, 3115-3132),
  through every entry point that evaluates them."""
import ctypes

import numpy as np
import pytest

import iqtree2_amd as ia
from iqtree2_amd import _capi
from iqtree2_amd.serving import Request, Server
from oracle.binding import OracleTree

pytestmark = pytest.mark.gpu

LNL_RTOL = 1e-9
LOG_SCALING_THRESHOLD = np.log(2.0 ** -256)


def guard_workload():
    return ia.Workload(seq_type=ia.DNA, ntaxa=24, nsite=3000, ncat=4, mean_len=0.12, alpha=0.8,
                       seed=45)


def oracle_pass(w, orc, branches, min_len, max_len, safe=False):
    """optimizeAllBranches' pass replayed on the oracle, branch by branch."""
    planner = ia.PhyloTree(None, w)
    a0, b0 = branches[0]
    ops, node, dad, length = planner.plan(b0, a0)
    orc.likelihood_branch(ops, node, dad, length)
    lens, counts = [], []
    for a, b in branches:
        ops, node, dad, length = planner.plan(b, a)
        x, _, nd, st = orc.optimize_branch(ops, node, dad, length, min_len, max_len)
        assert st == 0
        planner.set_length(a, b, x)
        lens.append(x)
        counts.append(nd)
    return np.array(lens), np.array(counts, np.int32), orc.likelihood_from_buffer(lens[-1])


def test_divergence_guard_single_branch(ctx):
    w = guard_workload()
    planner = ia.PhyloTree(None, w)
    inst = ia.Instance.from_workload(ctx, w)
    orc = OracleTree.from_workload(w)
    kept = restored = 0
    try:
        for a, b in planner.branches()[::3]:
            ops, node, dad, length = planner.plan(b, a)
            max_len = 0.6 * length
            # Newton from the length itself (above the bracket) and from
            # inside the bracket; the partials do not depend on the start
            for start in (length, 0.25 * length):
                xg, lg, ng = inst.optimize_branch(ops, node, dad, start, 1e-6, max_len)
                xo, lo, no, st = orc.optimize_branch(ops, node, dad, start, 1e-6, max_len)
                assert st == 0 and ng == no, (a, b, start, ng, no)
                assert abs(xg - xo) <= 1e-7 * max(1.0, xo), (a, b, start, xg, xo)
                assert abs(lg - lo) <= LNL_RTOL * abs(lo)
                if xo == start and start > max_len:
                    restored += 1
                elif xo > 0.95 * max_len:
                    kept += 1
    finally:
        inst.close()
    assert restored >= 2 and kept >= 1, (restored, kept)


def test_divergence_guard_device_sweep(ctx):
    w = guard_workload()
    max_len = 0.08
    tree = ia.PhyloTree(ctx, w)
    try:
        tree.set_branch_bounds(1e-6, max_len)
        lnl_dev, _ = tree.optimize_all_branches(iterations=1, tolerance=1e-3, max_nr=100)
        counts_dev = tree.pass_counts()
        branches = tree.branches()
        lens_dev = np.array([tree.get_length(a, b) for a, b in branches])
    finally:
        tree.close()
    orc = OracleTree.from_workload(w)
    lens_orc, counts_orc, lnl_orc = oracle_pass(w, orc, branches, 1e-6, max_len)
    init = dict(zip(zip(*w.edges()[:2]), w.edges()[2]))
    start = np.array([init.get((a, b), init.get((b, a))) for a, b in branches])
    assert np.sum((lens_orc == start) & (start > max_len)) >= 2      # restored by the guard
    np.testing.assert_array_equal(counts_dev, counts_orc)
    np.testing.assert_allclose(lens_dev, lens_orc, rtol=1e-7, atol=1e-12)
    assert abs(lnl_dev - lnl_orc) <= LNL_RTOL * abs(lnl_orc)


def test_divergence_guard_serving(ctx):
    w = guard_workload()
    planner = ia.PhyloTree(None, w)
    nb = len(planner.branches())
    reqs = [Request(w, "newton", branch=k, max_len=0.08) for k in range(0, nb, 2)]
    with Server(ctx) as srv:
        res = [srv.wait(t) for t in [srv.submit(r) for r in reqs]]
    inst = ia.Instance.from_workload(ctx, w)
    try:
        for r, got in zip(reqs, res):
            assert got["status"] == 0
            x, lnl, nd = inst.optimize_branch(r.ops, r.node, r.dad, r.length, 1e-6, 0.08)
            assert (got["length"], got["n_derv"]) == (x, nd)
            assert got["lnl"] == pytest.approx(lnl, rel=1e-12)
    finally:
        inst.close()


def zero_lh_setup(ctx, safe):
    w = ia.Workload(seq_type=ia.DNA, ntaxa=12, nsite=1500, ncat=4, mean_len=0.15, alpha=0.6,
                    seed=53)
    m = w.model_arrays()
    tip = m["tip"].copy()
    tip[2] = 0.0          # state G impossible: patterns showing G have likelihood 0
    model = dict(m, tip=tip)
    inst = ia.Instance(ctx, w.ntaxa, 4, 4, w.npatterns, w.ntaxa - 2, 18, safe_numeric=safe)
    inst.set_patterns(w.get("states"), w.get("freq"), np.zeros(w.npatterns))
    inst.set_model(*[model[k] for k in ("eval", "evec", "inv_evec", "tip", "rates", "props")])
    orc = OracleTree.from_arrays(w.ntaxa, 4, 4, 18, w.get("states"), w.get("freq"),
                                 np.zeros(w.npatterns), model, safe_numeric=safe)
    return w, model, inst, orc


@pytest.mark.parametrize("safe", [False, True])
def test_zero_likelihood_patterns(ctx, safe):
    w, model, inst, orc = zero_lh_setup(ctx, safe)
    planner = ia.PhyloTree(None, w)
    try:
        a, b = planner.branches()[4]
        ops, node, dad, length = planner.plan(a, b)
        with pytest.raises(ia.PlkError) as ei:
            inst.likelihood_branch(ops, node, dad, length)
        assert ei.value.status == 4
        # the fixed-up values are returned with the warning status
        lnl = ctypes.c_double()
        plh = np.zeros(w.npatterns)
        arr = np.ascontiguousarray(ops, dtype=ia.engine.OP_DTYPE)
        st = ia.lib().plk_likelihood_branch(ctx.handle, inst.id, arr.ctypes.data, len(arr), node,
                                            dad, length, ctypes.byref(lnl), _capi.dptr(plh))
        assert st == 4
        lo, po = orc.likelihood_branch(ops, node, dad, length, want_pattern_lh=True)
        zero = po == 4 * LOG_SCALING_THRESHOLD
        assert zero.sum() > 10
        np.testing.assert_array_equal(plh[zero], po[zero])
        np.testing.assert_allclose(plh, po, rtol=0, atol=1e-9)
        assert abs(lnl.value - lo) <= LNL_RTOL * abs(lo)
        # scale counters of the zero patterns grow at every step, bit-exact
        for op in ops[-3:]:
            slot = int(op["dad_slot"])
            np.testing.assert_array_equal(inst.read_partial(slot, partial=False)[1],
                                          orc.read_partial(slot, partial=False)[1])
        inst.likelihood_buffer(None, node, dad)
        orc.likelihood_buffer(None, node, dad)
        with pytest.raises(ia.PlkError) as ei:
            inst.likelihood_derv(length)
        assert ei.value.status == 4
        assert orc.likelihood_derv(length) == (0.0, 0.0)
        xo, _, no, so = orc.optimize_branch(ops, node, dad, length)
        if safe:
            x, _, nd = inst.optimize_branch(ops, node, dad, length)
            assert so == 0 and (x, nd) == (xo, no) == (length, 1)
        else:
            assert so == 4
            with pytest.raises(ia.PlkError) as ei:
                inst.optimize_branch(ops, node, dad, length)
            assert ei.value.status == 4
    finally:
        inst.close()


@pytest.mark.parametrize("safe", [False, True])
def test_zero_likelihood_sweep_and_server(ctx, safe):
    w, model, inst, orc = zero_lh_setup(ctx, safe)
    inst.close()
    tree = ia.PhyloTree(ctx, w, safe_numeric=safe)
    try:
        # the tree's instance gets the impossible-state model and no +I
        L = ia.lib()
        tid = tree.instance_id
        arrs = [np.ascontiguousarray(model[k]) for k in ("eval", "evec", "inv_evec", "tip", "rates",
                                                         "props")]
        _capi.check(L.plk_set_model(ctx.handle, tid, *[_capi.dptr(x) for x in arrs]), "set_model")
        st_ = np.ascontiguousarray(w.get("states"), np.uint8)
        fq = np.ascontiguousarray(w.get("freq"))
        zr = np.zeros(w.npatterns)
        _capi.check(L.plk_set_patterns(ctx.handle, tid, _capi.u8ptr(st_), _capi.dptr(fq),
                                       _capi.dptr(zr)), "set_patterns")
        tree.clear_all_partial()
        branches = tree.branches()
        if safe:
            # every derivative is 0/0: each branch keeps its (bracketed) length
            lnl_dev, nd = tree.optimize_all_branches(iterations=1, tolerance=1e-3, max_nr=100)
            lens_orc, counts_orc, lnl_orc = oracle_pass(w, orc, branches, 1e-6, 10.0)
            np.testing.assert_array_equal(tree.pass_counts(), counts_orc)
            assert np.all(counts_orc == 1)
            np.testing.assert_array_equal([tree.get_length(a, b) for a, b in branches], lens_orc)
            assert abs(lnl_dev - lnl_orc) <= LNL_RTOL * abs(lnl_orc)
        else:
            with pytest.raises(ia.PlkError) as ei:
                tree.optimize_all_branches(iterations=1, tolerance=1e-3, max_nr=100)
            assert ei.value.status == 4
        # partials left by a stopped sweep are never reused: the next
        # evaluation recomputes them (warning status, fixed-up value)
        with pytest.raises(ia.PlkError) as ei:
            tree.compute_likelihood()
        assert ei.value.status == 4
    finally:
        tree.close()
    # the request server: the safe request completes at its starting length,
    # the normal one fails alone
    req = Request(w, "newton", branch=5, safe_numeric=safe)
    req.invar = np.zeros(w.npatterns)
    req.model[3] = np.ascontiguousarray(model["tip"])
    other = Request(w, "newton", branch=5)
    with Server(ctx) as srv:
        r1, r2 = [srv.wait(t) for t in [srv.submit(req), srv.submit(other)]]
    assert r2["status"] == 0
    if safe:
        assert r1["status"] == 0 and r1["length"] == req.length and r1["n_derv"] == 1
    else:
        assert r1["status"] == 4
