"""The continuous-batching server (include/plk_serve.h) on the GPU: every
request's result equals the direct single-request call, is the same alone
and in a mix, admission respects the memory budget, and a failing request
does not disturb the others."""
import numpy as np
import pytest

import iqtree2_amd as ia
from iqtree2_amd.serving import Request, Server
from oracle.binding import OracleTree

pytestmark = pytest.mark.gpu


def request_mix(n, seed=11, max_taxa=48, max_sites=3000):
    rng = np.random.default_rng(seed)
    reqs = []
    for k in range(n):
        seq = [ia.DNA, ia.DNA, ia.PROTEIN, ia.CODON][k % 4]
        ntaxa = int(rng.integers(5, max_taxa))
        nsite = int(rng.integers(200, max_sites if seq != ia.CODON else max_sites // 4))
        w = ia.Workload(seq_type=seq, ntaxa=ntaxa, nsite=nsite, ncat=4, mean_len=0.1, alpha=0.6,
                        p_inv=0.1 if seq == ia.DNA else 0.0, seed=900 + k)
        kind = ["lnl", "newton", "rell"][k % 3]
        weights = None
        if kind == "rell":
            f = w.get("freq")
            weights = np.stack([rng.multinomial(int(f.sum()), f / f.sum()) for _ in range(3)])
        reqs.append((w, Request(w, kind, branch=k, weights=weights,
                                length_scale=3.0 if kind == "newton" else 1.0)))
    return reqs


def direct(ctx, w, req):
    """The same evaluation through the single-instance entry points."""
    safe = w.nstates not in (4, 20)
    inst = ia.Instance.from_workload(ctx, w, safe_numeric=safe)
    try:
        if req.kind == ia.serving.PLKS_NEWTON:
            x, lnl, nd = inst.optimize_branch(req.ops, req.node, req.dad, req.length)
            return {"lnl": lnl, "length": x, "n_derv": nd}
        lnl = inst.likelihood_branch(req.ops, req.node, req.dad, req.length)
        out = {"lnl": lnl}
        if req.kind == ia.serving.PLKS_RELL:
            out["rell"] = inst.rell(req.weights)
        return out
    finally:
        inst.close()


def run_all(ctx, reqs, **kw):
    with Server(ctx, **kw) as srv:
        tickets = [srv.submit(r) for _, r in reqs]
        res = [srv.wait(t) for t in tickets]
        stats = srv.stats()
    return res, stats


def test_results_equal_direct_calls(ctx):
    reqs = request_mix(18)
    res, stats = run_all(ctx, reqs)
    assert stats["completed"] == len(reqs)
    assert stats["peak_inflight"] > 1, "requests were meant to be resident together"
    for (w, r), got in zip(reqs, res):
        assert got["status"] == 0, got
        ref = direct(ctx, w, r)
        if r.kind == ia.serving.PLKS_NEWTON:
            # same theta, bit-identical batched derivatives: the same Newton path
            assert got["length"] == ref["length"] and got["n_derv"] == ref["n_derv"]
            assert got["lnl"] == ref["lnl"]
        else:
            # fused traversal vs the lone traversal kernel: rounding only
            assert abs(got["lnl"] - ref["lnl"]) <= 1e-12 * abs(ref["lnl"])
        if r.kind == ia.serving.PLKS_RELL:
            np.testing.assert_allclose(got["rell"], ref["rell"], rtol=1e-11)
    # and against the oracle for the plain lnL requests
    for (w, r), got in list(zip(reqs, res))[:6]:
        if r.kind != ia.serving.PLKS_LNL:
            continue
        orc = OracleTree.from_workload(w, safe_numeric=w.nstates not in (4, 20))
        lo = orc.likelihood_branch(r.ops, r.node, r.dad, r.length)
        assert abs(got["lnl"] - lo) <= 1e-9 * abs(lo)


def test_alone_and_in_mix_identical(ctx):
    reqs = request_mix(12, seed=5)
    mixed, _ = run_all(ctx, reqs)
    with Server(ctx) as srv:
        alone = [srv.wait(srv.submit(r)) for _, r in reqs]
    for a, m in zip(alone, mixed):
        assert a["status"] == m["status"] == 0
        assert a["lnl"] == m["lnl"]
        assert a["length"] == m["length"] and a["n_derv"] == m["n_derv"]
        if a["rell"] is not None:
            np.testing.assert_array_equal(a["rell"], m["rell"])


def test_admission_budget_and_failure_isolation(ctx):
    reqs = request_mix(10, seed=8, max_taxa=30, max_sites=2000)
    budget = int(2.5 * max(r.footprint() for _, r in reqs))
    # one request too large for the budget, one with an invalid op list
    big_w = ia.Workload(seq_type=ia.PROTEIN, ntaxa=60, nsite=20000, seed=3)
    big = Request(big_w, "lnl")
    assert big.footprint() > budget
    bad_w, bad = reqs[1][0], Request(reqs[1][0], "lnl", branch=2)
    bad.ops = bad.ops.copy()
    bad.ops[0]["dad_slot"] = 10_000
    with Server(ctx, hbm_budget=budget) as srv:
        t_ok = [srv.submit(r) for _, r in reqs[:5]]
        t_big = srv.submit(big)
        t_bad = srv.submit(bad)
        t_ok += [srv.submit(r) for _, r in reqs[5:]]
        res_ok = [srv.wait(t) for t in t_ok]
        res_big, res_bad = srv.wait(t_big), srv.wait(t_bad)
        stats = srv.stats()
    assert res_big["status"] == 2            # PLK_ERR_NOMEM
    assert res_bad["status"] == 1            # PLK_ERR_ARG
    assert stats["peak_bytes"] <= budget
    alone, _ = run_all(ctx, reqs)
    for got, ref in zip(res_ok, alone):
        assert got["status"] == 0
        assert got["lnl"] == ref["lnl"] and got["length"] == ref["length"]
