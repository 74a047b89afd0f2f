"""The state-basis kernels (non-reversible and mixture models) on the device
against the oracle (itself pinned by brute force in
tests/test_oracle_nonrev.py): partials and scale counters, tree and pattern
lnL, derivatives, Newton, unrooted and rooted trees (the root leaf, children
toward the root), DNA and protein, category counts 1-8 (mixtures fused and
not), safe scaling, the FP32 form, and the batched calls. Tolerances as the
other parity tests: lnL 1e-9 relative, pattern lnL 1e-9 absolute, partials
1e-10 of the pattern maximum, derivatives 1e-7 relative, counters exact."""
import numpy as np
import pytest

import iqtree2_amd as ia
from oracle.binding import OracleTree
from refmath import indicators
from statebasis import adjacency, categories, nonrev_model, parents, plan
from test_oracle_nonrev import internal_branches, oracle

pytestmark = pytest.mark.gpu


def device(ctx, w, nm, states=None, rooted=False, safe=False, precision=ia.PREC_F64):
    st = w.get("states") if states is None else states
    C = len(categories(nm))
    inst = ia.Instance(ctx, w.ntaxa, w.nstates, C, st.shape[1], w.ntaxa, w.state_unknown, precision, safe)
    inst.set_patterns(st, w.get("freq"), np.zeros(st.shape[1]))
    inst.set_model_nonrev(nm, indicators(w.seq_type, w.nstates, w.state_unknown), rooted)
    return inst


def rooted_workload(seq_type, ntaxa, nsite, seed, mean_len=0.15):
    """Taxon 0 is the root leaf (zero-length branch, state 0 everywhere)."""
    w0 = ia.Workload(seq_type=seq_type, ntaxa=ntaxa, nsite=nsite, ncat=4, mean_len=mean_len, seed=seed)
    u, v, ln = (np.array(x) for x in w0.edges())
    ln = ln.copy()
    ln[(u == 0) | (v == 0)] = 0.0
    w = ia.Workload(seq_type=seq_type, ntaxa=ntaxa, nsite=nsite, ncat=4, mean_len=mean_len, seed=seed,
                    edges=(u, v, ln))
    st = w.get("states").copy()
    st[0] = 0
    root = int(v[u == 0][0]) if np.any(u == 0) else int(u[v == 0][0])
    par = parents(adjacency(w.edges()), root)
    branches = [(root, 0)] + [(p, x) for x, p in par.items() if p is not None and x != 0]
    return w, st, root, branches


def compare(inst, orc, w, branches, plan_kw, check_partials=True):
    for a, b in branches:
        ops, node, dad, length = plan(w.ntaxa, w.edges(), a, b, **plan_kw)
        lg, pg = inst.likelihood_branch(ops, node, dad, length, want_pattern_lh=True)
        lo, po = orc.likelihood_branch(ops, node, dad, length, want_pattern_lh=True)
        assert abs(lg - lo) <= 1e-9 * abs(lo), (a, b, lg, lo)
        np.testing.assert_allclose(pg, po, rtol=0, atol=1e-9)
        if check_partials:
            pd, sd = inst.read_partial(node.id)
            po_, so = orc.read_partial(node.id)
            assert np.array_equal(sd, so), (a, b)
            scale = np.maximum(np.abs(po_).max(axis=1, keepdims=True), 1e-300)
            assert np.max(np.abs(pd - po_) / scale) <= 1e-10, (a, b)
        inst.likelihood_buffer(ops, node, dad)
        orc.likelihood_buffer(ops, node, dad)
        for t in (length, 0.4 * length + 1e-4, 2.0 * length + 0.02):
            dfg, ddfg, lng = inst.likelihood_derv(t)
            dfo, ddfo = orc.likelihood_derv(t)
            assert abs(dfg - dfo) <= 1e-7 * max(1.0, abs(dfo)), (a, b, t, dfg, dfo)
            assert abs(ddfg - ddfo) <= 1e-7 * max(1.0, abs(ddfo)), (a, b, t, ddfg, ddfo)
            lbo = orc.likelihood_from_buffer(t)
            assert abs(lng - lbo) <= 1e-9 * abs(lbo)
            assert abs(inst.likelihood_from_buffer(t) - lbo) <= 1e-9 * abs(lbo)


UNROOTED = [(ia.DNA, 40, 3000, 1, 1, False, False), (ia.DNA, 30, 2000, 2, 1, False, False),
            (ia.DNA, 30, 2000, 2, 2, False, False), (ia.DNA, 30, 2000, 4, 4, True, False),
            (ia.DNA, 24, 1500, 2, 4, False, False), (ia.DNA, 64, 1500, 2, 2, False, True),
            (ia.PROTEIN, 16, 800, 2, 2, False, False), (ia.PROTEIN, 12, 600, 1, 4, False, True)]


@pytest.mark.parametrize("seq_type,ntaxa,nsite,nmix,nrates,fused,safe", UNROOTED)
def test_unrooted_vs_oracle(ctx, seq_type, ntaxa, nsite, nmix, nrates, fused, safe):
    w = ia.Workload(seq_type=seq_type, ntaxa=ntaxa, nsite=nsite, ncat=4,
                    mean_len=0.6 if safe else 0.15, seed=700 + ntaxa + nmix)
    nm = nonrev_model(w.nstates, nmix, nrates, fused, seed=ntaxa + nrates)
    inst = device(ctx, w, nm, safe=safe)
    orc = oracle(w, nm, safe=safe)
    try:
        compare(inst, orc, w, internal_branches(w)[::5], {})
        # Newton: the same length and step count
        for u, v in internal_branches(w)[1::11]:
            ops, node, dad, length = plan(w.ntaxa, w.edges(), u, v)
            xg, lg, ng = inst.optimize_branch(ops, node, dad, length)
            xo, lo, no, so = orc.optimize_branch(ops, node, dad, length)
            assert so == 0 and ng == no, (u, v, ng, no)
            assert abs(xg - xo) <= 1e-7 * max(1.0, xo)
            assert abs(lg - lo) <= 1e-9 * abs(lo)
    finally:
        inst.close()


@pytest.mark.parametrize("seq_type,ntaxa,nsite,nmix,nrates", [(ia.DNA, 32, 2000, 1, 4), (ia.DNA, 20, 1500, 2, 2),
                                                            (ia.PROTEIN, 10, 500, 1, 4)])
def test_rooted_vs_oracle(ctx, seq_type, ntaxa, nsite, nmix, nrates):
    w, st, root, branches = rooted_workload(seq_type, ntaxa, nsite, seed=31 + ntaxa)
    nm = nonrev_model(w.nstates, nmix, nrates, False, seed=ntaxa)
    inst = device(ctx, w, nm, states=st, rooted=True)
    orc = oracle(w, nm, states=st, rooted=True)
    try:
        compare(inst, orc, w, branches[::3], dict(root=root, root_leaf=0))
        # every branch of a rooted tree gives the same lnL
        vals = [inst.likelihood_branch(*plan(w.ntaxa, w.edges(), a, b, root=root, root_leaf=0))
                for a, b in branches]
        assert np.ptp(vals) <= 1e-9 * abs(vals[0])
    finally:
        inst.close()


def test_reversible_model_through_state_basis(ctx):
    """kernel_nonrev on the workload's reversible model: the state-basis path
    gives the eigen-basis path's lnL."""
    w = ia.Workload(seq_type=ia.DNA, ntaxa=48, nsite=4000, ncat=4, mean_len=0.12, seed=88)
    m = w.model_arrays()
    Q, pi = w.get("Q"), w.get("pi")
    nm = dict(Q=Q[None], pi=pi[None], ceval=m["eval"][None].astype(complex),
              cevec=m["evec"][None].astype(complex), cinv_evec=m["inv_evec"][None].astype(complex),
              mixw=np.ones(1), rates=m["rates"], props=m["props"], fused=False)
    sb = device(ctx, w, nm)
    rev = ia.Instance.from_workload(ctx, w)
    planner = ia.PhyloTree(None, w)
    try:
        for a, b in internal_branches(w)[::7]:
            ops, node, dad, length = plan(w.ntaxa, w.edges(), a, b)
            l1 = sb.likelihood_branch(ops, node, dad, length)
            l2 = rev.likelihood_branch(ops, node, dad, length)
            assert abs(l1 - l2) <= 1e-10 * abs(l2), (a, b, l1, l2)
    finally:
        sb.close()
        rev.close()
        planner.close()


def test_fp32_and_batches(ctx):
    w = ia.Workload(seq_type=ia.DNA, ntaxa=40, nsite=3000, ncat=4, mean_len=0.12, seed=5)
    nm = nonrev_model(4, 2, 2, False, seed=6)
    inst = device(ctx, w, nm)
    inst32 = device(ctx, w, nm, precision=ia.PREC_F32)
    plain = ia.Instance.from_workload(ctx, w)
    try:
        u, v = internal_branches(w)[3]
        ops, node, dad, length = plan(w.ntaxa, w.edges(), u, v)
        lnl = inst.likelihood_branch(ops, node, dad, length)
        l32 = inst32.likelihood_branch(ops, node, dad, length)
        assert abs(l32 - lnl) <= 2e-6 * abs(lnl)
        # a batch with a state-basis item: that item is evaluated alone, bit for bit
        out = ia.likelihood_branch_batch(ctx, [(inst, ops, node, dad, length), (plain, ops, node, dad, length)])
        assert out[0] == lnl
        lp = plain.likelihood_branch(ops, node, dad, length)
        assert abs(out[1] - lp) <= 1e-12 * abs(lp)
        # the batched derivative call refuses it
        inst.likelihood_buffer(ops, node, dad)
        with pytest.raises(ia.PlkError):
            ia.likelihood_derv_batch(ctx, [(inst, length)])
        # RELL over the state-basis pattern lnL
        _, plh = inst.likelihood_branch(ops, node, dad, length, want_pattern_lh=True)
        wts = np.random.default_rng(1).integers(0, 3, (8, w.npatterns)).astype(np.float32)
        np.testing.assert_allclose(inst.rell(wts), wts.astype(np.float64) @ plh, rtol=1e-12)
        # back to the reversible kernels
        m = w.model_arrays()
        C = len(categories(nm))
        assert C == 4
        inst.set_model(m["eval"], m["evec"], m["inv_evec"], m["tip"], m["rates"], m["props"])
        planner = ia.PhyloTree(None, w)
        try:
            o2, n2, d2, l2 = planner.plan(*planner.branches()[0])
            assert abs(inst.likelihood_branch(o2, n2, d2, l2) - plain.likelihood_branch(o2, n2, d2, l2)) \
                <= 1e-12 * abs(lp)
        finally:
            planner.close()
    finally:
        for x in (inst, inst32, plain):
            x.close()


@pytest.mark.parametrize("safe", [False, True])
def test_asc_vs_oracle(ctx, safe):
    """+ASC with a non-reversible mixture: the constant columns unobserved."""
    from test_oracle_asc import asc_patterns
    w = ia.Workload(seq_type=ia.DNA, ntaxa=30, nsite=2500, ncat=4, mean_len=0.4 if safe else 0.15,
                    seed=77)
    st, fq, inv, nasc = asc_patterns(w)
    nm = nonrev_model(4, 2, 2, False, seed=5)
    C = len(categories(nm))
    ind = indicators(ia.DNA, 4, w.state_unknown)
    inst = ia.Instance(ctx, w.ntaxa, 4, C, st.shape[1], w.ntaxa, w.state_unknown, ia.PREC_F64, safe)
    inst.set_patterns(st, fq, inv)
    inst.set_model_nonrev(nm, ind)
    inst.set_asc(nasc)
    orc = OracleTree(w.ntaxa, 4, C, st.shape[1], w.ntaxa, w.state_unknown, safe, 2)
    orc.set_patterns(st, fq, inv)
    orc.set_model_nonrev(nm, ind)
    orc.set_asc(nasc)
    pobs = st.shape[1] - nasc
    try:
        for u, v in internal_branches(w)[::6]:
            ops, node, dad, length = plan(w.ntaxa, w.edges(), u, v)
            lg, pg = inst.likelihood_branch(ops, node, dad, length, want_pattern_lh=True)
            lo, po = orc.likelihood_branch(ops, node, dad, length, want_pattern_lh=True)
            assert abs(lg - lo) <= 1e-9 * abs(lo), (u, v, lg, lo)
            np.testing.assert_allclose(pg[:pobs], po, rtol=0, atol=1e-9)
            inst.likelihood_buffer(ops, node, dad)
            orc.likelihood_buffer(ops, node, dad)
            for t in (length, 0.5 * length + 1e-3):
                dfg, ddfg, lng = inst.likelihood_derv(t)
                dfo, ddfo = orc.likelihood_derv(t)
                assert abs(dfg - dfo) <= 1e-7 * max(1.0, abs(dfo))
                assert abs(ddfg - ddfo) <= 1e-7 * max(1.0, abs(ddfo))
                assert abs(lng - orc.likelihood_from_buffer(t)) <= 1e-9 * abs(lg)
        u, v = internal_branches(w)[2]
        ops, node, dad, length = plan(w.ntaxa, w.edges(), u, v)
        xg, lg, ng = inst.optimize_branch(ops, node, dad, length)
        xo, lo, no, so = orc.optimize_branch(ops, node, dad, length)
        assert so == 0 and ng == no
        assert abs(xg - xo) <= 1e-7 * max(1.0, xo)
    finally:
        inst.close()


@pytest.mark.parametrize("seq_type,nmix,nrates,safe", [(ia.DNA, 3, 2, False), (ia.DNA, 5, 1, True),
                                                      (ia.PROTEIN, 3, 1, False)])
def test_category_count_padded(ctx, seq_type, nmix, nrates, safe):
    """A mixture whose ncat_mix (6, 5, 3) is not a built category count runs
    in an instance of the next one up, the extra categories weight-0 copies:
    the oracle at the model's own count gives the same lnL and derivatives."""
    w = ia.Workload(seq_type=seq_type, ntaxa=24, nsite=1500, ncat=4, mean_len=0.3 if safe else 0.15,
                    seed=55 + nmix)
    nm = nonrev_model(w.nstates, nmix, nrates, False, seed=nmix + 17)
    Cm = len(categories(nm))
    Cpad = 1 << (Cm - 1).bit_length()
    ind = indicators(seq_type, w.nstates, w.state_unknown)
    inst = ia.Instance(ctx, w.ntaxa, w.nstates, Cpad, w.npatterns, w.ntaxa, w.state_unknown, ia.PREC_F64, safe)
    inst.set_patterns(w.get("states"), w.get("freq"), np.zeros(w.npatterns))
    inst.set_model_nonrev(nm, ind)
    orc = oracle(w, nm, safe=safe)
    try:
        for u, v in internal_branches(w)[::5]:
            ops, node, dad, length = plan(w.ntaxa, w.edges(), u, v)
            lg, pg = inst.likelihood_branch(ops, node, dad, length, want_pattern_lh=True)
            lo, po = orc.likelihood_branch(ops, node, dad, length, want_pattern_lh=True)
            assert abs(lg - lo) <= 1e-9 * abs(lo), (u, v, lg, lo)
            np.testing.assert_allclose(pg, po, rtol=0, atol=1e-9)
            inst.likelihood_buffer(ops, node, dad)
            orc.likelihood_buffer(ops, node, dad)
            dfg, ddfg, _ = inst.likelihood_derv(length)
            dfo, ddfo = orc.likelihood_derv(length)
            assert abs(dfg - dfo) <= 1e-7 * max(1.0, abs(dfo))
            assert abs(ddfg - ddfo) <= 1e-7 * max(1.0, abs(ddfo))
    finally:
        inst.close()
