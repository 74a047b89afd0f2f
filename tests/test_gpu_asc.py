"""Lewis's ascertainment bias correction (+ASC) on the device against the
oracle (itself pinned by tests/test_oracle_asc.py): tree lnL, pattern lnL,
branch derivatives, lnL from the buffer, Newton, RELL, the FP32 form, and the
batched calls (an +ASC item evaluated alone inside a batch; the batched
derivative call refusing it). Tolerances as the other parity tests: lnL 1e-9
relative, pattern lnL 1e-9 absolute, derivatives 1e-7 relative."""
import numpy as np
import pytest

import iqtree2_amd as ia
from oracle.binding import OracleTree
from test_oracle_asc import asc_oracle, asc_patterns

pytestmark = pytest.mark.gpu


def asc_instance(ctx, w, st, fq, inv, nasc, precision=ia.PREC_F64, safe=False):
    inst = ia.Instance(ctx, w.ntaxa, w.nstates, w.ncat, st.shape[1], w.ntaxa, w.state_unknown,
                       precision, safe)
    inst.set_patterns(st, fq, inv)
    m = w.model_arrays()
    inst.set_model(m["eval"], m["evec"], m["inv_evec"], m["tip"], m["rates"], m["props"])
    inst.set_asc(nasc)
    return inst


@pytest.mark.parametrize("seq_type,ntaxa,nsite,safe", [(ia.DNA, 40, 3000, False),
                                                       (ia.DNA, 24, 2000, True),
                                                       (ia.PROTEIN, 20, 1200, False),
                                                       (ia.CODON, 10, 150, True)])
def test_asc_vs_oracle(ctx, seq_type, ntaxa, nsite, safe):
    w = ia.Workload(seq_type=seq_type, ntaxa=ntaxa, nsite=nsite, ncat=4, mean_len=0.12, alpha=0.6,
                    seed=401 + ntaxa)
    st, fq, inv, nasc = asc_patterns(w)
    pobs = st.shape[1] - nasc
    inst = asc_instance(ctx, w, st, fq, inv, nasc, safe=safe)
    orc = asc_oracle(w, st, fq, inv, nasc, safe)
    planner = ia.PhyloTree(None, w)
    try:
        for a, b in planner.branches()[::5]:
            ops, node, dad, length = planner.plan(a, b)
            lg, pg = inst.likelihood_branch(ops, node, dad, length, want_pattern_lh=True)
            lo, po = orc.likelihood_branch(ops, node, dad, length, want_pattern_lh=True)
            assert abs(lg - lo) <= 1e-9 * abs(lo), (a, b, lg, lo)
            np.testing.assert_allclose(pg[:pobs], po, rtol=0, atol=1e-9)
            assert np.all(pg[pobs:] == 0.0)
            inst.likelihood_buffer(ops, node, dad)
            orc.likelihood_buffer(ops, node, dad)
            for t in (length, 0.3 * length + 1e-4, 2.0 * length + 0.01):
                dfg, ddfg, lng = inst.likelihood_derv(t)
                dfo, ddfo = orc.likelihood_derv(t)
                assert abs(dfg - dfo) <= 1e-7 * max(1.0, abs(dfo))
                assert abs(ddfg - ddfo) <= 1e-7 * max(1.0, abs(ddfo))
                lbo = orc.likelihood_from_buffer(t)
                assert abs(lng - lbo) <= 1e-9 * abs(lbo)
                assert abs(inst.likelihood_from_buffer(t) - lbo) <= 1e-9 * abs(lbo)
        # Newton: the same length and step count
        for a, b in planner.branches()[1::7]:
            ops, node, dad, length = planner.plan(b, a)
            xg, lg, ng = inst.optimize_branch(ops, node, dad, length)
            xo, lo, no, so = orc.optimize_branch(ops, node, dad, length)
            assert so == 0 and ng == no
            assert abs(xg - xo) <= 1e-7 * max(1.0, xo)
            assert abs(lg - lo) <= 1e-9 * abs(lo)
    finally:
        inst.close()
        planner.close()


def test_asc_rell_fp32_and_batches(ctx):
    w = ia.Workload(seq_type=ia.DNA, ntaxa=30, nsite=2500, ncat=4, mean_len=0.1, alpha=0.5,
                    seed=77)
    st, fq, inv, nasc = asc_patterns(w)
    pobs = st.shape[1] - nasc
    inst = asc_instance(ctx, w, st, fq, inv, nasc)
    inst32 = asc_instance(ctx, w, st, fq, inv, nasc, precision=ia.PREC_F32)
    plain = ia.Instance.from_workload(ctx, w)           # an item without +ASC
    planner = ia.PhyloTree(None, w)
    try:
        ops, node, dad, length = planner.plan(*planner.branches()[0])
        lnl, plh = inst.likelihood_branch(ops, node, dad, length, want_pattern_lh=True)
        # RELL reweights the conditioned pattern lnL
        rng = np.random.default_rng(3)
        wts = np.zeros((20, st.shape[1]), np.float32)
        for k in range(20):
            wts[k, :pobs] = rng.multinomial(int(fq[:pobs].sum()), fq[:pobs] / fq[:pobs].sum())
        np.testing.assert_allclose(inst.rell(wts), wts.astype(np.float64) @ plh, rtol=1e-12)
        np.testing.assert_allclose(inst.rell(fq[None, :].astype(np.float32)), [lnl], rtol=1e-12)
        # FP32 storage: the same conditioning
        l32 = inst32.likelihood_branch(ops, node, dad, length)
        assert abs(l32 - lnl) <= 1e-6 * abs(lnl)
        # a batch with an +ASC item: that item is evaluated alone, bit for bit
        out = ia.likelihood_branch_batch(ctx, [(inst, ops, node, dad, length),
                                               (plain, ops, node, dad, length)])
        assert out[0] == lnl
        lp = plain.likelihood_branch(ops, node, dad, length)
        assert abs(out[1] - lp) <= 1e-12 * abs(lp)
        # the batched derivative call leaves +ASC instances to plk_likelihood_derv
        inst.likelihood_buffer(ops, node, dad)
        with pytest.raises(ia.PlkError):
            ia.likelihood_derv_batch(ctx, [(inst, length)])
        # turning the correction off gives the observed patterns' plain lnL
        inst.set_asc(0)
        l0 = inst.likelihood_branch(ops, node, dad, length)
        orc = OracleTree.from_arrays(w.ntaxa, w.nstates, w.ncat, w.state_unknown, st, fq, inv,
                                     w.model_arrays(), nthreads=2)
        lo = orc.likelihood_branch(ops, node, dad, length)
        assert abs(l0 - lo) <= 1e-9 * abs(lo)
        assert l0 < lnl
    finally:
        for x in (inst, inst32, plain):
            x.close()
        planner.close()


@pytest.mark.parametrize("seq_type,ntaxa,nsite,safe", [(ia.DNA, 12, 3000, False),
                                                       (ia.DNA, 9, 2000, True),
                                                       (ia.PROTEIN, 4, 4000, False)])
def test_asc_informative_vs_oracle(ctx, seq_type, ntaxa, nsite, safe):
    """+ASC_INF: the informative columns, then every uninformative pattern
    as the unobserved set (the oracle's form is pinned by brute force in
    tests/test_oracle_asc_informative.py)."""
    from test_oracle_asc_informative import informative_set
    w = ia.Workload(seq_type=seq_type, ntaxa=ntaxa, nsite=nsite, ncat=4, mean_len=0.3, alpha=0.6,
                    seed=913 + ntaxa)
    st, fq, inv, nasc = informative_set(w)
    pobs = st.shape[1] - nasc
    assert pobs > 10
    inst = asc_instance(ctx, w, st, fq, inv, nasc, safe=safe)
    orc = asc_oracle(w, st, fq, inv, nasc, safe)
    planner = ia.PhyloTree(None, w)
    try:
        for a, b in planner.branches()[::3]:
            ops, node, dad, length = planner.plan(a, b)
            lg, pg = inst.likelihood_branch(ops, node, dad, length, want_pattern_lh=True)
            lo, po = orc.likelihood_branch(ops, node, dad, length, want_pattern_lh=True)
            assert abs(lg - lo) <= 1e-9 * abs(lo), (a, b, lg, lo)
            np.testing.assert_allclose(pg[:pobs], po, rtol=0, atol=1e-9)
            inst.likelihood_buffer(ops, node, dad)
            orc.likelihood_buffer(ops, node, dad)
            for t in (length, 0.5 * length + 1e-4):
                dfg, ddfg, _ = inst.likelihood_derv(t)
                dfo, ddfo = orc.likelihood_derv(t)
                assert abs(dfg - dfo) <= 1e-7 * max(1.0, abs(dfo))
                assert abs(ddfg - ddfo) <= 1e-7 * max(1.0, abs(ddfo))
        a, b = planner.branches()[1]
        ops, node, dad, length = planner.plan(b, a)
        xg, lg, ng = inst.optimize_branch(ops, node, dad, length)
        xo, lo, no, so = orc.optimize_branch(ops, node, dad, length)
        assert so == 0 and ng == no
        assert abs(xg - xo) <= 1e-7 * max(1.0, xo)
        assert abs(lg - lo) <= 1e-9 * abs(lo)
    finally:
        inst.close()
        planner.close()
