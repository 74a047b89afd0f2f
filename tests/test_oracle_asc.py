"""Lewis's ascertainment bias correction (+ASC, ASC_VARIANT) in the oracle,
pinned by mathematics it does not share code with: the conditional lnL
sum_p f_p log L_p - nsite * log(1 - sum_s L(constant column of state s)), every
L from brute-force state summation (tiny trees) or real-basis expm pruning,
and its branch derivatives from the analytic derivatives of those pieces and
from central differences of the conditional lnL itself."""
import numpy as np
import pytest

import iqtree2_amd as ia
from oracle.binding import OracleTree
from refmath import Pruner, brute_force_lnl, indicators


def asc_patterns(w):
    """The workload's variable patterns (constant columns dropped, as in SNP
    data), then the nstates constant patterns as the unobserved ones
    (weight 0, +I mass 0)."""
    states, freq, n = w.get("states"), w.get("freq"), w.nstates
    const = np.all(states == states[:1], axis=0) & (states[0] < n)
    obs = np.ascontiguousarray(states[:, ~const])
    cst = np.tile(np.arange(n, dtype=np.uint8), (w.ntaxa, 1))
    st = np.ascontiguousarray(np.concatenate([obs, cst], axis=1))
    fq = np.concatenate([freq[~const], np.zeros(n)])
    return st, fq, np.zeros(st.shape[1]), n


def asc_oracle(w, st, fq, inv, nasc, safe=False):
    orc = OracleTree.from_arrays(w.ntaxa, w.nstates, w.ncat, w.state_unknown, st, fq, inv,
                                 w.model_arrays(), safe_numeric=safe, nthreads=2)
    orc.set_asc(nasc)
    return orc


def pruners(w, st, fq, nasc):
    ind = indicators(w.seq_type, w.nstates, w.state_unknown)
    args = (w.ntaxa, w.edges())
    model = (w.get("Q"), w.get("pi"), w.get("rates"), w.get("props"), 0.0)
    pobs = st.shape[1] - nasc
    obs = Pruner(*args, st[:, :pobs], *model, fq[:pobs], ind)
    cst = Pruner(*args, np.ascontiguousarray(st[:, pobs:]), *model, np.ones(nasc), ind)
    return obs, cst, float(fq[:pobs].sum())


def conditional(obs, cst, nsite, a, b, t):
    """lnL, df, ddf of the conditional likelihood on branch (a, b) at t."""
    lnl, df, ddf = obs.derivs(a, b, t)
    log_lh, d1n, d2n = cst.branch_terms(a, b, t)
    lh = np.exp(log_lh)                       # the constant columns' likelihoods
    pc, dpc, ddpc = lh.sum(), (d1n * lh).sum(), (d2n * lh).sum()
    pv = 1.0 - pc
    return (lnl - nsite * np.log(pv), df + nsite * dpc / pv,
            ddf + nsite * (ddpc / pv + (dpc / pv) ** 2))


@pytest.mark.parametrize("seq_type,ntaxa,nsite", [(ia.DNA, 5, 400), (ia.DNA, 6, 250),
                                                  (ia.PROTEIN, 4, 200)])
def test_asc_lnl_by_brute_force(seq_type, ntaxa, nsite):
    w = ia.Workload(seq_type=seq_type, ntaxa=ntaxa, nsite=nsite, ncat=4, mean_len=0.3, alpha=0.7,
                    seed=31 + ntaxa)
    st, fq, inv, nasc = asc_patterns(w)
    pobs = st.shape[1] - nasc
    orc = asc_oracle(w, st, fq, inv, nasc)
    planner = ia.PhyloTree(None, w)
    ind = indicators(seq_type, w.nstates, w.state_unknown)
    model = (w.get("Q"), w.get("pi"), w.get("rates"), w.get("props"), 0.0)
    lobs, plh_obs = brute_force_lnl(w.ntaxa, w.edges(), st[:, :pobs], *model, fq[:pobs], ind)
    _, plh_c = brute_force_lnl(w.ntaxa, w.edges(), np.ascontiguousarray(st[:, pobs:]), *model,
                               np.ones(nasc), ind)
    corr = np.log(1.0 - np.exp(plh_c).sum())
    ref = lobs - fq[:pobs].sum() * corr
    assert corr < 0.0
    for a, b in planner.branches():
        lnl, plh = orc.likelihood_branch(*planner.plan(a, b), want_pattern_lh=True)
        assert abs(lnl - ref) <= 1e-11 * abs(ref), (a, b, lnl, ref)
        np.testing.assert_allclose(plh, plh_obs - corr, rtol=0, atol=1e-10)
    # without the correction: the plain lnL of the observed patterns
    orc.set_asc(0)
    lnl0 = orc.likelihood_branch(*planner.plan(*planner.branches()[0]))
    assert abs(lnl0 - lobs) <= 1e-11 * abs(lobs)


@pytest.mark.parametrize("seq_type,ntaxa,nsite,safe", [(ia.DNA, 14, 900, False),
                                                       (ia.DNA, 10, 600, True),
                                                       (ia.PROTEIN, 8, 400, False)])
def test_asc_derivatives(seq_type, ntaxa, nsite, safe):
    w = ia.Workload(seq_type=seq_type, ntaxa=ntaxa, nsite=nsite, ncat=4, mean_len=0.15,
                    alpha=0.6, seed=57 + ntaxa)
    st, fq, inv, nasc = asc_patterns(w)
    orc = asc_oracle(w, st, fq, inv, nasc, safe)
    obs, cst, nsite_obs = pruners(w, st, fq, nasc)
    planner = ia.PhyloTree(None, w)
    for a, b in planner.branches()[::4]:
        ops, node, dad, length = planner.plan(a, b)
        lnl_branch = orc.likelihood_branch(ops, node, dad, length)
        orc.likelihood_buffer(ops, node, dad)
        for t in (length, 0.4 * length + 1e-4, 2.5 * length + 0.02):
            df, ddf = orc.likelihood_derv(t)
            lnl_b = orc.likelihood_from_buffer(t)
            ref_lnl, ref_df, ref_ddf = conditional(obs, cst, nsite_obs, a, b, t)
            assert abs(lnl_b - ref_lnl) <= 1e-10 * abs(ref_lnl)
            assert abs(df - ref_df) <= 1e-8 * max(1.0, abs(ref_df))
            assert abs(ddf - ref_ddf) <= 1e-8 * max(1.0, abs(ref_ddf))
            # the correction's derivative terms by central differences of the
            # conditional lnL (sign and factor of nsite * P'/(1 - P) and its
            # derivative)
            h = 1e-5 * max(t, 1e-3)
            fp = conditional(obs, cst, nsite_obs, a, b, t + h)[0]
            fm = conditional(obs, cst, nsite_obs, a, b, t - h)[0]
            assert abs((fp - fm) / (2 * h) - ref_df) <= 1e-5 * max(1.0, abs(ref_df))
        assert abs(lnl_branch - orc.likelihood_from_buffer(length)) <= 1e-10 * abs(lnl_branch)


def test_asc_newton_moves_toward_the_conditional_optimum():
    """optimizeOneBranch on the conditional lnL stops where its derivative
    vanishes (or at a bound)."""
    w = ia.Workload(seq_type=ia.DNA, ntaxa=12, nsite=800, ncat=4, mean_len=0.1, alpha=0.5, seed=5)
    st, fq, inv, nasc = asc_patterns(w)
    orc = asc_oracle(w, st, fq, inv, nasc)
    obs, cst, nsite_obs = pruners(w, st, fq, nasc)
    planner = ia.PhyloTree(None, w)
    for a, b in planner.branches()[::5]:
        ops, node, dad, length = planner.plan(b, a)
        x, lnl, nd, status = orc.optimize_branch(ops, node, dad, length)
        assert status == 0 and nd > 0
        _, df, ddf = conditional(obs, cst, nsite_obs, a, b, x)
        assert abs(df) <= 1e-3 * max(1.0, abs(ddf)) or x <= 1.0001e-6 or x >= 9.99
