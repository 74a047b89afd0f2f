"""Pins the CPU oracle (oracle/oracle.c) against mathematics it does not share
code with: brute-force state summation, real-basis expm pruning, analytic
branch derivatives, closed-form model checks, and planted mistakes that those
checks must catch. The reference ships no golden vectors for this path."""
import numpy as np
import pytest
from scipy.special import gammainc, gammaincinv

import iqtree2_amd as ia
from oracle.binding import OracleTree
from refmath import Pruner, brute_force_lnl, indicators, invar_by_definition


def p_inv_of(w):
    return float(w.get("params")[1])


def pruner_for(w):
    """Independent pruning; +I from p_inv and the tip states, not the
    product's ptn_invar."""
    return Pruner(w.ntaxa, w.edges(), w.get("states"), w.get("Q"), w.get("pi"), w.get("rates"),
                  w.get("props"), p_inv_of(w), w.get("freq"),
                  indicators(w.seq_type, w.nstates, w.state_unknown))


def oracle_branch(w, planner, orc, a, b):
    ops, node, dad, length = planner.plan(a, b)
    return orc.likelihood_branch(ops, node, dad, length, want_pattern_lh=True)


@pytest.mark.parametrize("seq_type,ntaxa,nsite,p_inv,gap", [
    (ia.DNA, 5, 300, 0.2, 0.05),
    (ia.DNA, 6, 200, 0.0, 0.0),
    (ia.PROTEIN, 4, 150, 0.1, 0.0),
    (ia.CODON, 4, 80, 0.0, 0.0),
])
def test_brute_force_small_trees(seq_type, ntaxa, nsite, p_inv, gap):
    w = ia.Workload(seq_type=seq_type, ntaxa=ntaxa, nsite=nsite, ncat=4, mean_len=0.3,
                    alpha=0.7, p_inv=p_inv, gap_frac=gap, seed=11 + ntaxa)
    planner = ia.PhyloTree(None, w)
    # codon models run with per-category scaling, as IQ-TREE forces for them
    orc = OracleTree.from_workload(w, safe_numeric=seq_type == ia.CODON)
    ref, ref_p = brute_force_lnl(w.ntaxa, w.edges(), w.get("states"), w.get("Q"), w.get("pi"),
                                 w.get("rates"), w.get("props"), p_inv, w.get("freq"),
                                 indicators(seq_type, w.nstates, w.state_unknown))
    for a, b in planner.branches():
        lnl, plh = oracle_branch(w, planner, orc, a, b)
        assert abs(lnl - ref) <= 1e-11 * abs(ref), (a, b, lnl, ref)
        np.testing.assert_allclose(plh, ref_p, rtol=0, atol=1e-10)


@pytest.mark.parametrize("seq_type,ntaxa,nsite,mean_len,p_inv", [
    (ia.DNA, 40, 2000, 0.1, 0.2),
    (ia.PROTEIN, 24, 600, 0.2, 0.0),
    (ia.CODON, 16, 300, 0.15, 0.1),
])
def test_expm_pruning(seq_type, ntaxa, nsite, mean_len, p_inv):
    w = ia.Workload(seq_type=seq_type, ntaxa=ntaxa, nsite=nsite, ncat=4, mean_len=mean_len,
                    alpha=0.5, p_inv=p_inv, gap_frac=0.02 if seq_type == ia.DNA else 0.0, seed=7)
    planner = ia.PhyloTree(None, w)
    orc = OracleTree.from_workload(w, safe_numeric=seq_type == ia.CODON, nthreads=4)
    pr = pruner_for(w)
    branches = planner.branches()
    for a, b in branches[:: max(1, len(branches) // 6)]:
        lnl, plh = oracle_branch(w, planner, orc, a, b)
        ref, ref_p = pr.lnl(a, b, planner.get_length(a, b))
        assert abs(lnl - ref) <= 1e-10 * abs(ref)
        np.testing.assert_allclose(plh, ref_p, rtol=1e-10, atol=1e-10)


def test_scaling_exact_and_safe_mode():
    """A tree deep enough that partials cross 2^-256: scaling must be exact."""
    w = ia.Workload(seq_type=ia.DNA, ntaxa=300, nsite=400, ncat=4, mean_len=0.6, alpha=1.0,
                    seed=5)
    planner = ia.PhyloTree(None, w)
    orc = OracleTree.from_workload(w, nthreads=4)
    safe = OracleTree.from_workload(w, safe_numeric=True, nthreads=4)
    pr = pruner_for(w)
    a, b = planner.branches()[0]
    ops, node, dad, length = planner.plan(a, b)
    lnl = orc.likelihood_branch(ops, node, dad, length)
    lnl_safe = safe.likelihood_branch(ops, node, dad, length)
    scales = [orc.read_partial(op["dad_slot"])[1].max() for op in ops]
    assert max(scales) >= 1, "workload was meant to trigger rescaling"
    ref, _ = pr.lnl(a, b, length)
    assert abs(lnl - ref) <= 1e-10 * abs(ref)
    assert abs(lnl_safe - ref) <= 1e-10 * abs(ref)


def test_derivatives_analytic():
    w = ia.Workload(seq_type=ia.DNA, ntaxa=20, nsite=1500, ncat=4, mean_len=0.1, alpha=0.5,
                    p_inv=0.15, gap_frac=0.02, seed=3)
    planner = ia.PhyloTree(None, w)
    orc = OracleTree.from_workload(w)
    pr = pruner_for(w)
    for a, b in planner.branches()[::5]:
        ops, node, dad, length = planner.plan(a, b)
        orc.likelihood_buffer(ops, node, dad)
        for t in (length, 0.5 * length + 1e-4, 3 * length + 0.01):
            df, ddf = orc.likelihood_derv(t)
            lnl_b = orc.likelihood_from_buffer(t)
            ref_lnl, ref_df, ref_ddf = pr.derivs(a, b, t)
            assert abs(lnl_b - ref_lnl) <= 1e-10 * abs(ref_lnl)
            assert abs(df - ref_df) <= 1e-8 * max(1.0, abs(ref_df))
            assert abs(ddf - ref_ddf) <= 1e-8 * max(1.0, abs(ref_ddf))


@pytest.mark.parametrize("ncat", [1, 3, 4])
def test_codon_derivatives_and_odd_state_quirk(ncat):
    """61 states: with an odd block (ncat odd) the oracle's derivatives are
    the analytic ones. With an even block the reference's dotProductTriple
    (tree/phylokernelnew.h:668-716 called with N = block) counts the last
    (category, state) term of lh, d1 and d2 twice; the oracle must match the
    analytic derivatives with exactly that term added, computed here from
    real-basis partials and the last row of U^-1."""
    w = ia.Workload(seq_type=ia.CODON, ntaxa=6, nsite=200, ncat=ncat, mean_len=0.1, alpha=0.5,
                    p_inv=0.1, seed=5)
    planner = ia.PhyloTree(None, w)
    orc = OracleTree.from_workload(w, safe_numeric=True)
    pr = pruner_for(w)
    ev, inv = w.get("eval"), w.get("inv_evec")
    rates, props, freq, invar = w.get("rates"), w.get("props"), w.get("freq"), pr.invar
    block_even = (w.nstates * ncat) % 2 == 0
    for a, b in planner.branches()[::3]:
        ops, node, dad, length = planner.plan(a, b)
        orc.likelihood_buffer(ops, node, dad)
        for t in (length, 2.5 * length + 1e-3):
            df, ddf = orc.likelihood_derv(t)
            A, sa = pr.side(a, b)
            B, sb = pr.side(b, a)
            Pt = pr.P(t)
            rQ = rates[:, None, None] * w.get("Q")[None]
            P1 = np.einsum("cxz,czy->cxy", rQ, Pt)
            P2 = np.einsum("cxz,czy->cxy", rQ, P1)
            pi = w.get("pi")
            sc = np.exp(sa + sb)
            s = [np.einsum("c,x,cpx,cxy,cpy->p", props, pi, A, M, B) * sc for M in (Pt, P1, P2)]
            if block_even:
                c = ncat - 1
                cof = ev[-1] * rates[c]
                x = props[c] * np.exp(cof * t) * (A[c] @ inv[-1]) * (B[c] @ inv[-1]) * sc
                s = [s[0] + x, s[1] + cof * x, s[2] + cof * cof * x]
            lh = np.abs(s[0]) + invar
            ref_df = float(np.dot(freq, s[1] / lh))
            ref_ddf = float(np.dot(freq, s[2] / lh - (s[1] / lh) ** 2))
            assert abs(df - ref_df) <= 1e-8 * max(1.0, abs(ref_df)), (ncat, a, b, t, df, ref_df)
            assert abs(ddf - ref_ddf) <= 1e-8 * max(1.0, abs(ref_ddf)), (ncat, a, b, t, ddf, ref_ddf)


def test_newton_reaches_stationary_point():
    w = ia.Workload(seq_type=ia.DNA, ntaxa=12, nsite=800, ncat=4, mean_len=0.1, alpha=0.5,
                    seed=9)
    planner = ia.PhyloTree(None, w)
    orc = OracleTree.from_workload(w)
    pr = pruner_for(w)
    for a, b in planner.branches()[:6]:
        ops, node, dad, length = planner.plan(a, b)
        x, lnl, nd, st = orc.optimize_branch(ops, node, dad, length)
        assert st == 0 and nd >= 1
        _, df, ddf = pr.derivs(a, b, x)
        if 1e-6 < x < 9.5:
            assert abs(df) < 1e-2 and ddf < 0
        # not worse than any point on a grid around it
        grid = [max(1e-6, x * f) for f in (0.5, 0.9, 1.1, 2.0)]
        assert all(pr.lnl(a, b, g)[0] <= lnl + 1e-7 for g in grid)


def test_minimize_newton_on_a_parabola():
    """minimizeNewton on f(x) = (x - 0.3)^2 lands on 0.3 and honours the bracket."""
    import ctypes
    from oracle import binding

    calls = []

    @ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_double, ctypes.POINTER(ctypes.c_double),
                      ctypes.POINTER(ctypes.c_double))
    def fn(user, x, f, df):
        calls.append(x)
        f[0] = 2 * (x - target[0])
        df[0] = 2.0

    L = binding.lib()
    L.orc_minimize_newton.restype = ctypes.c_double
    L.orc_minimize_newton.argtypes = [ctypes.c_double] * 4 + [
        ctypes.POINTER(ctypes.c_double), ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
        ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
    for target, lo, hi, expect in (([0.3], 1e-6, 10.0, 0.3), ([25.0], 1e-6, 10.0, 10.0)):
        d2l, ne, st = ctypes.c_double(), ctypes.c_int(), ctypes.c_int()
        x = L.orc_minimize_newton(lo, 1.0, hi, 1e-6, ctypes.byref(d2l), 100,
                                  ctypes.cast(fn, ctypes.c_void_p), None, ctypes.byref(ne),
                                  ctypes.byref(st))
        assert st.value == 0
        assert abs(x - expect) < 1e-5


@pytest.mark.parametrize("alpha,p_inv", [(0.5, 0.0), (0.8, 0.0), (0.5, 0.2), (2.0, 0.4)])
def test_gamma_rates_against_scipy(alpha, p_inv):
    C = 4
    rates, props = np.zeros(C), np.zeros(C)
    assert ia.lib().plkh_gamma_rates(C, alpha, p_inv, rates.ctypes.data_as(
        __import__("ctypes").POINTER(__import__("ctypes").c_double)),
        props.ctypes.data_as(__import__("ctypes").POINTER(__import__("ctypes").c_double))) == 0
    # Yang (1994) mean rates from scipy's incomplete gamma
    cuts = gammaincinv(alpha, np.arange(1, C) / C) / alpha
    cdf1 = gammainc(alpha + 1, cuts * alpha)
    ref = np.diff(np.concatenate([[0.0], cdf1, [1.0]])) * C
    ref *= (C / (1 - p_inv)) / ref.sum()
    np.testing.assert_allclose(rates, ref, rtol=1e-12)
    np.testing.assert_allclose(props, (1 - p_inv) / C, rtol=0, atol=0)
    assert abs(np.dot(rates, props) - 1.0) < 1e-13


@pytest.mark.parametrize("seq_type", [ia.DNA, ia.PROTEIN, ia.CODON])
def test_eigensystem(seq_type):
    w = ia.Workload(seq_type=seq_type, ntaxa=4, nsite=10, seed=21)
    Q, pi, lam, U, Ui = w.get("Q"), w.get("pi"), w.get("eval"), w.get("evec"), w.get("inv_evec")
    n = w.nstates
    np.testing.assert_allclose(Q.sum(axis=1), 0, atol=1e-13)
    assert abs(-np.dot(pi, np.diag(Q)) - 1.0) < 1e-13                 # mean rate 1
    np.testing.assert_allclose(pi[:, None] * Q, (pi[:, None] * Q).T, atol=1e-14)  # reversibility
    np.testing.assert_allclose(U @ Ui, np.eye(n), atol=1e-11)
    np.testing.assert_allclose(Q @ U, U * lam[None, :], atol=1e-11)
    assert np.all(np.diff(lam) >= 0) and abs(lam[-1]) < 1e-12         # ascending, one zero
    S = np.sqrt(pi)[:, None] * Q / np.sqrt(pi)[None, :]
    np.testing.assert_allclose(lam, np.linalg.eigvalsh((S + S.T) / 2), atol=1e-12)
    tip = w.get("tip")
    np.testing.assert_allclose(tip, indicators(seq_type, n, w.state_unknown) @ Ui.T, atol=1e-14)


def test_planted_mistakes_are_caught():
    """Each plausible mistake, planted in the oracle's inputs, must break the
    brute-force agreement that test_brute_force_small_trees relies on."""
    w = ia.Workload(seq_type=ia.DNA, ntaxa=5, nsite=300, ncat=4, mean_len=0.3, alpha=0.7,
                    p_inv=0.2, gap_frac=0.05, seed=16)
    planner = ia.PhyloTree(None, w)
    ref, _ = brute_force_lnl(w.ntaxa, w.edges(), w.get("states"), w.get("Q"), w.get("pi"),
                             w.get("rates"), w.get("props"), 0.2, w.get("freq"),
                             indicators(ia.DNA, 4, 18))
    a, b = planner.branches()[0]
    ops, node, dad, length = planner.plan(a, b)
    m = w.model_arrays()
    mistakes = {
        "transposed inv_evec": dict(m, inv_evec=m["inv_evec"].T.copy()),
        "swapped eigenvalues": dict(m, eval=m["eval"][[1, 0, 2, 3]].copy()),
        "dropped category weight": dict(m, props=m["props"] * np.array([1, 1, 1, 0.0])),
        "rates not rescaled for +I": dict(m, rates=m["rates"] * 0.8),
        "ambiguity tip as unknown": dict(m, tip=np.where(np.arange(19)[:, None] == 8,
                                                           m["tip"][18], m["tip"])),
    }
    for name, mm in mistakes.items():
        orc = OracleTree.from_arrays(w.ntaxa, 4, 4, 18, w.get("states"), w.get("freq"),
                                     w.get("invar"), mm)
        lnl = orc.likelihood_branch(ops, node, dad, length)
        assert abs(lnl - ref) > 1e-8 * abs(ref), f"planted mistake not detected: {name}"
    # a dropped +I term
    orc = OracleTree.from_arrays(w.ntaxa, 4, 4, 18, w.get("states"), w.get("freq"),
                                 np.zeros(w.npatterns), m)
    assert abs(orc.likelihood_branch(ops, node, dad, length) - ref) > 1e-8 * abs(ref)
