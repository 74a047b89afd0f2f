"""The oracle's state-basis kernels (non-reversible and mixture models,
tree/phylokernelnonrev.h) pinned by brute-force summation over internal
states with expm transition matrices: unrooted trees (the frequencies at the
node end of the evaluated branch), rooted trees (the root leaf and children
toward the root), mixtures (fused and not), a reversible model through the
state-basis path against the reversible eigen-basis path, derivatives against
central differences, and the safe kernel against the normal one."""
import numpy as np
import pytest

import iqtree2_amd as ia
from oracle.binding import OracleTree
from refmath import indicators
from statebasis import brute_force_rooted, categories, nonrev_model, plan

TOL = 1e-11


def oracle(w, nm, states=None, rooted=False, safe=False, nthreads=2):
    st = w.get("states") if states is None else states
    C = len(categories(nm))
    t = OracleTree(w.ntaxa, w.nstates, C, st.shape[1], w.ntaxa, w.state_unknown, safe, nthreads)
    t.set_patterns(st, w.get("freq"), np.zeros(st.shape[1]))
    t.set_model_nonrev(nm, indicators(w.seq_type, w.nstates, w.state_unknown), rooted)
    return t


def internal_branches(w):
    u, v, _ = w.edges()
    out = []
    for a, b in zip(u.tolist(), v.tolist()):
        if a >= w.ntaxa:
            out.append((a, b))
        if b >= w.ntaxa:
            out.append((b, a))
    return out


CASES = [(ia.DNA, 6, 1, 4, False), (ia.DNA, 5, 2, 2, False), (ia.DNA, 5, 4, 4, True),
         (ia.PROTEIN, 4, 1, 2, False)]


@pytest.mark.parametrize("seq_type,ntaxa,nmix,nrates,fused", CASES)
def test_unrooted_by_brute_force(seq_type, ntaxa, nmix, nrates, fused):
    w = ia.Workload(seq_type=seq_type, ntaxa=ntaxa, nsite=300, ncat=4, mean_len=0.3, seed=3 + ntaxa)
    nm = nonrev_model(w.nstates, nmix, nrates, fused, seed=11 + nmix)
    orc = oracle(w, nm)
    ind = indicators(seq_type, w.nstates, w.state_unknown)
    st, freq = w.get("states"), w.get("freq")
    for u, v in internal_branches(w):
        ops, node, dad, length = plan(w.ntaxa, w.edges(), u, v)
        lnl, plh = orc.likelihood_branch(ops, node, dad, length, want_pattern_lh=True)
        ref = np.log(brute_force_rooted(w.ntaxa, w.edges(), u, st, nm, ind))
        assert abs(lnl - ref @ freq) <= TOL * abs(lnl), (u, v)
        np.testing.assert_allclose(plh, ref, rtol=0, atol=1e-10)


def rooted_setup(seed=5, ntaxa=6, nmix=1, nrates=4):
    """Taxon 0 is the root leaf: zero-length branch, state 0 everywhere."""
    w0 = ia.Workload(seq_type=ia.DNA, ntaxa=ntaxa, nsite=300, ncat=4, mean_len=0.3, seed=seed)
    u, v, ln = (np.array(x) for x in w0.edges())
    ln = ln.copy()
    ln[(u == 0) | (v == 0)] = 0.0
    w = ia.Workload(seq_type=ia.DNA, ntaxa=ntaxa, nsite=300, ncat=4, mean_len=0.3, seed=seed,
                    edges=(u, v, ln))
    st = w.get("states").copy()
    st[0] = 0
    root = int(v[u == 0][0]) if np.any(u == 0) else int(u[v == 0][0])
    return w, st, root, nonrev_model(4, nmix, nrates, False, seed=seed)


def rooted_branches(w, root):
    """(upper, lower) per branch; the root leaf's branch as (root node, 0)."""
    from statebasis import adjacency, parents
    par = parents(adjacency(w.edges()), root)
    out = [(root, 0)]
    for x, p in par.items():
        if p is not None and x != 0:
            out.append((p, x))
    return out


@pytest.mark.parametrize("nmix,nrates", [(1, 4), (2, 2)])
def test_rooted_by_brute_force(nmix, nrates):
    w, st, root, nm = rooted_setup(nmix=nmix, nrates=nrates)
    orc = oracle(w, nm, states=st, rooted=True)
    ind = indicators(ia.DNA, 4, w.state_unknown)
    ref_plh = np.log(brute_force_rooted(w.ntaxa, w.edges(), root, st, nm, ind, skip_leaf=0))
    ref = ref_plh @ w.get("freq")
    for a, b in rooted_branches(w, root):
        ops, node, dad, length = plan(w.ntaxa, w.edges(), a, b, root=root, root_leaf=0)
        lnl, plh = orc.likelihood_branch(ops, node, dad, length, want_pattern_lh=True)
        assert abs(lnl - ref) <= TOL * abs(ref), (a, b, lnl, ref)
        np.testing.assert_allclose(plh, ref_plh, rtol=0, atol=1e-10)


def test_reversible_model_through_state_basis():
    """kernel_nonrev on a reversible model: the same lnL as the eigen-basis
    kernels (the root position does not matter for a reversible model)."""
    w = ia.Workload(seq_type=ia.DNA, ntaxa=12, nsite=800, ncat=4, mean_len=0.2, seed=21)
    nm = nonrev_model(4, 1, 4, False, seed=4, reversible=True)
    # the eigen-basis kernels need U^-1 = U^T diag(pi): the symmetric form
    Q, pi = nm["Q"][0], nm["pi"][0]
    s = np.sqrt(pi)
    ev, Us = np.linalg.eigh(Q * s[:, None] / s[None, :])
    V, Vi = Us / s[:, None], Us.T * s[None, :]
    nm.update(ceval=ev[None].astype(complex), cevec=V[None].astype(complex),
              cinv_evec=Vi[None].astype(complex))
    orc = oracle(w, nm)
    ind = indicators(ia.DNA, 4, w.state_unknown)
    rev = OracleTree(w.ntaxa, 4, 4, w.npatterns, w.ntaxa, w.state_unknown, False, 2)
    rev.set_patterns(w.get("states"), w.get("freq"), np.zeros(w.npatterns))
    rev.set_model(ev, V, Vi, ind @ Vi.T, nm["rates"], nm["props"])
    for u, v in internal_branches(w)[::3]:
        ops, node, dad, length = plan(w.ntaxa, w.edges(), u, v)
        a = orc.likelihood_branch(ops, node, dad, length)
        b = rev.likelihood_branch(ops, node, dad, length)
        assert abs(a - b) <= 1e-12 * abs(b), (u, v, a, b)


@pytest.mark.parametrize("rooted", [False, True])
def test_derivatives_and_newton(rooted):
    if rooted:
        w, st, root, nm = rooted_setup(seed=9, ntaxa=8)
        orc = oracle(w, nm, states=st, rooted=True)
        branches = rooted_branches(w, root)[1:]
        kw = dict(root=root, root_leaf=0)
    else:
        w = ia.Workload(seq_type=ia.DNA, ntaxa=10, nsite=600, ncat=4, mean_len=0.2, seed=13)
        nm = nonrev_model(4, 2, 2, False, seed=8)
        orc = oracle(w, nm)
        branches = internal_branches(w)
        kw = {}
    for a, b in branches[::2]:
        ops, node, dad, length = plan(w.ntaxa, w.edges(), a, b, **kw)
        lnl = orc.likelihood_branch(ops, node, dad, length)
        orc.likelihood_buffer(ops, node, dad)
        assert abs(orc.likelihood_from_buffer(length) - lnl) <= 1e-13 * abs(lnl)
        for t in (length, 0.5 * length + 1e-3, 2.0 * length + 0.05):
            df, ddf = orc.likelihood_derv(t)
            h = 1e-5 * max(t, 1e-2)
            fp, fm = (orc.likelihood_from_buffer(x) for x in (t + h, t - h))
            assert abs((fp - fm) / (2 * h) - df) <= 1e-5 * max(1.0, abs(df)), (a, b, t)
            h2 = 1e-3 * max(t, 0.05)
            fp, fm, f0 = (orc.likelihood_from_buffer(x) for x in (t + h2, t - h2, t))
            assert abs((fp - 2 * f0 + fm) / h2 ** 2 - ddf) <= 1e-3 * max(1.0, abs(ddf)), (a, b, t)
        x, lo, nd, status = orc.optimize_branch(ops, node, dad, length)
        assert status == 0 and nd > 0
        orc.likelihood_buffer(ops, node, dad)
        df, ddf = orc.likelihood_derv(x)
        # a zero of the derivative, or the lower bound with lnL falling
        assert abs(df) <= 1e-3 * max(1.0, abs(ddf)) or (df < 0 and x <= 1e-5), (a, b, x, df)
        assert lo >= lnl - 1e-9 * abs(lnl)


def test_safe_kernel_matches_normal():
    """A tree deep enough that partials are rescaled: the safe kernel's
    per-category counters give the normal kernel's lnL and derivatives."""
    w = ia.Workload(seq_type=ia.DNA, ntaxa=120, nsite=400, ncat=4, mean_len=0.8, seed=17)
    nm = nonrev_model(4, 2, 2, False, seed=3)
    norm, safe = oracle(w, nm), oracle(w, nm, safe=True)
    scaled = False
    for u, v in internal_branches(w)[::23]:
        ops, node, dad, length = plan(w.ntaxa, w.edges(), u, v)
        a, pa = norm.likelihood_branch(ops, node, dad, length, want_pattern_lh=True)
        b, pb = safe.likelihood_branch(ops, node, dad, length, want_pattern_lh=True)
        assert abs(a - b) <= 1e-11 * abs(a)
        np.testing.assert_allclose(pa, pb, rtol=0, atol=1e-9)
        norm.likelihood_buffer(ops, node, dad)
        safe.likelihood_buffer(ops, node, dad)
        da, dda = norm.likelihood_derv(length)
        db, ddb = safe.likelihood_derv(length)
        assert abs(da - db) <= 1e-9 * max(1.0, abs(da))
        assert abs(dda - ddb) <= 1e-9 * max(1.0, abs(dda))
        _, sc = norm.read_partial(node.id, partial=False)
        scaled = scaled or sc.max() > 0
    assert scaled


def test_asc_by_brute_force():
    """+ASC with a non-reversible model: the constant columns unobserved,
    lnL - nsite log(1 - sum_s L(const_s)), every L by brute force rooted at
    the node end; derivatives against central differences."""
    w = ia.Workload(seq_type=ia.DNA, ntaxa=6, nsite=500, ncat=4, mean_len=0.3, seed=41)
    nm = nonrev_model(4, 2, 2, False, seed=12)
    states, freq = w.get("states"), w.get("freq")
    const = np.all(states == states[:1], axis=0) & (states[0] < 4)
    cst = np.tile(np.arange(4, dtype=np.uint8), (w.ntaxa, 1))
    st = np.ascontiguousarray(np.concatenate([states[:, ~const], cst], axis=1))
    fq = np.concatenate([freq[~const], np.zeros(4)])
    pobs = st.shape[1] - 4
    C = len(categories(nm))
    orc = OracleTree(w.ntaxa, 4, C, st.shape[1], w.ntaxa, w.state_unknown, False, 2)
    orc.set_patterns(st, fq, np.zeros(st.shape[1]))
    orc.set_model_nonrev(nm, indicators(ia.DNA, 4, w.state_unknown))
    orc.set_asc(4)
    ind = indicators(ia.DNA, 4, w.state_unknown)
    for u, v in internal_branches(w)[::2]:
        ops, node, dad, length = plan(w.ntaxa, w.edges(), u, v)
        lnl, plh = orc.likelihood_branch(ops, node, dad, length, want_pattern_lh=True)
        L = brute_force_rooted(w.ntaxa, w.edges(), u, st, nm, ind)
        corr = np.log(1.0 - L[pobs:].sum())
        ref = np.log(L[:pobs]) @ fq[:pobs] - fq[:pobs].sum() * corr
        assert abs(lnl - ref) <= TOL * abs(ref), (u, v, lnl, ref)
        np.testing.assert_allclose(plh, np.log(L[:pobs]) - corr, rtol=0, atol=1e-10)
        orc.likelihood_buffer(ops, node, dad)
        df, ddf = orc.likelihood_derv(length)
        h = 1e-5 * max(length, 1e-2)
        fp, fm = (orc.likelihood_from_buffer(x) for x in (length + h, length - h))
        assert abs((fp - fm) / (2 * h) - df) <= 1e-5 * max(1.0, abs(df))
