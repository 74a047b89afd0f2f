"""+ASC_INF (ASC_INFORMATIVE): Lewis's correction conditioned on
parsimony-informative sites. The likelihood is the +ASC one
(tree/phylokernelnew.h:2272-2273 treats ASC_INFORMATIVE as ASC_VARIANT); only
the unobserved set differs: every parsimony-uninformative pattern, from
Alignment::getUnobservedConstPatterns (alignment/alignment.cpp:4181-4196).

The host enumerator plkh_uninformative_patterns is pinned by the definition
of an informative site (alignment/alignment.cpp:943-969: at least two states,
each in at least two sequences) over every one of the nstates^nseq patterns,
by hand-derived first rows of the reference's order, and by the one case the
reference's rule leaves out (a pattern with nseq distinct states is kept only
when it holds state 0). The conditional lnL is pinned by brute-force state
summation, with the closure sum over all nstates^nseq patterns of L = 1."""
import itertools

import numpy as np
import pytest

import iqtree2_amd as ia
from oracle.binding import OracleTree
from refmath import brute_force_lnl, indicators


def by_definition(nstates, nseq):
    """All nstates^nseq patterns as rows, and which are informative."""
    rows = np.array(list(itertools.product(range(nstates), repeat=nseq)), np.uint8)
    counts = np.stack([(rows == s).sum(axis=1) for s in range(nstates)], axis=1)
    return rows, (counts >= 2).sum(axis=1) >= 2


def informative_columns(states, nstates):
    """Alignment::computeConst's is_informative per column (unambiguous
    states only are counted)."""
    counts = np.stack([(states == s).sum(axis=0) for s in range(nstates)], axis=0)
    return (counts >= 2).sum(axis=0) >= 2


@pytest.mark.parametrize("nstates,nseq", [(4, 5), (4, 6), (4, 3), (4, 4), (2, 7), (20, 3), (5, 4)])
def test_uninformative_set_by_definition(nstates, nseq):
    got = ia.uninformative_patterns(nstates, nseq)
    assert got.shape[1] == nseq
    rows, inf = by_definition(nstates, nseq)
    expect = {r.tobytes() for r in rows[~inf]}
    # the reference's rule: a pattern of nseq distinct states comes only from
    # repeat state 0 (the singleton set of size nseq - 1 is taken for repeat 0
    # alone), so those without state 0 are not in the set
    distinct = np.array([len(set(r)) == nseq for r in rows])
    dropped = {r.tobytes() for r in rows[distinct & ~np.any(rows == 0, axis=1)]}
    assert dropped <= expect
    assert (len(dropped) > 0) == (nseq <= nstates - 1)
    got_set = {r.tobytes() for r in got}
    assert len(got_set) == got.shape[0]           # no pattern twice
    assert got_set == expect - dropped


def test_uninformative_order():
    """The reference's order: repeat state outermost, singleton subsets in
    binary-counter order over the other states, sequences for the singletons
    chosen in increasing index, one after another."""
    got = ia.uninformative_patterns(4, 5)
    first = ["00000",                                       # repeat 0, {}
             "10000", "01000", "00100", "00010", "00001",   # {1}
             "20000", "02000", "00200", "00020", "00002",   # {2}
             "12000", "10200", "10020", "10002", "21000"]   # {1, 2}: 1 at 0, then 2 at 1..4; 1 at 1 ...
    assert ["".join(map(str, r)) for r in got[:len(first)]] == first
    # repeat 1's block starts with the constant pattern of state 1
    n0 = sum(1 for r in got if np.count_nonzero(r == 0) >= 2 or len(set(r)) == 5)
    assert "".join(map(str, got[n0])) == "11111"


def test_uninformative_limits():
    with pytest.raises(ia.PlkError):
        ia.uninformative_patterns(20, 12, max_patterns=1000)
    with pytest.raises(ia.PlkError):
        ia.uninformative_patterns(32, 3)


def informative_set(w):
    """The workload's informative columns, then every uninformative pattern
    as the unobserved ones (weight 0, +I mass 0)."""
    states, freq, n = w.get("states"), w.get("freq"), w.nstates
    keep = informative_columns(states, n)
    unobs = ia.uninformative_patterns(n, w.ntaxa).T
    st = np.ascontiguousarray(np.concatenate([states[:, keep], unobs], axis=1))
    fq = np.concatenate([freq[keep], np.zeros(unobs.shape[1])])
    return st, fq, np.zeros(st.shape[1]), unobs.shape[1]


@pytest.mark.parametrize("seq_type,ntaxa,nsite", [(ia.DNA, 5, 1500), (ia.DNA, 6, 800)])
def test_asc_inf_lnl_by_brute_force(seq_type, ntaxa, nsite):
    w = ia.Workload(seq_type=seq_type, ntaxa=ntaxa, nsite=nsite, ncat=4, mean_len=0.4, alpha=0.7,
                    seed=71 + ntaxa)
    st, fq, inv, nasc = informative_set(w)
    pobs = st.shape[1] - nasc
    assert pobs > 10
    ind = indicators(seq_type, w.nstates, w.state_unknown)
    model = (w.get("Q"), w.get("pi"), w.get("rates"), w.get("props"), 0.0)
    lobs, plh_obs = brute_force_lnl(w.ntaxa, w.edges(), st[:, :pobs], *model, fq[:pobs], ind)
    _, plh_u = brute_force_lnl(w.ntaxa, w.edges(), np.ascontiguousarray(st[:, pobs:]), *model,
                               np.ones(nasc), ind)
    # closure: the likelihoods of all nstates^ntaxa patterns sum to one, so
    # 1 - P(uninformative) is P(informative) summed directly
    rows, inf = by_definition(w.nstates, w.ntaxa)
    _, plh_all = brute_force_lnl(w.ntaxa, w.edges(), np.ascontiguousarray(rows.T), *model,
                                 np.ones(rows.shape[0]), ind)
    assert abs(np.exp(plh_all).sum() - 1.0) < 1e-12
    p_inf = np.exp(plh_all[inf]).sum()
    assert abs((1.0 - np.exp(plh_u).sum()) - p_inf) < 1e-12
    corr = np.log(p_inf)
    ref = lobs - fq[:pobs].sum() * corr
    orc = OracleTree.from_arrays(w.ntaxa, w.nstates, w.ncat, w.state_unknown, st, fq, inv,
                                 w.model_arrays(), nthreads=2)
    orc.set_asc(nasc)
    planner = ia.PhyloTree(None, w)
    try:
        for a, b in planner.branches():
            lnl, plh = orc.likelihood_branch(*planner.plan(a, b), want_pattern_lh=True)
            assert abs(lnl - ref) <= 1e-11 * abs(ref), (a, b, lnl, ref)
            np.testing.assert_allclose(plh, plh_obs - corr, rtol=0, atol=1e-10)
    finally:
        planner.close()
