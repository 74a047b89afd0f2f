"""PhyloSuperTree::optimizeAllBranches (tree/phylosupertree.cpp:773-792) with
every partition's Newton loop advanced together on the device, against each
partition optimised alone (the host Newton loop of PhyloTree, already tested
against the oracle) and against the oracle's replay of optimizeAllBranches."""
import numpy as np
import pytest

import iqtree2_amd as ia
from oracle.binding import OracleTree

pytestmark = pytest.mark.gpu


def partitions(n=6, ntaxa=24, seed=4242, codon=False):
    """n partitions on one tree, alternating DNA GTR+G4 and protein +G4
    (independent model parameters), 300-1,500 sites; optionally one codon."""
    rng = np.random.default_rng(seed)
    base = ia.Workload(seq_type=ia.DNA, ntaxa=ntaxa, nsite=10, mean_len=0.1, seed=seed)
    edges = base.edges()
    base.close()
    parts = []
    for k in range(n):
        st = ia.PROTEIN if k % 2 else ia.DNA
        if codon and k == n - 1:
            st = ia.CODON
        nsite = int(rng.integers(300, 1500)) if st != ia.CODON else 200
        parts.append(ia.Workload(seq_type=st, ntaxa=ntaxa, nsite=nsite, ncat=4, alpha=-1.0,
                                 seed=seed * 100 + k, edges=edges))
    return parts


def lengths(tree):
    return np.array([tree.get_length(a, b) for a, b in tree.branches()])


@pytest.mark.parametrize("codon", [False, True])
def test_supertree_equals_each_partition_alone(ctx, codon):
    parts = partitions(codon=codon)
    together = [ia.PhyloTree(ctx, w) for w in parts]
    alone = [ia.PhyloTree(ctx, w) for w in parts]
    try:
        lnl_b, nd_b = ia.supertree_optimize_all_branches(together, iterations=3, tolerance=1e-3,
                                                         max_nr=100)
        tol = 1e-3 / min(len(parts), 10)
        for k, t in enumerate(alone):
            t.set_device_newton(False)   # the host Newton loop, one engine call per evaluation
            lnl_a, nd_a = t.optimize_all_branches(iterations=3, tolerance=tol, max_nr=100)
            assert nd_b[k] == nd_a, f"partition {k}: {nd_b[k]} vs {nd_a} derivative evaluations"
            assert abs(lnl_b[k] - lnl_a) <= 1e-10 * abs(lnl_a), f"partition {k}: {lnl_b[k]} vs {lnl_a}"
            np.testing.assert_array_equal(together[k].pass_counts(), t.pass_counts())
            np.testing.assert_allclose(lengths(together[k]), lengths(t), rtol=1e-12, atol=0)
            # the optimised lengths raise (or keep) the lnL
            assert lnl_b[k] >= together[k].compute_likelihood() - 1e-6
    finally:
        for t in together + alone:
            t.close()
        for w in parts:
            w.close()


def test_supertree_one_pass_vs_oracle(ctx):
    """One pass over two partitions (DNA, protein): each optimised length and
    the final lnL against the oracle's optimizeOneBranch replay."""
    parts = partitions(n=2, ntaxa=20, seed=777)
    trees = [ia.PhyloTree(ctx, w) for w in parts]
    try:
        lnl, nd = ia.supertree_optimize_all_branches(trees, iterations=1, tolerance=1e-3, max_nr=100)
        for k, w in enumerate(parts):
            planner = ia.PhyloTree(None, w)
            orc = OracleTree.from_workload(w, nthreads=4)
            branches = planner.branches()
            a0, b0 = branches[0]
            orc.likelihood_branch(*planner.plan(b0, a0))
            n_orc = 0
            for a, b in branches:
                ops, node, dad, length = planner.plan(b, a)
                x, _, ndb, st = orc.optimize_branch(ops, node, dad, length)
                assert st == 0
                n_orc += ndb
                planner.set_length(a, b, x)
                assert abs(trees[k].get_length(a, b) - x) <= 1e-7 * max(x, 1e-6)
            lo = orc.likelihood_from_buffer(planner.get_length(*branches[-1]))
            assert abs(lnl[k] - lo) <= 1e-9 * abs(lo)
            assert nd[k] == n_orc
            planner.close()
    finally:
        for t in trees:
            t.close()
        for w in parts:
            w.close()


def test_supertree_rejects_a_tree_twice(ctx):
    w = partitions(n=1)[0]
    t = ia.PhyloTree(ctx, w)
    try:
        with pytest.raises(ia.PlkError):
            ia.supertree_optimize_all_branches([t, t])
    finally:
        t.close()
        w.close()
