"""Full-size workloads on the GPU, checked through properties of the
likelihood (the comparisons with the oracle at these sizes are in
test_full_size_oracle.py):

- pulley principle: for a reversible model the tree lnL is the same at every
  branch (tree/phylotree.cpp computeLikelihoodBranch on any branch);
- the tree lnL is the pattern-weighted sum of the pattern lnL it reports;
- derivatives agree with central differences of the lnL;
- one optimizeAllBranches pass (the device-resident sweep) never lowers the
  lnL and leaves every length inside the bracket;
- the FP32-storage form agrees with FP64 to FP32 rounding;
- the batched partition launch equals one evaluation per partition;
- large 61-state requests give the same bits on repeated evaluation.

Sizes: DNA GTR+I+G4 200 taxa x 100,000 sites; protein +G4 100 taxa x 20,000
sites; 500 partitions of 100-5,000 sites on one 64-taxon tree."""
import numpy as np
import pytest

import iqtree2_amd as ia
from iqtree2_amd import workloads as wl

pytestmark = pytest.mark.gpu


def dna_200x100k():
    return wl.dna_200x100k()


def check_tree_invariants(ctx, w, nbranch_checks, pulley_rtol):
    tree = ia.PhyloTree(ctx, w)
    try:
        lnl0, plh = tree.compute_likelihood(want_pattern_lh=True)
        assert np.isfinite(lnl0) and np.all(np.isfinite(plh))
        freq = w.get("freq")
        s = float(np.sum(freq * plh))
        assert abs(s - lnl0) <= 1e-12 * abs(lnl0)
        branches = tree.branches()
        pick = np.linspace(0, len(branches) - 1, nbranch_checks).astype(int)
        for k in pick:
            a, b = branches[k]
            lb = tree.likelihood_branch(a, b)
            assert abs(lb - lnl0) <= pulley_rtol * abs(lnl0), (a, b, lb, lnl0)
        # derivatives against central differences of the lnL at one branch
        a, b = branches[pick[len(pick) // 2]]
        t = tree.get_length(a, b)
        df, ddf = tree.likelihood_derv(a, b)
        h = 1e-4 * max(t, 1e-3)
        tree.set_length(a, b, t + h)
        lp = tree.likelihood_branch(a, b)
        tree.set_length(a, b, t - h)
        lm = tree.likelihood_branch(a, b)
        tree.set_length(a, b, t)
        l0 = tree.likelihood_branch(a, b)
        fd1 = (lp - lm) / (2 * h)
        fd2 = (lp - 2 * l0 + lm) / (h * h)
        assert abs(df - fd1) <= 1e-4 * max(1.0, abs(fd1)), (df, fd1)
        assert abs(ddf - fd2) <= 1e-2 * max(1.0, abs(fd2)), (ddf, fd2)
        return tree, lnl0
    except Exception:
        tree.close()
        raise


def test_dna_200x100k_invariants_and_sweep(ctx):
    w = dna_200x100k()
    tree, lnl0 = check_tree_invariants(ctx, w, nbranch_checks=8, pulley_rtol=1e-10)
    try:
        nb = len(tree.branches())
        lnl1, nd = tree.optimize_all_branches(iterations=1, tolerance=0.001, max_nr=100)
        assert lnl1 >= lnl0 - 1e-9 * abs(lnl0), (lnl0, lnl1)
        assert nd >= nb
        for a, b in tree.branches():
            assert 1e-6 <= tree.get_length(a, b) <= 10.0
        # the lnL the sweep reports is the tree lnL at the new lengths
        tree.clear_all_partial()
        assert abs(tree.compute_likelihood() - lnl1) <= 1e-10 * abs(lnl1)
    finally:
        tree.close()
    t32 = ia.PhyloTree(ctx, w, precision=ia.PREC_F32)
    try:
        l32 = t32.compute_likelihood()
        assert abs(l32 - lnl0) <= 1e-6 * abs(lnl0), (l32, lnl0)
    finally:
        t32.close()


def test_protein_100x20k_invariants(ctx):
    w = wl.protein_100x20k()
    tree, lnl0 = check_tree_invariants(ctx, w, nbranch_checks=6, pulley_rtol=1e-10)
    tree.close()
    t32 = ia.PhyloTree(ctx, w, precision=ia.PREC_F32)
    try:
        assert abs(t32.compute_likelihood() - lnl0) <= 1e-6 * abs(lnl0)
    finally:
        t32.close()


def test_codon_serving_requests_deterministic(ctx):
    """61-state requests of the serving workload (43 taxa x 34,991 patterns
    and 148 taxa x 27,677 patterns; each traversal step's output tile is
    written cooperatively by the workgroup, and the next step must read it
    only after a barrier): repeated evaluations give the same bits, FP64
    and FP32."""
    import bench
    reqs = bench.make_requests(ia, 62)
    for k in (29, 61):
        r = reqs[k]
        assert r.nstates == 61
        d = r.desc
        for prec in (ia.PREC_F64, ia.PREC_F32):
            inst = ia.Instance(ctx, d.ntaxa, d.nstates, d.ncat, d.npatterns, d.nslots, d.state_unknown,
                               prec, safe_numeric=True)
            try:
                inst.set_patterns(r.states, r.freq, r.invar)
                inst.set_model(*r.model)
                vals = [inst.likelihood_branch(r.ops, r.node, r.dad, r.length) for _ in range(3)]
                assert vals[0] == vals[1] == vals[2], (k, prec, vals)
            finally:
                inst.close()


def test_partitions_500_batched_equals_per_partition(ctx):
    parts = wl.partitions_500()
    plans = []
    for w in parts:
        planner = ia.PhyloTree(None, w)
        plans.append(planner.plan(*planner.branches()[0]))
    for prec, rtol in ((ia.PREC_F64, 1e-12), (ia.PREC_F32, 0.0)):
        insts = [ia.Instance.from_workload(ctx, w, precision=prec) for w in parts]
        try:
            jobs = [(inst, *pl) for inst, pl in zip(insts, plans)]
            plan = ia.BatchPlan(ctx, jobs)
            first = plan.run()
            again = plan.run()
            np.testing.assert_array_equal(first, again)
            single = np.array([inst.likelihood_branch(*pl) for inst, pl in zip(insts, plans)])
            if prec == ia.PREC_F64:
                np.testing.assert_allclose(first, single, rtol=rtol, atol=0)
                ref64 = first
            else:
                # FP32 partials: per tree within FP32 rounding of the FP64 lnL
                np.testing.assert_allclose(first, ref64, rtol=1e-5, atol=0)
                np.testing.assert_allclose(first, single, rtol=1e-12, atol=0)
        finally:
            for inst in insts:
                inst.close()
