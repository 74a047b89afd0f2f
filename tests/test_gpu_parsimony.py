"""Fitch parsimony on the device against the oracle: the weighted score and
the changes on the evaluated branch, equal (integer frequencies: exact sums)
on every branch from a full op list, for DNA (gaps and ambiguity codes),
protein and codon; the tree driver's incremental evaluation (only stale
slots recomputed) over every branch in order and after an NNI evaluation;
and the 200-taxon x 100,000-site DNA workload at full size."""
import numpy as np
import pytest

import iqtree2_amd as ia
from iqtree2_amd import workloads as wl
from oracle.binding import OracleTree

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("seq_type,ntaxa,nsite", [(ia.DNA, 40, 4000), (ia.PROTEIN, 30, 1500),
                                                  (ia.CODON, 12, 300), (ia.DNA, 7, 50)])
def test_parsimony_vs_oracle(ctx, seq_type, ntaxa, nsite):
    w = ia.Workload(seq_type=seq_type, ntaxa=ntaxa, nsite=nsite, ncat=4, mean_len=0.15, alpha=0.6,
                    gap_frac=0.02 if seq_type == ia.DNA else 0.0, seed=600 + ntaxa)
    print(f"parsimony seq_type={seq_type} ntaxa={ntaxa} nsite={nsite}", flush=True)
    inst = ia.Instance.from_workload(ctx, w, safe_numeric=seq_type == ia.CODON)
    orc = OracleTree.from_workload(w)
    planner = ia.PhyloTree(None, w)
    try:
        for a, b in planner.branches()[::3]:
            ops, node, dad = planner.plan_full(a, b)
            sg, bg = inst.parsimony_branch(ops, node, dad)
            so, bo = orc.parsimony_branch(ops, node, dad)
            assert (sg, bg) == (so, bo), (a, b)
        assert so > 0
    finally:
        inst.close()
        planner.close()


def test_parsimony_hbm_form(ctx):
    """The forms that keep the sets in HBM: a protein tree whose working set
    does not fit the LDS form (380 taxa), and an op list that writes a slot
    twice (run in its own order)."""
    w = ia.Workload(seq_type=ia.PROTEIN, ntaxa=380, nsite=300, ncat=4, mean_len=0.1, alpha=0.6,
                    seed=661)
    print("parsimony protein 380 taxa", flush=True)
    inst = ia.Instance.from_workload(ctx, w)
    orc = OracleTree.from_workload(w)
    planner = ia.PhyloTree(None, w)
    try:
        for a, b in planner.branches()[::40]:
            ops, node, dad = planner.plan_full(a, b)
            assert inst.parsimony_branch(ops, node, dad) == orc.parsimony_branch(ops, node, dad)
            twice = np.concatenate([ops, ops[-1:]])
            assert inst.parsimony_branch(twice, node, dad) == orc.parsimony_branch(twice, node, dad)
    finally:
        inst.close()
        planner.close()


@pytest.mark.parametrize("seq_type", [ia.DNA, ia.PROTEIN, ia.CODON])
def test_parsimony_driver_incremental(ctx, seq_type):
    """Every branch in turn through the tree driver, which recomputes only
    the slots whose direction changed, then again after an NNI evaluation
    (which swaps and restores the topology and drops every slot)."""
    w = ia.Workload(seq_type=seq_type, ntaxa=24 if seq_type != ia.CODON else 10,
                    nsite=2000 if seq_type != ia.CODON else 200, ncat=4, mean_len=0.12, alpha=0.6,
                    gap_frac=0.02, seed=640 + seq_type)
    tree = ia.PhyloTree(ctx, w, safe_numeric=seq_type == ia.CODON)
    orc = OracleTree.from_workload(w)
    try:
        branches = tree.branches()
        ref = {}
        for a, b in branches:
            ref[(a, b)] = orc.parsimony_branch(*tree.plan_full(a, b))
            assert tree.parsimony_branch(a, b) == ref[(a, b)], (a, b)
        for a, b in reversed(branches):
            assert tree.parsimony_branch(b, a) == ref[(a, b)], (b, a)
        tree.compute_likelihood()
        a, b = tree.inner_branches()[0]
        tree.nni_evaluate(a, b, 3)
        for a, b in branches[::4]:
            assert tree.parsimony_branch(a, b) == ref[(a, b)], (a, b)
    finally:
        tree.close()


def test_parsimony_full_size(ctx):
    w = wl.dna_200x100k()
    print("parsimony dna 200x100k", flush=True)
    tree = ia.PhyloTree(ctx, w)
    orc = OracleTree.from_workload(w, nthreads=8)
    try:
        lnl = tree.compute_likelihood()
        a, b = tree.branches()[0]
        score, sub = tree.parsimony_branch(a, b)
        assert (score, sub) == orc.parsimony_branch(*tree.plan_full(a, b))
        # the same tree score on another branch; the likelihood path untouched
        a2, b2 = tree.branches()[150]
        assert tree.parsimony_branch(a2, b2)[0] == score
        assert tree.compute_likelihood() == lnl
    finally:
        tree.close()
        w.close()
