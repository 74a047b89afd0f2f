"""Sankoff parsimony on the device against the oracle (pinned by brute force
and, for the uniform cost matrix, by the Fitch score in
tests/test_oracle_sankoff.py): the tree score and the cost on the evaluated
branch, exactly, on every few branches, for DNA and protein, symmetric and
asymmetric cost matrices, gaps and ambiguity codes included; the uniform
matrix against the device's own Fitch kernel; the 200 x 100k workload."""
import numpy as np
import pytest

import iqtree2_amd as ia
from oracle.binding import OracleTree
from test_oracle_sankoff import random_cost

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("seq_type,ntaxa,nsite,symmetric", [(ia.DNA, 40, 3000, True), (ia.DNA, 24, 2000, False),
                                                             (ia.PROTEIN, 20, 1500, True),
                                                             (ia.PROTEIN, 12, 800, False),
                                                             (ia.CODON, 10, 300, True)])
def test_sankoff_vs_oracle(ctx, seq_type, ntaxa, nsite, symmetric):
    w = ia.Workload(seq_type=seq_type, ntaxa=ntaxa, nsite=nsite, ncat=4, mean_len=0.3, gap_frac=0.05,
                    seed=200 + ntaxa)
    cost = random_cost(w.nstates, np.random.default_rng(ntaxa), symmetric)
    inst = ia.Instance.from_workload(ctx, w, safe_numeric=seq_type == ia.CODON)   # 61 states are always safe
    inst.set_cost_matrix(cost)
    orc = OracleTree.from_workload(w)
    orc.set_cost_matrix(cost, seq_type)
    planner = ia.PhyloTree(None, w)
    try:
        for a, b in planner.branches()[::4]:
            ops, node, dad = planner.plan_full(a, b)
            sg, bg = inst.parsimony_sankoff(ops, node, dad)
            so, bo = orc.parsimony_sankoff(ops, node, dad)
            assert sg == so and bg == bo, (a, b, sg, so, bg, bo)
        if symmetric:
            # the uniform matrix gives the Fitch score of the device's Fitch kernel
            inst.set_cost_matrix((1 - np.eye(w.nstates)).astype(np.uint32))
            ops, node, dad = planner.plan_full(*planner.branches()[0])
            assert inst.parsimony_sankoff(ops, node, dad)[0] == inst.parsimony_branch(ops, node, dad)[0]
    finally:
        inst.close()
        planner.close()


def test_sankoff_full_size(ctx):
    from iqtree2_amd import workloads as wl
    w = wl.dna_200x100k()
    cost = random_cost(4, np.random.default_rng(1))
    inst = ia.Instance.from_workload(ctx, w)
    inst.set_cost_matrix(cost)
    orc = OracleTree.from_workload(w, nthreads=8)
    orc.set_cost_matrix(cost, ia.DNA)
    planner = ia.PhyloTree(None, w)
    try:
        ops, node, dad = planner.plan_full(*planner.branches()[0])
        assert inst.parsimony_sankoff(ops, node, dad) == orc.parsimony_sankoff(ops, node, dad)
    finally:
        inst.close()
        planner.close()
        w.close()
