"""The planning of multifurcating nodes (a node of degree k > 3 as k - 2
binary steps joined by zero-length branches) checked in the oracle against
brute-force summation over the multifurcating tree itself: the tree lnL at
every branch, degree 4 and 5, DNA and protein. The device form of the same
check is tests/test_gpu_polytomy.py."""
import numpy as np
import pytest

import iqtree2_amd as ia
from oracle.binding import OracleTree
from refmath import brute_force_lnl, indicators
from test_gpu_polytomy import U, V, contract


@pytest.mark.parametrize("seq_type,zero", [(ia.DNA, [4]), (ia.DNA, [4, 6]), (ia.PROTEIN, [4])])
def test_polytomy_planning_in_the_oracle(seq_type, zero):
    rng = np.random.default_rng(11 + len(zero) + seq_type)
    t = rng.uniform(0.05, 0.4, len(U))
    t[zero] = 0.0
    w = ia.Workload(seq_type=seq_type, ntaxa=6, nsite=300 if seq_type == ia.DNA else 120, ncat=4,
                    alpha=0.6, seed=5 + seq_type, edges=(U, V, t))
    pu, pv, pt = contract(U.tolist(), V.tolist(), t.tolist(), zero)
    ref, _ = brute_force_lnl(6, (pu, pv, pt), w.get("states"), w.get("Q"), w.get("pi"), w.get("rates"),
                             w.get("props"), 0.0, w.get("freq"),
                             indicators(seq_type, w.nstates, w.state_unknown))
    orc = OracleTree.from_workload(w)
    planner = ia.PhyloTree(None, w)
    for a, b in planner.branches():
        lnl = orc.likelihood_branch(*planner.plan(a, b))
        assert abs(lnl - ref) <= 1e-10 * abs(ref), (a, b, lnl, ref)
    # contracting the zero-length edges changes the tree: a positive length
    # there gives another likelihood
    t2 = t.copy()
    t2[zero] = 0.2
    w2 = ia.Workload(seq_type=seq_type, ntaxa=6, nsite=300 if seq_type == ia.DNA else 120, ncat=4,
                     alpha=0.6, seed=5 + seq_type, edges=(U, V, t2))
    orc2 = OracleTree.from_arrays(6, w.nstates, 4, w.state_unknown, w.get("states"), w.get("freq"),
                                  w.get("invar"), w2.model_arrays())
    assert abs(orc2.likelihood_branch(*ia.PhyloTree(None, w2).plan(*planner.branches()[0])) - ref) > 1e-3
