"""Multifurcating nodes (the reference's degree > 3 branch,
tree/phylokernelnew.h:1405-1631: a node's partial is the product over all of
its children). The engine's ops are binary; a node of degree k > 3 is planned
as k - 2 binary ops joined by zero-length branches (transition matrix = the
identity), which is the same product. Pinned by brute-force summation over the
multifurcating tree itself: the tree lnL at every branch, for a degree-4 and a
degree-5 node, DNA and protein."""
import numpy as np
import pytest

import iqtree2_amd as ia
from refmath import brute_force_lnl, indicators

pytestmark = pytest.mark.gpu

# a 6-taxon binary tree: taxa 0-5, internal nodes 6-9
U = np.array([0, 1, 6, 2, 7, 3, 8, 4, 5], np.int32)
V = np.array([6, 6, 7, 7, 8, 8, 9, 9, 9], np.int32)


def contract(u, v, t, zero):
    """Edges of the tree with the zero-length edges contracted (their two
    ends become one node, the smaller id kept)."""
    rep = {}

    def find(x):
        while x in rep:
            x = rep[x]
        return x
    for k in zero:
        a, b = find(u[k]), find(v[k])
        rep[max(a, b)] = min(a, b)
    keep = [k for k in range(len(u)) if k not in zero]
    return (np.array([find(u[k]) for k in keep]), np.array([find(v[k]) for k in keep]),
            np.array([t[k] for k in keep]))


@pytest.mark.parametrize("seq_type,zero", [(ia.DNA, [4]), (ia.DNA, [4, 6]), (ia.PROTEIN, [4])])
def test_polytomy_lnl_by_brute_force(ctx, seq_type, zero):
    rng = np.random.default_rng(11 + len(zero) + seq_type)
    t = rng.uniform(0.05, 0.4, len(U))
    t[zero] = 0.0
    w = ia.Workload(seq_type=seq_type, ntaxa=6, nsite=300 if seq_type == ia.DNA else 120, ncat=4,
                    alpha=0.6, seed=5 + seq_type, edges=(U, V, t))
    print(f"polytomy seq_type={seq_type} zero={zero}", flush=True)
    pu, pv, pt = contract(U.tolist(), V.tolist(), t.tolist(), zero)
    degree = max(np.bincount(np.concatenate([pu, pv])))
    assert degree == 2 + len(zero) + 1
    states, freq = w.get("states"), w.get("freq")
    ref, _ = brute_force_lnl(6, (pu, pv, pt), states, w.get("Q"), w.get("pi"), w.get("rates"),
                             w.get("props"), 0.0, freq, indicators(seq_type, w.nstates, w.state_unknown))
    tree = ia.PhyloTree(ctx, w)
    try:
        for a, b in tree.branches():
            lnl = tree.likelihood_branch(a, b)
            assert abs(lnl - ref) <= 1e-10 * abs(ref), (a, b, lnl, ref)
    finally:
        tree.close()
        w.close()
