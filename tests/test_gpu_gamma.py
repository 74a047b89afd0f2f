"""ModelFinder's parameter loop for the Gamma shape (RateGamma::
optimizeParameters, model/rategamma.cpp:232-246) with each likelihood a full
traversal on the device, against the oracle driving the restated Brent
(oracle/brent.py) with its own likelihood: the same shape, lnL and number of
evaluations; and a shape set directly against the oracle at that shape."""
import numpy as np
import pytest

import iqtree2_amd as ia
from oracle import brent
from oracle.binding import OracleTree

pytestmark = pytest.mark.gpu


def oracle_neg_lnl(w, tree, p_inv):
    orc = OracleTree.from_workload(w, nthreads=4)
    ops, node, dad, length = tree.plan(*tree.branches()[0])
    m = w.model_arrays()

    def f(alpha):
        rates, props = ia.gamma_rates(w.ncat, alpha, p_inv)
        orc.set_model(m["eval"], m["evec"], m["inv_evec"], m["tip"], rates, props)
        return -orc.likelihood_branch(ops, node, dad, length)
    return f


@pytest.mark.parametrize("seq_type,ntaxa,nsite,alpha,p_inv", [(ia.DNA, 24, 3000, 0.6, 0.0),
                                                               (ia.DNA, 16, 2000, 1.5, 0.2),
                                                               (ia.PROTEIN, 16, 800, 0.8, 0.0)])
def test_gamma_shape_vs_oracle(ctx, seq_type, ntaxa, nsite, alpha, p_inv):
    w = ia.Workload(seq_type=seq_type, ntaxa=ntaxa, nsite=nsite, ncat=4, mean_len=0.15, alpha=alpha,
                    p_inv=p_inv, seed=900 + ntaxa)
    print(f"gamma shape seq_type={seq_type} ntaxa={ntaxa}", flush=True)
    planner = ia.PhyloTree(None, w)
    tree = ia.PhyloTree(ctx, w)
    try:
        f = oracle_neg_lnl(w, planner, p_inv)
        # a shape set directly
        tree.set_gamma_shape(1.7)
        lg = tree.compute_likelihood()
        assert abs(lg + f(1.7)) <= 1e-9 * abs(lg)
        tree.set_gamma_shape(alpha)
        # the optimisation from the generating shape
        ag, lg, ng = tree.optimize_gamma_shape()
        ao, lo, no = brent.optimize_gamma_shape(f, alpha)
        assert ng == no
        assert abs(ag - ao) <= 1e-9 * ao
        assert abs(lg - lo) <= 1e-9 * abs(lo)
        assert lg >= tree.compute_likelihood() - 1e-9 * abs(lg)
    finally:
        tree.close()
        planner.close()


def test_models_batch_equals_alone(ctx):
    """Every tree of the batched form (one traversal launch per round for
    all trees still optimising) ends where it ends alone, bit for bit."""
    specs = [(ia.DNA, 20, 2000, 0.5, 0.0, 31), (ia.DNA, 30, 1500, 2.0, 0.1, 32),
             (ia.PROTEIN, 12, 600, 0.9, 0.0, 33), (ia.DNA, 16, 900, 0.3, 0.0, 34)]
    print("gamma shape batch", flush=True)
    ws = [ia.Workload(seq_type=s, ntaxa=n, nsite=m, ncat=4, mean_len=0.12, alpha=a, p_inv=p, seed=sd)
          for s, n, m, a, p, sd in specs]
    alone, batch = [ia.PhyloTree(ctx, w) for w in ws], [ia.PhyloTree(ctx, w) for w in ws]
    try:
        ref = [t.optimize_gamma_shape() for t in alone]
        a, lnl, ne, st = ia.models_optimize_gamma_shape(batch)
        assert list(st) == [0] * len(ws)
        for k, (ra, rl, rn) in enumerate(ref):
            assert (a[k], lnl[k], ne[k]) == (ra, rl, rn), k
            assert batch[k].compute_likelihood() == alone[k].compute_likelihood()
    finally:
        for t in alone + batch:
            t.close()
        for w in ws:
            w.close()
