"""NNI evaluation (getBestNNIForBran, tree/phylotree.cpp:4069-4300) on the
GPU against the oracle.

The reference's procedure for one inner branch (node1, node2), with its
defaults nni5 = true, nni5_num_eval = 1 and NNI_MAX_NR_STEP = 10
(utils/tools.cpp:1155-1156, utils/pllnni.cpp:42): node1's and node2's
partials move to the two nni_partial_lh buffers; for each of the two moves
(node1's first neighbour exchanged with node2's first, then its second)
the partials seen across the central branch are cleared, each of node1's
other branches is optimised (optimizeOneBranch with clearLH = false) after
clearing node1's partial seen from it, then the central branch, then each
of node2's other branches; the score is computeLikelihoodFromBuffer; the
move is undone but the lengths carry over to the next move; at the end the
topology, lengths and buffers are restored.

replay() below performs that procedure through the planner's primitives
(plkh_tree_nni_begin / _swap / clear_partial / plan / _nni_end) and runs
every Newton optimisation on the engine (plk_optimize_branch) and on the
oracle with the same op lists, slot numbers included. The driver's own
forms - one engine call per evaluation (plkh_tree_nni_evaluate) and every
inner branch in one device sweep (plkh_tree_nni_evaluate_all) - are then
compared with the oracle's replay: lengths within 1e-7, evaluation counts
equal, lnL within 1e-9 relative."""
import numpy as np
import pytest

import iqtree2_amd as ia
from oracle.binding import OracleTree

pytestmark = pytest.mark.gpu

NNI_MAX_NR_STEP = 10


def replay(planner, run_branch, from_buffer, node1, node2):
    planner.nni_begin(node1, node2)
    n1 = [x for x in planner.neighbors(node1) if x != node2]
    n2 = [y for y in planner.neighbors(node2) if y != node1]
    out = []

    def optimize(u, v):
        # optimizeOneBranch(u, v, clearLH = false): theta across (v, u)
        ops, node, dad, length = planner.plan(v, u)
        x, nd = run_branch(ops, node, dad, length)
        planner.set_length_keep(u, v, x)
        return x, nd

    for cnt in range(2):
        a, c = n1[0], n2[cnt]
        planner.nni_swap(node1, a, node2, c)
        planner.clear_partial(node2, node1)
        planner.clear_partial(node1, node2)
        lens, nds = [0.0] * 5, [0] * 5
        i = 1
        for x in [v for v in planner.neighbors(node1) if v != node2]:
            planner.clear_partial(node1, x)
            lens[i], nds[i] = optimize(node1, x)
            i += 1
        planner.clear_partial(node1, node2)
        lens[0], nds[0] = optimize(node1, node2)
        last = None
        for y in [v for v in planner.neighbors(node2) if v != node1]:
            planner.clear_partial(node2, y)
            lens[i], nds[i] = optimize(node2, y)
            i += 1
            last = y
        planner.clear_partial(node2, node1)
        score = from_buffer(planner.get_length(node2, last))
        planner.nni_swap(node1, c, node2, a)
        out.append((score, lens, nds))
    planner.nni_end()
    return out


def oracle_replay_all(w, orc, inner):
    planner = ia.PhyloTree(None, w)

    def run(ops, node, dad, length):
        x, _, nd, st = orc.optimize_branch(ops, node, dad, length, 1e-6, 10.0, NNI_MAX_NR_STEP)
        assert st == 0
        return x, nd

    res = [replay(planner, run, orc.likelihood_from_buffer, a, b) for a, b in inner]
    planner.close()
    return res


@pytest.mark.parametrize("seq_type", [ia.DNA, ia.PROTEIN])
def test_nni_replay_engine_and_oracle(ctx, seq_type):
    """The procedure through the C ABI: engine and oracle in lock step."""
    w = ia.Workload(seq_type=seq_type, ntaxa=24, nsite=1500, ncat=4, mean_len=0.15, alpha=0.7,
                    p_inv=0.1 if seq_type == ia.DNA else 0.0, seed=61 + seq_type)
    planner = ia.PhyloTree(None, w)
    inst = ia.Instance.from_workload(ctx, w)
    orc = OracleTree.from_workload(w)
    nslots = w.ntaxa - 2
    seen_nni_slot = set()

    def run(ops, node, dad, length):
        for op in ops:
            if op["dad_slot"] >= nslots:
                seen_nni_slot.add(int(op["dad_slot"]))
        xg, lg, ng = inst.optimize_branch(ops, node, dad, length, 1e-6, 10.0, NNI_MAX_NR_STEP)
        xo, lo, no, st = orc.optimize_branch(ops, node, dad, length, 1e-6, 10.0, NNI_MAX_NR_STEP)
        assert st == 0 and ng == no, (ng, no)
        assert abs(xg - xo) <= 1e-7 * max(1.0, xo), (xg, xo)
        assert abs(lg - lo) <= 1e-9 * abs(lo)
        return xo, no

    def from_buffer(t):
        lg, lo = inst.likelihood_from_buffer(t), orc.likelihood_from_buffer(t)
        assert abs(lg - lo) <= 1e-9 * abs(lo)
        return lo

    try:
        inner = planner.inner_branches()
        lens0 = {(a, b): planner.get_length(a, b) for a, b in planner.branches()}
        for node1, node2 in inner[::3]:
            out = replay(planner, run, from_buffer, node1, node2)
            assert all(np.isfinite(s) for s, _, _ in out)
        assert seen_nni_slot == {nslots, nslots + 1}
        # topology and lengths restored
        assert {(a, b): planner.get_length(a, b) for a, b in planner.branches()} == lens0
    finally:
        inst.close()


@pytest.mark.parametrize("seq_type", [ia.DNA, ia.PROTEIN, ia.CODON])
def test_nni_driver_host_and_device_vs_oracle(ctx, seq_type):
    ntaxa, nsite = {ia.DNA: (40, 2500), ia.PROTEIN: (40, 800), ia.CODON: (14, 300)}[seq_type]
    w = ia.Workload(seq_type=seq_type, ntaxa=ntaxa, nsite=nsite, ncat=4, mean_len=0.12, alpha=0.6,
                    seed=71 + seq_type)
    print(f"nni driver seq_type={seq_type}", flush=True)
    tree = ia.PhyloTree(ctx, w)
    try:
        inner = tree.inner_branches()
        assert len(inner) == w.ntaxa - 3
        host = [tree.nni_evaluate(a, b, NNI_MAX_NR_STEP) for a, b in inner]
        dev_lnl, dev_lens, dev_nd = tree.nni_evaluate_all(NNI_MAX_NR_STEP)
        # the tree is unchanged afterwards: its lnL equals a fresh evaluation.
        # Partials the sweep computed away from the two NNI nodes stay cached;
        # the sweep's per-thread arithmetic and the protein traversal's
        # matrix-core tile round them differently, so the two agree to
        # rounding, not bit for bit.
        lnl_after = tree.compute_likelihood()
        tree.clear_all_partial()
        lnl_fresh = tree.compute_likelihood()
        assert abs(lnl_fresh - lnl_after) <= 1e-12 * abs(lnl_fresh)
    finally:
        tree.close()
    orc = OracleTree.from_workload(w, safe_numeric=seq_type == ia.CODON)
    ref = oracle_replay_all(w, orc, inner)
    for j, ((lh, Lh, nh), r) in enumerate(zip(host, ref)):
        for cnt in range(2):
            score, lens, nds = r[cnt]
            np.testing.assert_array_equal(nh[cnt], nds, err_msg=f"host counts, branch {j}")
            np.testing.assert_array_equal(dev_nd[j, cnt], nds, err_msg=f"device counts, branch {j}")
            np.testing.assert_allclose(Lh[cnt], lens, rtol=1e-7, atol=1e-12)
            np.testing.assert_allclose(dev_lens[j, cnt], lens, rtol=1e-7, atol=1e-12)
            assert abs(lh[cnt] - score) <= 1e-9 * abs(score)
            assert abs(dev_lnl[j, cnt] - score) <= 1e-9 * abs(score)
