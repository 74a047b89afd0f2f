"""The benchmark configurations at full size, GPU engine against the oracle.

Tolerances as for the small cases (FP64 form): tree lnL 1e-9 relative,
pattern lnL 1e-9 absolute, derivatives 1e-7 relative, scale counters
bit-exact, partials 1e-10 relative to their pattern's largest entry, Newton
lengths 1e-7 with identical evaluation counts.

- DNA GTR+I+G4, 200 taxa x 100,000 sites: full traversal (every one of the
  198 partial slots), derivatives on ten branches, one optimizeAllBranches
  pass on the device against the oracle's branch-by-branch replay;
- protein +G4, 100 taxa x 20,000 sites, Exp(mean 0.2) lengths: full
  traversal (98 slots) and derivatives;
- 500 partitions of 100-5,000 sites on one 64-taxon tree: every
  partition's batched FP64 lnL.

The oracle runs on up to 16 host threads; each configuration costs it a few
seconds."""
import os

import numpy as np
import pytest

import iqtree2_amd as ia
from iqtree2_amd import workloads as wl
from oracle.binding import OracleTree

pytestmark = pytest.mark.gpu

LNL_RTOL = 1e-9
DERV_RTOL = 1e-7
NT = max(1, min(16, os.cpu_count() or 1))


def traversal_parity(ctx, w, nslots_partials, nderv):
    planner = ia.PhyloTree(None, w)
    inst = ia.Instance.from_workload(ctx, w)
    orc = OracleTree.from_workload(w, nthreads=NT)
    try:
        branches = planner.branches()
        a, b = branches[0]
        ops, node, dad, length = planner.plan(a, b)
        assert len(ops) == w.ntaxa - 2
        lg, pg = inst.likelihood_branch(ops, node, dad, length, want_pattern_lh=True)
        lo, po = orc.likelihood_branch(ops, node, dad, length, want_pattern_lh=True)
        print(f"[{w.ntaxa}x{w.nsite}] lnL gpu {lg:.10f} oracle {lo:.10f}", flush=True)
        assert abs(lg - lo) <= LNL_RTOL * abs(lo)
        np.testing.assert_allclose(pg, po, rtol=0, atol=1e-9)
        # scale counters of every slot, bit for bit
        scaled = 0
        for op in ops:
            slot = int(op["dad_slot"])
            sg = inst.read_partial(slot, partial=False)[1]
            so = orc.read_partial(slot, partial=False)[1]
            bad = np.flatnonzero(sg != so)
            assert bad.size == 0, f"slot {slot}: counters differ at {bad[:10]}"
            scaled += int(so.sum() > 0)
        # partials of a spread of slots, post-order position first to last
        for k in np.linspace(0, len(ops) - 1, nslots_partials).astype(int):
            slot = int(ops[k]["dad_slot"])
            pgp = inst.read_partial(slot)[0]
            pop = orc.read_partial(slot)[0]
            denom = np.maximum(np.abs(pop).max(axis=1, keepdims=True), 1e-300)
            err = float(np.max(np.abs(pgp - pop) / denom))
            assert err <= 1e-10, f"slot {slot}: partial error {err}"
        # derivatives and lnL from the theta buffer on a spread of branches
        for k in np.linspace(0, len(branches) - 1, nderv).astype(int):
            a, b = branches[k]
            ops, node, dad, length = planner.plan(a, b)
            inst.likelihood_buffer(ops, node, dad)
            orc.likelihood_buffer(ops, node, dad)
            for t in (length, 1.7 * length + 1e-3):
                dg, ddg, lng = inst.likelihood_derv(t)
                do, ddo = orc.likelihood_derv(t)
                assert abs(dg - do) <= DERV_RTOL * max(1.0, abs(do)), (a, b, t, dg, do)
                assert abs(ddg - ddo) <= DERV_RTOL * max(1.0, abs(ddo)), (a, b, t, ddg, ddo)
                lfb = orc.likelihood_from_buffer(t)
                assert abs(lng - lfb) <= LNL_RTOL * abs(lfb)
        return scaled
    finally:
        inst.close()


def test_dna_200x100k_traversal_vs_oracle(ctx):
    w = wl.dna_200x100k()
    traversal_parity(ctx, w, nslots_partials=12, nderv=10)


def test_dna_200x100k_newton_pass_vs_oracle(ctx):
    """One optimizeAllBranches pass: the device-resident sweep against the
    oracle's optimizeOneBranch replayed branch by branch on the same
    pre-order (every length within 1e-7, every evaluation count equal)."""
    w = wl.dna_200x100k()
    tree = ia.PhyloTree(ctx, w)
    try:
        lnl_dev, nd_dev = tree.optimize_all_branches(iterations=1, tolerance=1e-3, max_nr=100)
        counts_dev = tree.pass_counts()
        branches = tree.branches()
        lens_dev = np.array([tree.get_length(a, b) for a, b in branches])
    finally:
        tree.close()
    planner = ia.PhyloTree(None, w)
    orc = OracleTree.from_workload(w, nthreads=NT)
    a0, b0 = branches[0]
    ops, node, dad, length = planner.plan(b0, a0)
    lnl0 = orc.likelihood_branch(ops, node, dad, length)
    lens_orc, counts_orc = [], []
    for a, b in branches:
        ops, node, dad, length = planner.plan(b, a)
        x, _, nd, st = orc.optimize_branch(ops, node, dad, length)
        assert st == 0
        planner.set_length(a, b, x)
        lens_orc.append(x)
        counts_orc.append(nd)
    lnl_orc = orc.likelihood_from_buffer(lens_orc[-1])
    lens_orc = np.array(lens_orc)
    print(f"newton pass: lnL {lnl0:.6f} -> gpu {lnl_dev:.10f} oracle {lnl_orc:.10f}; "
          f"evaluations gpu {int(np.sum(counts_dev))} oracle {int(np.sum(counts_orc))}", flush=True)
    np.testing.assert_array_equal(counts_dev, np.array(counts_orc, np.int32))
    assert nd_dev == int(np.sum(counts_orc))
    np.testing.assert_allclose(lens_dev, lens_orc, rtol=1e-7, atol=1e-12)
    assert abs(lnl_dev - lnl_orc) <= LNL_RTOL * abs(lnl_orc)


def test_protein_100x20k_vs_oracle(ctx):
    w = wl.protein_100x20k()
    traversal_parity(ctx, w, nslots_partials=10, nderv=5)


def test_partitions_500_batched_vs_oracle(ctx):
    parts = wl.partitions_500()
    jobs, orcs = [], []
    insts = []
    try:
        for w in parts:
            planner = ia.PhyloTree(None, w)
            a, b = planner.branches()[0]
            plan = planner.plan(a, b)
            inst = ia.Instance.from_workload(ctx, w)
            insts.append(inst)
            jobs.append((inst, *plan))
        got = ia.likelihood_branch_batch(ctx, jobs)
        ref = np.zeros(len(parts))
        for k, w in enumerate(parts):
            o = OracleTree.from_workload(w, nthreads=NT)
            ref[k] = o.likelihood_branch(*jobs[k][1:])
            if k % 25 == 0:
                orcs.append((k, o))   # kept for the derivative checks below
        rel = np.abs(got - ref) / np.abs(ref)
        print(f"500 partitions: max relative |dlnL| {rel.max():.3e}", flush=True)
        assert np.all(rel <= LNL_RTOL), np.argsort(-rel)[:5]
        # the theta form of the same launch, then each item's derivatives
        modes = np.ones(len(jobs), np.int32)
        ia.batch_evaluate(ctx, jobs, modes)
        for k, o in orcs:
            inst, ops, node, dad, length = jobs[k]
            o.likelihood_buffer(None, node, dad)
            dg, ddg, _ = inst.likelihood_derv(length)
            do, ddo = o.likelihood_derv(length)
            assert abs(dg - do) <= DERV_RTOL * max(1.0, abs(do)), (k, dg, do)
            assert abs(ddg - ddo) <= DERV_RTOL * max(1.0, abs(ddo)), (k, ddg, ddo)
    finally:
        for inst in insts:
            inst.close()
