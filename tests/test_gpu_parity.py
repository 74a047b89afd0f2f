"""GPU parity: the HIP engine (through the C-ABI) against the CPU oracle on
identical inputs. Tolerances (FP64 form): lnL 1e-9 relative, per-pattern
lnL 1e-9 absolute, derivatives 1e-7 relative, scale counters bit-exact,
partials 1e-10 relative to their pattern's largest entry."""
import numpy as np
import pytest

import iqtree2_amd as ia
from oracle.binding import OracleTree

pytestmark = pytest.mark.gpu

LNL_RTOL = 1e-9
DERV_RTOL = 1e-7


def setup(ctx, w, precision=ia.PREC_F64, safe=False, nthreads=4):
    planner = ia.PhyloTree(None, w)
    inst = ia.Instance.from_workload(ctx, w, precision=precision, safe_numeric=safe)
    orc = OracleTree.from_workload(w, safe_numeric=safe, nthreads=nthreads)
    return planner, inst, orc


def compare_partials(inst, orc, ops, tol=1e-10):
    for op in ops:
        slot = int(op["dad_slot"])
        pg, sg = inst.read_partial(slot)
        po, so = orc.read_partial(slot)
        mism = np.flatnonzero((sg != so).reshape(sg.shape[0], -1).any(axis=1))
        assert mism.size == 0, f"slot {slot}: scale counters differ at patterns {mism[:10]}"
        denom = np.maximum(np.abs(po).max(axis=1, keepdims=True), 1e-300)
        err = np.max(np.abs(pg - po) / denom)
        assert err <= tol, f"slot {slot}: partial error {err}"


def all_branch_parity(w, planner, inst, orc, check_derv=True, stride=1):
    for a, b in planner.branches()[::stride]:
        ops, node, dad, length = planner.plan(a, b)
        lg, pg = inst.likelihood_branch(ops, node, dad, length, want_pattern_lh=True)
        lo, po = orc.likelihood_branch(ops, node, dad, length, want_pattern_lh=True)
        assert abs(lg - lo) <= LNL_RTOL * abs(lo), (a, b, lg, lo)
        np.testing.assert_allclose(pg, po, rtol=0, atol=1e-9)
        compare_partials(inst, orc, ops)
        if check_derv:
            inst.likelihood_buffer(None, node, dad)
            orc.likelihood_buffer(None, node, dad)
            for t in (length, 2.0 * length + 1e-3):
                dg, ddg, lng = inst.likelihood_derv(t)
                do, ddo = orc.likelihood_derv(t)
                assert abs(dg - do) <= DERV_RTOL * max(1.0, abs(do)), (a, b, t, dg, do)
                assert abs(ddg - ddo) <= DERV_RTOL * max(1.0, abs(ddo)), (a, b, t, ddg, ddo)
                lfb = orc.likelihood_from_buffer(t)
                assert abs(lng - lfb) <= LNL_RTOL * abs(lfb)
                assert abs(inst.likelihood_from_buffer(t) - lfb) <= LNL_RTOL * abs(lfb)


def test_config1_dna_gtr_g4_all_branches(ctx):
    """16 taxa x 1,000 sites, GTR+G4, lnL and df/ddf on every branch."""
    w = ia.Workload(seq_type=ia.DNA, ntaxa=16, nsite=1000, ncat=4, mean_len=0.1, alpha=0.5,
                    seed=20261203)
    all_branch_parity(w, *setup(ctx, w))


def test_dna_gtr_i_g4_gaps(ctx):
    w = ia.Workload(seq_type=ia.DNA, ntaxa=60, nsite=5000, ncat=4, mean_len=0.1, alpha=0.5,
                    p_inv=0.2, gap_frac=0.02, seed=20261204)
    all_branch_parity(w, *setup(ctx, w), stride=7)


def test_protein_g4(ctx):
    w = ia.Workload(seq_type=ia.PROTEIN, ntaxa=30, nsite=2000, ncat=4, mean_len=0.2, alpha=0.8,
                    seed=20261205)
    all_branch_parity(w, *setup(ctx, w), stride=5)


@pytest.mark.parametrize("safe", [False, True])
def test_protein_scaling_counts_bit_exact(ctx, safe):
    """Protein partials on the matrix cores cross 2^-256 (150 taxa, long
    branches): counters bit-exact in normal (per pattern) and safe (per
    category) mode."""
    w = ia.Workload(seq_type=ia.PROTEIN, ntaxa=150, nsite=400, ncat=4, mean_len=0.5, alpha=1.0,
                    seed=17)
    planner, inst, orc = setup(ctx, w, safe=safe)
    a, b = planner.branches()[0]
    ops, node, dad, length = planner.plan(a, b)
    lg = inst.likelihood_branch(ops, node, dad, length)
    lo = orc.likelihood_branch(ops, node, dad, length)
    assert abs(lg - lo) <= LNL_RTOL * abs(lo)
    assert max(orc.read_partial(int(op["dad_slot"]))[1].max() for op in ops) >= 1
    compare_partials(inst, orc, ops, tol=1e-9)
    inst.likelihood_buffer(None, node, dad)
    orc.likelihood_buffer(None, node, dad)
    dg, ddg, _ = inst.likelihood_derv(length)
    do, ddo = orc.likelihood_derv(length)
    assert abs(dg - do) <= DERV_RTOL * max(1.0, abs(do))
    assert abs(ddg - ddo) <= DERV_RTOL * max(1.0, abs(ddo))


@pytest.mark.parametrize("ncat", [1, 2, 8])
def test_protein_category_counts(ctx, ncat):
    w = ia.Workload(seq_type=ia.PROTEIN, ntaxa=14, nsite=700, ncat=ncat, mean_len=0.2, alpha=0.8,
                    seed=40 + ncat)
    all_branch_parity(w, *setup(ctx, w), stride=5)


def test_protein_fp32_storage(ctx):
    w = ia.Workload(seq_type=ia.PROTEIN, ntaxa=40, nsite=3000, ncat=4, mean_len=0.2, alpha=0.8,
                    seed=20261208)
    planner, inst32, orc = setup(ctx, w, precision=ia.PREC_F32)
    for a, b in planner.branches()[::9]:
        ops, node, dad, length = planner.plan(a, b)
        lg = inst32.likelihood_branch(ops, node, dad, length)
        lo = orc.likelihood_branch(ops, node, dad, length)
        assert abs(lg - lo) <= 1e-6 * abs(lo), (a, b, lg, lo, abs(lg - lo))


def test_codon_gy_g4(ctx):
    """61-state GY94+I+G4 on the matrix-core path (per-category scaling)."""
    w = ia.Workload(seq_type=ia.CODON, ntaxa=14, nsite=800, ncat=4, mean_len=0.1, alpha=0.5,
                    p_inv=0.1, seed=20261206)
    all_branch_parity(w, *setup(ctx, w, safe=True), stride=3)


@pytest.mark.parametrize("ncat", [1, 2, 8])
def test_codon_category_counts(ctx, ncat):
    w = ia.Workload(seq_type=ia.CODON, ntaxa=9, nsite=400, ncat=ncat, mean_len=0.2, alpha=0.7,
                    seed=60 + ncat)
    all_branch_parity(w, *setup(ctx, w, safe=True), stride=4)


def test_codon_scaling_counts_bit_exact(ctx):
    """Long branches on 120 taxa: codon partials cross 2^-256 per category.
    Partials are compared at 1e-7 of their pattern's largest entry: in the
    slowest rate category the eigen-basis entries of such deep subtrees are
    sums of terms that cancel to ~1e-9 of their size, and an 80-bit
    recomputation of the worst pattern puts both this engine (7e-10) and the
    oracle (2.2e-9) that far from the exact value. Counters stay bit-exact
    and lnL stays within 1e-9."""
    w = ia.Workload(seq_type=ia.CODON, ntaxa=120, nsite=300, ncat=4, mean_len=0.5, alpha=1.0,
                    seed=7)
    planner, inst, orc = setup(ctx, w, safe=True)
    a, b = planner.branches()[0]
    ops, node, dad, length = planner.plan(a, b)
    lg = inst.likelihood_branch(ops, node, dad, length)
    lo = orc.likelihood_branch(ops, node, dad, length)
    assert abs(lg - lo) <= LNL_RTOL * abs(lo)
    assert max(orc.read_partial(int(op["dad_slot"]))[1].max() for op in ops) >= 1
    compare_partials(inst, orc, ops, tol=1e-7)
    inst.likelihood_buffer(None, node, dad)
    orc.likelihood_buffer(None, node, dad)
    dg, ddg, _ = inst.likelihood_derv(length)
    do, ddo = orc.likelihood_derv(length)
    assert abs(dg - do) <= DERV_RTOL * max(1.0, abs(do))
    assert abs(ddg - ddo) <= DERV_RTOL * max(1.0, abs(ddo))


def test_codon_fp32_storage(ctx):
    w = ia.Workload(seq_type=ia.CODON, ntaxa=20, nsite=600, ncat=4, mean_len=0.15, alpha=0.5,
                    seed=20261207)
    planner, inst32, orc = setup(ctx, w, precision=ia.PREC_F32, safe=True)
    for a, b in planner.branches()[::5]:
        ops, node, dad, length = planner.plan(a, b)
        lg = inst32.likelihood_branch(ops, node, dad, length)
        lo = orc.likelihood_branch(ops, node, dad, length)
        assert abs(lg - lo) <= 1e-6 * abs(lo), (a, b, lg, lo, abs(lg - lo))


@pytest.mark.parametrize("ncat", [1, 4])
def test_codon_optimize_all_branches(ctx, ncat):
    """The driver runs the device sweep for 61 states; it matches the
    oracle replaying optimizeAllBranches (tree/phylotree.cpp:2698-2773) with
    the same plans, including the revert of every length when the pass
    lowered the lnL. With ncat = 4 the reference's derivative double-counts
    one theta term (dotProductTriple with an odd state count and an even
    block), so its Newton steps can lower the lnL and the pass is reverted."""
    w = ia.Workload(seq_type=ia.CODON, ntaxa=10, nsite=500, ncat=ncat, mean_len=0.1, alpha=0.5,
                    seed=31)
    tolerance = 1e-3
    gpu_tree = ia.PhyloTree(ctx, w)
    lnl0 = gpu_tree.compute_likelihood()
    lnl, nd = gpu_tree.optimize_all_branches(iterations=1, tolerance=tolerance, max_nr=100)
    assert nd > 0
    planner = ia.PhyloTree(None, w)
    orc = OracleTree.from_workload(w, safe_numeric=True, nthreads=4)
    branches = planner.branches()
    initial = {br: planner.get_length(*br) for br in branches}
    a0, b0 = branches[0]
    ops, node, dad, length = planner.plan(b0, a0)
    lo0 = orc.likelihood_branch(ops, node, dad, length)
    assert abs(lnl0 - lo0) <= LNL_RTOL * abs(lo0)
    for a, b in branches:
        ops, node, dad, length = planner.plan(b, a)
        x, _, _, st = orc.optimize_branch(ops, node, dad, length)
        assert st == 0
        planner.set_length(a, b, x)
    lo = orc.likelihood_from_buffer(planner.get_length(*branches[-1]))
    reverted = lo < lo0 - tolerance * 0.1
    if reverted:
        expect_len, expect_lnl = initial, lo0
    else:
        expect_len, expect_lnl = {br: planner.get_length(*br) for br in branches}, lo
    for br in branches:
        x = expect_len[br]
        assert abs(gpu_tree.get_length(*br) - x) <= 1e-7 * max(x, 1e-6), (br, reverted)
    assert abs(lnl - expect_lnl) <= LNL_RTOL * abs(expect_lnl), (lnl, expect_lnl, reverted)


@pytest.mark.parametrize("safe", [False, True])
def test_scaling_counts_bit_exact(ctx, safe):
    """Deep tree with long branches: partials cross 2^-256 and get rescaled."""
    w = ia.Workload(seq_type=ia.DNA, ntaxa=300, nsite=600, ncat=4, mean_len=0.6, alpha=1.0,
                    seed=5)
    planner, inst, orc = setup(ctx, w, safe=safe)
    a, b = planner.branches()[0]
    ops, node, dad, length = planner.plan(a, b)
    lg = inst.likelihood_branch(ops, node, dad, length)
    lo = orc.likelihood_branch(ops, node, dad, length)
    assert abs(lg - lo) <= LNL_RTOL * abs(lo)
    assert max(orc.read_partial(int(op["dad_slot"]))[1].max() for op in ops) >= 1
    compare_partials(inst, orc, ops)
    inst.likelihood_buffer(None, node, dad)
    orc.likelihood_buffer(None, node, dad)
    dg, ddg, _ = inst.likelihood_derv(length)
    do, ddo = orc.likelihood_derv(length)
    assert abs(dg - do) <= DERV_RTOL * max(1.0, abs(do))
    assert abs(ddg - ddo) <= DERV_RTOL * max(1.0, abs(ddo))


@pytest.mark.parametrize("ncat", [1, 2, 8])
def test_category_counts(ctx, ncat):
    w = ia.Workload(seq_type=ia.DNA, ntaxa=24, nsite=1500, ncat=ncat, mean_len=0.15, alpha=0.6,
                    p_inv=0.1, seed=30 + ncat)
    all_branch_parity(w, *setup(ctx, w), stride=9)


def test_three_taxa(ctx):
    w = ia.Workload(seq_type=ia.DNA, ntaxa=3, nsite=200, ncat=4, seed=77)
    all_branch_parity(w, *setup(ctx, w))


def test_newton_branch_optimisation(ctx):
    w = ia.Workload(seq_type=ia.DNA, ntaxa=40, nsite=3000, ncat=4, mean_len=0.1, alpha=0.5,
                    p_inv=0.2, gap_frac=0.02, seed=1234)
    planner, inst, orc = setup(ctx, w)
    for a, b in planner.branches()[::4]:
        ops, node, dad, length = planner.plan(a, b)
        xg, lg, ng = inst.optimize_branch(ops, node, dad, length * 3.0)
        xo, lo, no, st = orc.optimize_branch(ops, node, dad, length * 3.0)
        assert st == 0
        assert abs(xg - xo) <= 1e-7 * max(xo, 1e-6), (a, b, xg, xo)
        assert abs(lg - lo) <= LNL_RTOL * abs(lo)
        assert ng == no


def test_optimize_all_branches_driver(ctx):
    """PhyloTree::optimizeAllBranches through the C++ driver vs the oracle
    driven by the same plans."""
    w = ia.Workload(seq_type=ia.DNA, ntaxa=20, nsite=1500, ncat=4, mean_len=0.1, alpha=0.5,
                    seed=99)
    gpu_tree = ia.PhyloTree(ctx, w)
    lnl0 = gpu_tree.compute_likelihood()
    lnl, nd = gpu_tree.optimize_all_branches(iterations=1, tolerance=1e-3, max_nr=100)
    assert lnl >= lnl0 - 1e-6 and nd > 0
    # the oracle replays optimizeAllBranches with the same planner
    planner = ia.PhyloTree(None, w)
    orc = OracleTree.from_workload(w, nthreads=4)
    branches = planner.branches()
    a0, b0 = branches[0]
    ops, node, dad, length = planner.plan(b0, a0)
    orc.likelihood_branch(ops, node, dad, length)
    for a, b in branches:
        ops, node, dad, length = planner.plan(b, a)
        x, _, _, st = orc.optimize_branch(ops, node, dad, length)
        assert st == 0
        planner.set_length(a, b, x)
        assert abs(gpu_tree.get_length(a, b) - x) <= 1e-7 * max(x, 1e-6)
    lo = orc.likelihood_from_buffer(planner.get_length(*branches[-1]))
    assert abs(lnl - lo) <= LNL_RTOL * abs(lo)


def test_rell(ctx):
    w = ia.Workload(seq_type=ia.DNA, ntaxa=16, nsite=2000, ncat=4, seed=3)
    planner, inst, orc = setup(ctx, w)
    ops, node, dad, length = planner.plan(*planner.branches()[0])
    _, plh = inst.likelihood_branch(ops, node, dad, length, want_pattern_lh=True)
    rng = np.random.default_rng(0)
    freq = w.get("freq")
    weights = np.stack([rng.multinomial(int(freq.sum()), freq / freq.sum())
                        for _ in range(50)]).astype(np.float32)
    out = inst.rell(weights)
    np.testing.assert_allclose(out, weights.astype(np.float64) @ plh, rtol=1e-12)


def test_fp32_storage_form(ctx):
    """FP32 partials/contractions, FP64 log, sums and derivatives."""
    w = ia.Workload(seq_type=ia.DNA, ntaxa=60, nsite=5000, ncat=4, mean_len=0.1, alpha=0.5,
                    p_inv=0.2, gap_frac=0.02, seed=20261204)
    planner, inst32, orc = setup(ctx, w, precision=ia.PREC_F32)
    for a, b in planner.branches()[::11]:
        ops, node, dad, length = planner.plan(a, b)
        lg = inst32.likelihood_branch(ops, node, dad, length)
        lo = orc.likelihood_branch(ops, node, dad, length)
        assert abs(lg - lo) <= 1e-6 * abs(lo), (a, b, lg, lo, abs(lg - lo))
        inst32.likelihood_buffer(None, node, dad)
        orc.likelihood_buffer(None, node, dad)
        dg, ddg, _ = inst32.likelihood_derv(length)
        do, ddo = orc.likelihood_derv(length)
        assert abs(dg - do) <= 1e-4 * max(1.0, abs(do))


def test_error_paths(ctx):
    w = ia.Workload(seq_type=ia.DNA, ntaxa=8, nsite=100, seed=1)
    planner, inst, orc = setup(ctx, w)
    with pytest.raises(ia.PlkError) as e:
        inst.likelihood_derv(0.1)
    assert e.value.status == 5          # theta not computed
    ops, node, dad, length = planner.plan(*planner.branches()[0])
    bad = ops.copy()
    bad[0]["dad_slot"] = 999
    with pytest.raises(ia.PlkError) as e:
        inst.likelihood_branch(bad, node, dad, length)
    assert e.value.status == 1
    small = ia.Context(0, hbm_budget=1 << 20)
    big = ia.Workload(seq_type=ia.DNA, ntaxa=64, nsite=20000, seed=2)
    with pytest.raises(ia.PlkError) as e:
        ia.Instance.from_workload(small, big)
    assert e.value.status == 2
    small.close()
    # the engine still works after the failures
    assert np.isfinite(inst.likelihood_branch(ops, node, dad, length))


@pytest.mark.parametrize("seq_type,precision,safe,ncat", [
    (ia.DNA, ia.PREC_F64, False, 4),
    (ia.DNA, ia.PREC_F64, True, 4),
    (ia.DNA, ia.PREC_F64, False, 1),
    (ia.PROTEIN, ia.PREC_F64, False, 4),
    (ia.DNA, ia.PREC_F32, False, 4),
])
def test_device_newton_sweep_matches_host_newton(ctx, seq_type, precision, safe, ncat):
    """One cooperative launch per optimizeAllBranches pass gives the same
    lengths, evaluation counts and lnL as minimizeNewton on the host."""
    w = ia.Workload(seq_type=seq_type, ntaxa=40, nsite=3000, ncat=ncat, mean_len=0.1, alpha=0.5,
                    p_inv=0.1 if seq_type == ia.DNA else 0.0, gap_frac=0.02, seed=4242)
    dev = ia.PhyloTree(ctx, w, precision=precision, safe_numeric=safe)
    host = ia.PhyloTree(ctx, w, precision=precision, safe_numeric=safe)
    host.set_device_newton(False)
    ld, nd = dev.optimize_all_branches(iterations=3, tolerance=1e-3, max_nr=100)
    lh, nh = host.optimize_all_branches(iterations=3, tolerance=1e-3, max_nr=100)
    assert nd == nh
    assert abs(ld - lh) <= 1e-9 * abs(lh), (ld, lh)
    for a, b in dev.branches():
        xd, xh = dev.get_length(a, b), host.get_length(a, b)
        assert abs(xd - xh) <= 1e-9 * max(xh, 1e-6), (a, b, xd, xh)
    # the tree lnL at the new lengths, recomputed from scratch
    dev.clear_all_partial()
    assert abs(dev.compute_likelihood() - ld) <= 1e-8 * abs(ld)


@pytest.mark.parametrize("precision,ncat", [(ia.PREC_F64, 4), (ia.PREC_F64, 2), (ia.PREC_F32, 1)])
def test_codon_device_sweep_matches_host_newton(ctx, precision, ncat):
    """61 states in the device sweep (tables built by the grid into HBM,
    the codon matrix-core tile, streamed theta) against minimizeNewton on the
    host: the same lengths, evaluation counts and lnL over two passes."""
    w = ia.Workload(seq_type=ia.CODON, ntaxa=16, nsite=600, ncat=ncat, mean_len=0.1, alpha=0.5,
                    seed=4343 + ncat)
    print(f"codon sweep precision={precision} ncat={ncat}", flush=True)
    dev = ia.PhyloTree(ctx, w, precision=precision)
    host = ia.PhyloTree(ctx, w, precision=precision)
    host.set_device_newton(False)
    try:
        ld, nd = dev.optimize_all_branches(iterations=2, tolerance=1e-3, max_nr=100)
        lh, nh = host.optimize_all_branches(iterations=2, tolerance=1e-3, max_nr=100)
        assert nd == nh
        assert abs(ld - lh) <= 1e-9 * abs(lh), (ld, lh)
        for a, b in dev.branches():
            xd, xh = dev.get_length(a, b), host.get_length(a, b)
            assert abs(xd - xh) <= 1e-9 * max(xh, 1e-6), (a, b, xd, xh)
        dev.clear_all_partial()
        assert abs(dev.compute_likelihood() - ld) <= 1e-8 * abs(ld)
    finally:
        dev.close()
        host.close()


def context_with(env):
    """A second context created with the environment knobs in env set (they
    are read once, when a context is created)."""
    import os
    old = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        return ia.Context(0)
    finally:
        for k, v in old.items():
            if v is None:
                del os.environ[k]
            else:
                os.environ[k] = v


@pytest.mark.parametrize("precision", [ia.PREC_F64, ia.PREC_F32])
def test_device_sweep_theta_in_lds_or_hbm(ctx, precision):
    """The device Newton pass with theta resident in LDS (one thread per
    pattern in the evaluations) and with theta in HBM (the path of
    alignments too large for LDS) take the same Newton path."""
    w = ia.Workload(seq_type=ia.DNA, ntaxa=48, nsite=20000, ncat=4, mean_len=0.1, alpha=0.5,
                    p_inv=0.1, gap_frac=0.02, seed=4343)
    ctx_hbm = context_with({"PLK_SWEEP_LDS": "0"})
    try:
        lds = ia.PhyloTree(ctx, w, precision=precision)
        hbm = ia.PhyloTree(ctx_hbm, w, precision=precision)
        l1, n1 = lds.optimize_all_branches(iterations=2, tolerance=1e-3, max_nr=100)
        l2, n2 = hbm.optimize_all_branches(iterations=2, tolerance=1e-3, max_nr=100)
        assert n1 == n2
        assert abs(l1 - l2) <= 1e-9 * abs(l2), (l1, l2)
        for a, b in lds.branches():
            x1, x2 = lds.get_length(a, b), hbm.get_length(a, b)
            assert abs(x1 - x2) <= 1e-9 * max(x2, 1e-6), (a, b, x1, x2)
        lds.close()
        hbm.close()
    finally:
        ctx_hbm.close()


@pytest.mark.parametrize("sub", ["2", "4"])
def test_protein_traversal_workgroup_shapes_bit_identical(ctx, sub):
    """Protein traversals with 2 or 4 sub-tiles of 16 patterns per workgroup
    give the same partials and lnL bits as one sub-tile per workgroup."""
    w = ia.Workload(seq_type=ia.PROTEIN, ntaxa=60, nsite=3000, ncat=4, mean_len=0.5, alpha=1.0,
                    seed=91)
    planner = ia.PhyloTree(None, w)
    ops, node, dad, length = planner.plan(*planner.branches()[3])
    ctx_sub = context_with({"PLK_MFMA_SUB": sub})
    try:
        for prec in (ia.PREC_F64, ia.PREC_F32):
            a = ia.Instance.from_workload(ctx, w, precision=prec)
            b = ia.Instance.from_workload(ctx_sub, w, precision=prec)
            assert a.likelihood_branch(ops, node, dad, length) == b.likelihood_branch(ops, node, dad, length)
            for op in ops[::7]:
                pa, sa = a.read_partial(int(op["dad_slot"]))
                pb, sb = b.read_partial(int(op["dad_slot"]))
                np.testing.assert_array_equal(pa, pb)
                np.testing.assert_array_equal(sa, sb)
            a.close()
            b.close()
    finally:
        ctx_sub.close()


def partition_set(n_parts, ntaxa=24, seed=7, f32_every=0):
    """Partitions of mixed data type and length on one shared tree."""
    base = ia.Workload(seq_type=ia.DNA, ntaxa=ntaxa, nsite=10, seed=seed)
    edges = base.edges()
    rng = np.random.default_rng(seed)
    parts = []
    for k in range(n_parts):
        st = ia.PROTEIN if k % 3 == 2 else ia.DNA
        nsite = int(rng.integers(60, 1500))
        w = ia.Workload(seq_type=st, ntaxa=ntaxa, nsite=nsite, ncat=4, alpha=-1.0,
                        p_inv=0.1 if st == ia.DNA and k % 2 else 0.0,
                        gap_frac=0.02 if st == ia.DNA else 0.0, seed=1000 + k, edges=edges)
        prec = ia.PREC_F32 if f32_every and k % f32_every == 0 else ia.PREC_F64
        parts.append((w, prec))
    return parts


def test_partition_batch_equals_per_partition(ctx):
    """plk_likelihood_branch_batch over mixed DNA/protein, FP64/FP32
    partitions equals one plk_likelihood_branch per partition and is
    reproducible."""
    parts = partition_set(40, f32_every=5)
    jobs, single = [], []
    for k, (w, prec) in enumerate(parts):
        planner = ia.PhyloTree(None, w)
        branches = planner.branches()
        a, b = branches[k % len(branches)]
        ops, node, dad, length = planner.plan(a, b)
        inst = ia.Instance.from_workload(ctx, w, precision=prec)
        jobs.append((inst, ops, node, dad, length))
    for inst, ops, node, dad, length in jobs:
        single.append(inst.likelihood_branch(ops, node, dad, length))
    # repeated calls of one shape (each a different launch-shape candidate
    # once a batch has enough tiles to be tuned) give the same bits
    first = ia.likelihood_branch_batch(ctx, jobs)
    second = ia.likelihood_branch_batch(ctx, jobs)
    third = ia.likelihood_branch_batch(ctx, jobs)
    np.testing.assert_array_equal(first, second)
    np.testing.assert_array_equal(first, third)
    np.testing.assert_allclose(first, np.array(single), rtol=1e-12, atol=0)
    # against the oracle for the FP64 partitions
    for k, (w, prec) in enumerate(parts[:8]):
        if prec != ia.PREC_F64:
            continue
        inst, ops, node, dad, length = jobs[k]
        orc = OracleTree.from_workload(w)
        lo = orc.likelihood_branch(ops, node, dad, length)
        assert abs(first[k] - lo) <= LNL_RTOL * abs(lo)
    for inst, *_ in jobs:
        inst.close()


def test_batch_protein_scaling_bit_identical(ctx):
    """Batched protein tiles (four 16-pattern MFMA sub-tiles per workgroup)
    with partials crossing the scaling threshold, in normal and safe mode,
    FP64 and FP32: every partial and scale counter the batch leaves equals
    the lone traversal kernel's, and the lnL agrees to rounding."""
    specs = [(False, ia.PREC_F64), (True, ia.PREC_F64), (False, ia.PREC_F32), (True, ia.PREC_F32)]
    jobs, twins = [], []
    for k, (safe, prec) in enumerate(specs):
        w = ia.Workload(seq_type=ia.PROTEIN, ntaxa=150, nsite=300 + 70 * k, ncat=4, mean_len=0.5,
                        alpha=1.0, seed=17 + k)
        planner = ia.PhyloTree(None, w)
        ops, node, dad, length = planner.plan(*planner.branches()[k])
        a = ia.Instance.from_workload(ctx, w, precision=prec, safe_numeric=safe)
        b = ia.Instance.from_workload(ctx, w, precision=prec, safe_numeric=safe)
        jobs.append((a, ops, node, dad, length))
        twins.append((b, ops, node, dad, length))
    batched = ia.likelihood_branch_batch(ctx, jobs)
    single = [b.likelihood_branch(ops, node, dad, length) for b, ops, node, dad, length in twins]
    np.testing.assert_allclose(batched, np.array(single), rtol=1e-12, atol=0)
    for (a, ops, *_), (b, *_) in zip(jobs, twins):
        scaled = 0
        for op in ops:
            pa, sa = a.read_partial(int(op["dad_slot"]))
            pb, sb = b.read_partial(int(op["dad_slot"]))
            np.testing.assert_array_equal(pa, pb)
            np.testing.assert_array_equal(sa, sb)
            scaled = max(scaled, int(sb.max()))
        assert scaled >= 1, "the workload was meant to cross the scaling threshold"
    for (a, *_), (b, *_) in zip(jobs, twins):
        a.close()
        b.close()


def test_batch_mixed_shapes(ctx):
    """A batch mixing fusable items with 61-state and 2-category items: the
    fused ones share one launch, the others run alone, and every result
    equals its own plk_likelihood_branch."""
    base = ia.Workload(seq_type=ia.DNA, ntaxa=12, nsite=10, seed=3)
    edges = base.edges()
    specs = [(ia.DNA, 4, 900), (ia.CODON, 4, 300), (ia.PROTEIN, 4, 400), (ia.DNA, 2, 700),
             (ia.CODON, 1, 200), (ia.DNA, 4, 1500)]
    jobs = []
    for k, (st, ncat, nsite) in enumerate(specs):
        w = ia.Workload(seq_type=st, ntaxa=12, nsite=nsite, ncat=ncat, alpha=0.7, seed=500 + k,
                        edges=edges)
        planner = ia.PhyloTree(None, w)
        a, b = planner.branches()[k % 5]
        ops, node, dad, length = planner.plan(a, b)
        inst = ia.Instance.from_workload(ctx, w, safe_numeric=st == ia.CODON)
        jobs.append((inst, ops, node, dad, length))
    single = [inst.likelihood_branch(ops, node, dad, length) for inst, ops, node, dad, length in jobs]
    batched = ia.likelihood_branch_batch(ctx, jobs)
    np.testing.assert_allclose(batched, np.array(single), rtol=1e-12, atol=0)
    for inst, *_ in jobs:
        inst.close()


def test_derv_batch_bit_identical(ctx):
    """plk_likelihood_derv_batch over mixed shapes (and an instance twice)
    gives exactly plk_likelihood_derv's df, ddf and lnL per item."""
    specs = [(ia.DNA, 4, ia.PREC_F64), (ia.PROTEIN, 4, ia.PREC_F64), (ia.CODON, 4, ia.PREC_F64),
             (ia.DNA, 2, ia.PREC_F32), (ia.DNA, 4, ia.PREC_F64), (ia.CODON, 1, ia.PREC_F32)]
    insts = []
    for k, (st, ncat, prec) in enumerate(specs):
        w = ia.Workload(seq_type=st, ntaxa=9 + k, nsite=300 + 97 * k, ncat=ncat, alpha=0.6,
                        seed=70 + k)
        planner = ia.PhyloTree(None, w)
        ops, node, dad, length = planner.plan(*planner.branches()[k % 3])
        inst = ia.Instance.from_workload(ctx, w, precision=prec, safe_numeric=st == ia.CODON)
        inst.likelihood_buffer(ops, node, dad)
        insts.append((inst, length))
    pairs = [(inst, x * f) for inst, x in insts for f in (0.5, 2.0)] + [(insts[0][0], 0.3)]
    df, ddf, lnl = ia.likelihood_derv_batch(ctx, pairs)
    for k, (inst, x) in enumerate(pairs):
        d1, d2, l1 = inst.likelihood_derv(x)
        assert (df[k], ddf[k], lnl[k]) == (d1, d2, l1), (k, df[k], d1, ddf[k], d2, lnl[k], l1)
    for inst, _ in insts:
        inst.close()


def test_batch_theta_mode_bit_identical(ctx):
    """plk_batch_evaluate with theta items next to lnL items: the theta it
    leaves gives exactly the derivatives of plk_likelihood_buffer's theta,
    and the lnL items equal plk_likelihood_branch_batch."""
    specs = [(ia.DNA, ia.PREC_F64), (ia.PROTEIN, ia.PREC_F64), (ia.DNA, ia.PREC_F32),
             (ia.CODON, ia.PREC_F64), (ia.DNA, ia.PREC_F64), (ia.PROTEIN, ia.PREC_F32)]
    jobs, refs, modes = [], [], []
    for k, (st, prec) in enumerate(specs):
        w = ia.Workload(seq_type=st, ntaxa=10 + 2 * k, nsite=400 + 150 * k, ncat=4, alpha=0.6,
                        seed=300 + k)
        planner = ia.PhyloTree(None, w)
        ops, node, dad, length = planner.plan(*planner.branches()[k % 4])
        safe = st == ia.CODON
        a = ia.Instance.from_workload(ctx, w, precision=prec, safe_numeric=safe)
        b = ia.Instance.from_workload(ctx, w, precision=prec, safe_numeric=safe)
        jobs.append((a, ops, node, dad, length))
        refs.append((b, ops, node, dad, length))
        modes.append(k % 2)
    out = ia.batch_evaluate(ctx, jobs, modes)
    lnl_ref = ia.likelihood_branch_batch(ctx, [r for r, m in zip(refs, modes) if m == 0])
    np.testing.assert_array_equal(out[np.array(modes) == 0], lnl_ref)
    for (a, *_), (b, ops, node, dad, length), m in zip(jobs, refs, modes):
        if m != 1:
            continue
        b.likelihood_buffer(ops, node, dad)
        for x in (0.5 * length, length, 3 * length + 0.01):
            assert a.likelihood_derv(x) == b.likelihood_derv(x)
    for (a, *_), (b, *_) in zip(jobs, refs):
        a.close()
        b.close()


@pytest.mark.parametrize("seq_type,ncat,safe", [(ia.DNA, 5, False), (ia.DNA, 3, True), (ia.PROTEIN, 6, False)])
def test_category_count_padded_reversible(ctx, seq_type, ncat, safe):
    """A rate model with a category count that is not built (FreeRate +R3,
    +R5, +R6) runs in an instance of the next built count, the extra
    categories copies of category 0 with weight 0: lnL, derivatives and Newton
    equal the oracle's at the model's own count."""
    w = ia.Workload(seq_type=seq_type, ntaxa=24, nsite=1500, ncat=4, mean_len=0.3 if safe else 0.15,
                    seed=300 + ncat)
    m = w.model_arrays()
    rng = np.random.default_rng(ncat)
    rates = np.sort(rng.uniform(0.1, 3.0, ncat))
    props = rng.dirichlet(np.full(ncat, 3.0))
    rates /= rates @ props
    cpad = 1 << (ncat - 1).bit_length()
    r_pad = np.concatenate([rates, np.full(cpad - ncat, rates[0])])
    p_pad = np.concatenate([props, np.zeros(cpad - ncat)])
    inst = ia.Instance(ctx, w.ntaxa, w.nstates, cpad, w.npatterns, w.ntaxa, w.state_unknown, ia.PREC_F64, safe)
    inst.set_patterns(w.get("states"), w.get("freq"), w.get("invar"))
    inst.set_model(m["eval"], m["evec"], m["inv_evec"], m["tip"], r_pad, p_pad)
    orc = OracleTree(w.ntaxa, w.nstates, ncat, w.npatterns, w.ntaxa, w.state_unknown, safe, 2)
    orc.set_patterns(w.get("states"), w.get("freq"), w.get("invar"))
    orc.set_model(m["eval"], m["evec"], m["inv_evec"], m["tip"], rates, props)
    planner = ia.PhyloTree(None, w)
    try:
        for a, b in planner.branches()[::6]:
            ops, node, dad, length = planner.plan(a, b)
            lg = inst.likelihood_branch(ops, node, dad, length)
            lo = orc.likelihood_branch(ops, node, dad, length)
            assert abs(lg - lo) <= 1e-9 * abs(lo), (a, b, lg, lo)
            inst.likelihood_buffer(ops, node, dad)
            orc.likelihood_buffer(ops, node, dad)
            dfg, ddfg, _ = inst.likelihood_derv(length)
            dfo, ddfo = orc.likelihood_derv(length)
            assert abs(dfg - dfo) <= 1e-7 * max(1.0, abs(dfo))
            assert abs(ddfg - ddfo) <= 1e-7 * max(1.0, abs(ddfo))
        ops, node, dad, length = planner.plan(*planner.branches()[1][::-1])
        xg, lg, ng = inst.optimize_branch(ops, node, dad, length)
        xo, lo, no, so = orc.optimize_branch(ops, node, dad, length)
        assert so == 0 and ng == no and abs(xg - xo) <= 1e-7 * max(1.0, xo)
    finally:
        inst.close()
        planner.close()
