"""plk_likelihood_derv_batch_noalloc (scratch in a module-scope device array,
no device allocation) against plk_likelihood_derv_batch, called through the
C ABI: bit-identical results on two calls in a row, the second smaller, and
a batch beyond the scratch size rejected before anything runs."""
import ctypes

import numpy as np
import pytest

import iqtree2_amd as ia
from iqtree2_amd._capi import PLK_OK, dptr, i32ptr, lib

pytestmark = pytest.mark.gpu

PLK_ERR_NOMEM = 2


def theta_instances(ctx, specs):
    """instances with their theta buffer computed at the first branch"""
    out = []
    for seq_type, ntaxa, nsite, seed in specs:
        w = ia.Workload(seq_type=seq_type, ntaxa=ntaxa, nsite=nsite, ncat=4, mean_len=0.1,
                        alpha=0.5, seed=seed)
        planner = ia.PhyloTree(None, w)
        inst = ia.Instance.from_workload(ctx, w, safe_numeric=seq_type == ia.CODON)
        a, b = planner.branches()[0]
        ops, node, dad, length = planner.plan(b, a)
        inst.likelihood_buffer(ops, node, dad)
        out.append((inst, length))
        planner.close()
        w.close()
    return out


def derv(fn, ctx, items):
    n = len(items)
    ids = np.array([inst.id for inst, _ in items], np.int32)
    lens = np.array([x for _, x in items], np.float64)
    df, ddf, lnl = np.zeros(n), np.zeros(n), np.zeros(n)
    st = fn(ctx.handle, i32ptr(ids), dptr(lens), n, dptr(df), dptr(ddf), dptr(lnl))
    return st, df, ddf, lnl


def test_noalloc_matches_allocating_form(ctx):
    big = theta_instances(ctx, [(ia.DNA, 24, 3000, 11), (ia.PROTEIN, 20, 1200, 12),
                                (ia.DNA, 40, 9000, 13), (ia.CODON, 12, 300, 14)])
    small = big[:2]
    L = lib()
    try:
        for items, scale in ((big, 1.0), (small, 0.7)):
            # several lengths per instance, as the Newton iterations of many
            # branch optimisations share one batch
            batch = [(inst, x * scale * f) for inst, x in items for f in (0.5, 1.0, 2.0)]
            st_a, df_a, ddf_a, l_a = derv(L.plk_likelihood_derv_batch, ctx, batch)
            st_b, df_b, ddf_b, l_b = derv(L.plk_likelihood_derv_batch_noalloc, ctx, batch)
            assert st_a == PLK_OK and st_b == PLK_OK
            np.testing.assert_array_equal(df_b, df_a)
            np.testing.assert_array_equal(ddf_b, ddf_a)
            np.testing.assert_array_equal(l_b, l_a)
            assert np.all(np.isfinite(l_a))
    finally:
        for inst, _ in big:
            inst.close()


def test_noalloc_rejects_a_batch_beyond_its_scratch(ctx):
    # 4 MiB hold the sums of about 170k tiles: 300 items of a ~90k-pattern
    # instance (about 1,400 tiles each) exceed it; the size check fails before
    # any copy or launch
    items = theta_instances(ctx, [(ia.DNA, 16, 100000, 21)])
    L = lib()
    try:
        batch = [items[0]] * 300
        st, _, _, _ = derv(L.plk_likelihood_derv_batch_noalloc, ctx, batch)
        assert st == PLK_ERR_NOMEM
        assert b"no-allocation" in L.plk_last_error()
        st, _, _, _ = derv(L.plk_likelihood_derv_batch_noalloc, ctx, batch[:1])
        assert st == PLK_OK
    finally:
        items[0][0].close()
