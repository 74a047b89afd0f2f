"""CPU tests of the host side: C-ABI exports, pattern compression (bit-exact
against the Python restatement of alignment.cpp), workload determinism and
the traversal planner's invariants."""
import os
import re

import numpy as np
import pytest

import iqtree2_amd as ia
from iqtree2_amd import _capi
from oracle import patterns as refpat

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def header_functions():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(plk[hs]?_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    lib = ia.lib()
    declared = header_functions()
    assert len(declared) > 30
    missing = [n for n in sorted(declared) if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python binding prototypes every one of them
    assert declared <= set(_capi.exported_names()), declared - set(_capi.exported_names())


def random_rows(rng, seq_type, nseq, nsite):
    if seq_type == ia.DNA:
        alpha = list("ACGT") * 8 + list("-NRYWSMKBHDV?")
        rows = ["".join(rng.choice(alpha, nsite)) for _ in range(nseq)]
        # constant and ambiguous-constant columns
        for site in range(0, nsite, 7):
            ch = rng.choice(list("ACGTRN-"))
            for r in range(nseq):
                rows[r] = rows[r][:site] + (ch if rng.random() < 0.8 else "-") + rows[r][site + 1:]
        return rows
    if seq_type == ia.PROTEIN:
        alpha = list("ARNDCQEGHILKMFPSTWYV") * 3 + list("BZJX-*?")
        rows = ["".join(rng.choice(alpha, nsite)) for _ in range(nseq)]
        for site in range(0, nsite, 5):
            ch = rng.choice(list("NDBZJ"))
            for r in range(nseq):
                rows[r] = rows[r][:site] + ch + rows[r][site + 1:]
        return rows
    codons = [a + b + c for a in "ACGT" for b in "ACGT" for c in "ACGT"
              if refpat.AMINO[a + b + c] != "*"]
    pool = codons[:8] + ["---", "NNN", "AC-"]
    return ["".join(rng.choice(pool, nsite)) for _ in range(nseq)]


@pytest.mark.parametrize("seq_type", [ia.DNA, ia.PROTEIN, ia.CODON])
def test_patterns_match_restatement(seq_type):
    rng = np.random.default_rng(100 + seq_type)
    rows = random_rows(rng, seq_type, 9, 400)
    p = ia.Patterns(seq_type, rows)
    cols, freq, site_pattern, consts = refpat.build_patterns(seq_type, rows)
    assert p.npatterns == len(cols)
    np.testing.assert_array_equal(p.states.T, np.array(cols, dtype=np.uint8))
    np.testing.assert_array_equal(p.freq, freq)
    np.testing.assert_array_equal(p.site_pattern, site_pattern)
    np.testing.assert_array_equal(p.const_char, consts)
    assert p.freq.sum() == p.nsite
    # round trip: site_pattern reproduces every column
    step = 3 if seq_type == ia.CODON else 1
    for site in range(0, p.nsite, 13):
        col = [refpat.convert_codon(r[3 * site:3 * site + 3]) if step == 3
               else refpat.convert_state(seq_type, r[site]) for r in rows]
        np.testing.assert_array_equal(p.states[:, p.site_pattern[site]], col)
    # first-occurrence order
    first = [site_pattern.index(k) for k in range(len(cols))]
    assert first == sorted(first)
    pi = rng.dirichlet(np.ones(refpat.NUM_STATES[seq_type]))
    for p_inv in (0.0, 0.25):
        np.testing.assert_array_equal(p.invar(pi, p_inv),
                                      np.array(refpat.ptn_invar(seq_type, consts, pi, p_inv)))


def read_phylip(path):
    with open(path) as f:
        ntax, nchar = map(int, f.readline().split())
        rows = []
        for line in f:
            parts = line.split()
            if len(parts) >= 2:
                rows.append("".join(parts[1:]))
    assert len(rows) == ntax and all(len(r) == nchar for r in rows)
    return rows


def test_reference_example_alignment():
    """example/example.phy of the reference (17 taxa x 1,998 DNA sites)."""
    rows = read_phylip(os.path.join(GOLDEN, "example.phy"))
    p = ia.Patterns(ia.DNA, rows)
    cols, freq, site_pattern, consts = refpat.build_patterns(ia.DNA, rows)
    assert p.nsite == 1998 and p.npatterns == len(cols)
    np.testing.assert_array_equal(p.freq, freq)
    np.testing.assert_array_equal(p.const_char, consts)
    np.testing.assert_array_equal(p.site_pattern, site_pattern)


def test_workload_is_deterministic():
    a = ia.Workload(seq_type=ia.DNA, ntaxa=30, nsite=500, seed=42, p_inv=0.1, gap_frac=0.02)
    b = ia.Workload(seq_type=ia.DNA, ntaxa=30, nsite=500, seed=42, p_inv=0.1, gap_frac=0.02)
    c = ia.Workload(seq_type=ia.DNA, ntaxa=30, nsite=500, seed=43, p_inv=0.1, gap_frac=0.02)
    for k in ("states", "freq", "invar", "eval", "evec", "edges_len", "rates"):
        np.testing.assert_array_equal(a.get(k), b.get(k))
    assert not np.array_equal(a.get("edges_len"), c.get("edges_len"))


def check_plan(planner, w, ops, node, dad, computed):
    """post-order: every internal child is computed before (or by an earlier op)"""
    N = w.ntaxa
    for op in ops:
        for k in (0, 1):
            if not op[f"c{k}_tip"]:
                assert op[f"c{k}_id"] in computed, "child used before it was computed"
            else:
                assert 0 <= op[f"c{k}_id"] < N
        computed.add(op["dad_slot"])
    assert node.is_tip == 0 and node.id in computed
    if not dad.is_tip:
        assert dad.id in computed


def test_planner_invariants():
    w = ia.Workload(seq_type=ia.DNA, ntaxa=50, nsite=100, seed=8)
    planner = ia.PhyloTree(None, w)
    branches = planner.branches()
    assert len(branches) == 2 * w.ntaxa - 3
    assert len({tuple(sorted(b)) for b in branches}) == len(branches)
    # first plan computes a full post-order: N - 2 internal partials
    ops, node, dad, _ = planner.plan(*branches[0])
    assert len(ops) == w.ntaxa - 2
    check_plan(planner, w, ops, node, dad, set())
    # re-planning the same branch needs nothing
    assert len(planner.plan(*branches[0])[0]) == 0
    # walking the pre-order re-orients O(1) partials per branch on average
    total = 0
    computed = set(range(w.ntaxa - 2))
    for a, b in branches:
        ops, node, dad, _ = planner.plan(a, b)
        check_plan(planner, w, ops, node, dad, computed)
        total += len(ops)
    assert total <= 3 * len(branches)
    # changing a length invalidates exactly the partials whose subtree holds it
    a, b = branches[5]
    planner.plan(a, b)
    planner.set_length(a, b, planner.get_length(a, b) * 1.5)
    ops, _, _, _ = planner.plan(a, b)
    assert len(ops) == 0
    far = branches[-1]
    ops, _, _, _ = planner.plan(*far)
    assert len(ops) > 0


def test_nni_planner_buffers_and_restore():
    """getBestNNIForBran's buffer handling in the planner: during an NNI
    evaluation node1 and node2 live in the two NNI slots (after the
    internal-node slots), their exchanged topology is planned there, and
    nni_end restores topology, lengths and the central partials (nothing
    to recompute afterwards)."""
    w = ia.Workload(seq_type=ia.DNA, ntaxa=14, nsite=60, seed=12)
    pl = ia.PhyloTree(None, w)
    central = w.ntaxa - 2
    nnodes = 2 * w.ntaxa - 2
    inner = pl.inner_branches()
    assert len(inner) == w.ntaxa - 3
    nb0 = {v: pl.neighbors(v) for v in range(nnodes)}
    br = pl.branches()
    len0 = {b: pl.get_length(*b) for b in br}
    pl.plan(*br[0])                               # every partial computed once
    for node1, node2 in inner[:4]:
        pl.nni_begin(node1, node2)
        n1 = [x for x in pl.neighbors(node1) if x != node2]
        n2 = [y for y in pl.neighbors(node2) if y != node1]
        pl.nni_swap(node1, n1[0], node2, n2[1])
        assert n2[1] in pl.neighbors(node1) and n1[0] in pl.neighbors(node2)
        assert pl.get_length(node1, n2[1]) == len0.get((node2, n2[1]), len0.get((n2[1], node2)))
        ops, node, dad, _ = pl.plan(n1[1], node1)
        dads = [int(o["dad_slot"]) for o in ops]
        # node1's partial seen from n1[1] is built in an NNI slot, from
        # node2's (the other NNI slot) and the swapped-in subtree
        assert dads[-1] in (central, central + 1)
        assert set(dads) & {central, central + 1}
        pl.set_length_keep(node1, n1[1], 0.33)
        pl.nni_swap(node1, n2[1], node2, n1[0])
        pl.nni_end()
        assert {v: pl.neighbors(v) for v in range(nnodes)} == nb0
        assert {b: pl.get_length(*b) for b in br} == len0
    # central partials were left intact: no op needed for the first branch
    assert len(pl.plan(*br[0])[0]) <= 2 * 4
