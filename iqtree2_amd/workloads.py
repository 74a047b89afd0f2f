"""Seeded synthetic workloads of the benchmark configurations.

Every generator draws the tree, the model parameters and the sequences from
fixed seeds (seed = 20261202 + k for configuration k), simulates the
sequences down the tree in FP64 and compresses them into site patterns
(alignment/alignment.cpp:1725-1929), so the GPU engine, the tests and the CPU
timing all see the same inputs.
"""
import numpy as np

from . import engine as _engine
from ._capi import CODON, DNA, PROTEIN

SEED = 20261202


def dna_16x1k():
    """DNA GTR+G4, 16 taxa x 1,000 sites, Exp(mean 0.1) branch lengths,
    Gamma alpha 0.5."""
    return _engine.Workload(seq_type=DNA, ntaxa=16, nsite=1000, ncat=4, mean_len=0.1, alpha=0.5,
                            seed=SEED)


def dna_200x100k(ntaxa=200, nsite=100_000):
    """DNA GTR+I+G4, 200 taxa x 100,000 sites (~60-100k unique patterns),
    p_inv 0.2, 2% of tip cells gaps or ambiguity codes, Gamma alpha 0.5,
    Exp(mean 0.1) branch lengths."""
    return _engine.Workload(seq_type=DNA, ntaxa=ntaxa, nsite=nsite, ncat=4, mean_len=0.1,
                            alpha=0.5, p_inv=0.2, gap_frac=0.02, seed=SEED + 1)


def protein_100x20k(ntaxa=100, nsite=20_000):
    """Protein +G4 with random exchangeabilities (Gamma(1,1)) and
    Dirichlet(2) frequencies, 100 taxa x 20,000 sites, Gamma alpha 0.8,
    Exp(mean 0.2) branch lengths."""
    return _engine.Workload(seq_type=PROTEIN, ntaxa=ntaxa, nsite=nsite, ncat=4, mean_len=0.2,
                            alpha=0.8, seed=SEED + 2)


def partitions_500(nparts=500, seed=SEED + 3):
    """500 partitions on one 64-taxon tree (Exp mean 0.1): partition lengths
    lognormal(ln 800, sigma 0.8) clipped to [100, 5000] sites, 70% DNA
    GTR+G4 / 30% protein +G4, each with independently seeded parameters
    (Gamma alpha ~ U(0.2, 2))."""
    rng = np.random.default_rng(seed)
    base = _engine.Workload(seq_type=DNA, ntaxa=64, nsite=10, mean_len=0.1, seed=seed)
    edges = base.edges()
    parts = []
    for k in range(nparts):
        nsite = int(np.clip(round(rng.lognormal(np.log(800.0), 0.8)), 100, 5000))
        st = PROTEIN if rng.random() < 0.3 else DNA
        parts.append(_engine.Workload(seq_type=st, ntaxa=64, nsite=nsite, ncat=4, alpha=-1.0,
                                      seed=seed * 1000 + k, edges=edges))
    return parts


def codon_64x5k():
    """Codon GY94 61-state +G4, 64 taxa x 5,000 codons, Exp(mean 0.1)."""
    return _engine.Workload(seq_type=CODON, ntaxa=64, nsite=5000, ncat=4, mean_len=0.1,
                            alpha=0.5, seed=77)
