"""MI355X-native Felsenstein-pruning likelihood engine for IQ-TREE 2.

The compute path is libplk.so (hand-written HIP kernels for gfx950 behind a
C ABI, include/plk.h); this package is its Python host mirror.
"""
from . import _capi
from ._capi import PREC_F32, PREC_F64, PlkError, lib
from .engine import (BatchPlan, Context, Instance, Patterns, PhyloTree, Workload, end,
                     batch_evaluate, likelihood_branch_batch, likelihood_derv_batch,
                     ops_array, supertree_optimize_all_branches, minimize_one_dimen, gamma_rates,
                     models_optimize_gamma_shape, uninformative_patterns)

DNA, PROTEIN, CODON = _capi.DNA, _capi.PROTEIN, _capi.CODON

__all__ = ["BatchPlan", "Context", "Instance", "Patterns", "PhyloTree", "Workload", "end", "ops_array",
           "likelihood_branch_batch", "likelihood_derv_batch", "batch_evaluate",
           "supertree_optimize_all_branches", "minimize_one_dimen", "gamma_rates",
           "models_optimize_gamma_shape", "uninformative_patterns",
           "PREC_F32", "PREC_F64", "PlkError", "lib", "DNA", "PROTEIN", "CODON"]
