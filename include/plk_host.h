/*
 * plk_host.h - host-side C ABI around the likelihood engine:
 *   - the data formats either side of the kernels (site patterns, model
 *     eigensystem, Gamma rates, eigen-basis tip vectors);
 *   - a seeded synthetic workload generator;
 *   - a PhyloTree-shaped driver (traversal planning with per-node buffer
 *     reorientation, branch Newton optimisation, optimizeAllBranches) that
 *     calls the engine in plk.h.
 * Functions returning int use the plk_status codes of plk.h.
 */
#ifndef PLK_HOST_H
#define PLK_HOST_H

#include <stddef.h>
#include <stdint.h>

#include "plk.h"

#ifdef __cplusplus
extern "C" {
#endif

enum { PLKH_DNA = 0, PLKH_PROTEIN = 1, PLKH_CODON = 2 };

/* ---- model inputs ------------------------------------------------------ */
int plkh_num_states(int seq_type);
int plkh_state_unknown(int seq_type);
/* Q from exchangeabilities (upper triangle) and pi, normalised to mean rate 1
 * (model/modelmarkov.cpp:1499-1512) */
int plkh_rate_matrix(int n, const double* exch_upper, const double* pi, double* Q);
/* eigensystem of a reversible Q (model/modelmarkov.cpp:1514-1556) */
int plkh_decompose(int n, const double* Q, const double* pi, double* eval, double* evec,
                   double* inv_evec);
/* discrete-Gamma mean rates and weights with +I rescaling
 * (model/rategamma.cpp:98-171, model/rategammainvar.cpp:25-42) */
int plkh_gamma_rates(int ncat, double alpha, double p_inv, double* rates, double* props);
/* Optimization::minimizeOneDimen (utils/optimization.cpp:322-373, with
 * brent_opt 185-320): the minimum of f(x, arg) on [xmin, xmax] from the
 * guess; *x the abscissa, *fx f there, *nevals the evaluations of f */
int plkh_minimize_one_dimen(double (*f)(double, void*), void* arg, double xmin, double xguess, double xmax,
                            double tol, double* x, double* fx, int* nevals);
/* eigen-basis tip vectors [SU+1][n] (tree/phylotreesse.cpp:455-466) */
int plkh_tip_vectors(int seq_type, const double* inv_evec, double* tip);
/* GY94 codon exchangeabilities (model/modelcodon.cpp:845-870) */
int plkh_codon_exchangeabilities(double kappa, double omega, double* exch_upper);

/* ---- site patterns (alignment/alignment.cpp:1725-1929, 925-1022) ---------- */
typedef struct plkh_patterns plkh_patterns;
/* chars: nseq rows of nchar characters, row-major */
plkh_patterns* plkh_patterns_build(int seq_type, int nseq, int nchar, const char* chars, int* status);
void plkh_patterns_destroy(plkh_patterns* p);
int plkh_patterns_dims(const plkh_patterns* p, int* nseq, int* nsite, int* npattern);
/* states uint8 [nseq][npattern], freq int32 [npattern], const_char int32
 * [npattern], site_pattern int32 [nsite]; any pointer may be NULL */
int plkh_patterns_get(const plkh_patterns* p, uint8_t* states, int32_t* freq, int32_t* const_char,
                      int32_t* site_pattern);
/* ptn_invar (tree/phylotreesse.cpp:549-617) for the first npattern entries */
int plkh_patterns_invar(const plkh_patterns* p, const double* pi, double p_inv, double* invar);
/* The unobserved patterns of Lewis's correction for parsimony-informative
 * sites (ASC_INFORMATIVE, Alignment::getUnobservedConstPatterns,
 * alignment/alignment.cpp:4181-4196): every parsimony-uninformative pattern
 * of nseq sequences over the nstates states, in the reference's order, as
 * rows of nseq states. *count gets the number; out (nullable) receives
 * min(count, cap) rows. PLK_ERR_UNSUPPORTED when there would be more than
 * max_patterns, or more than 31 states. Append them to the observed patterns
 * (weight 0, +I mass 0) and call plk_set_asc with their count. */
int plkh_uninformative_patterns(int nstates, int nseq, long max_patterns, uint8_t* out, long cap, long* count);

/* ---- synthetic workloads ---------------------------------------------- */
typedef struct {
    int seq_type;      /* PLKH_DNA / PLKH_PROTEIN / PLKH_CODON */
    int ntaxa;
    int nsite;         /* alignment columns (codons for PLKH_CODON) */
    int ncat;
    double mean_len;   /* Exp(mean) branch lengths */
    double alpha;      /* Gamma shape; <= 0 draws U(0.2, 2) */
    double p_inv;
    double gap_frac;   /* DNA: fraction of tip cells set to gap / ambiguity */
    uint64_t seed;
} plkh_workload_spec;

typedef struct plkh_workload plkh_workload;
plkh_workload* plkh_workload_create(const plkh_workload_spec* spec, int* status);
/* the same generator on a caller-given tree (edge list over ntaxa leaves) */
plkh_workload* plkh_workload_create_on_tree(const plkh_workload_spec* spec, int nedges,
                                            const int* u, const int* v, const double* len,
                                            int* status);
void plkh_workload_destroy(plkh_workload* w);
int plkh_workload_dims(const plkh_workload* w, int* ntaxa, int* nstates, int* ncat, int* npattern,
                       int* state_unknown, int* nsite);
/* arrays by name: "states" uint8 [ntaxa][P]; "freq", "invar" [P];
 * "eval" [n]; "evec", "inv_evec", "Q" [n][n]; "pi" [n]; "tip" [SU+1][n];
 * "rates", "props" [C]; "edges_u", "edges_v" int32 [2N-3]; "edges_len" [2N-3];
 * "site_pattern" int32 [nsite]; "const_char" int32 [P]; "params" double
 * [alpha, p_inv, kappa, omega] */
int plkh_workload_get(const plkh_workload* w, const char* name, void* out);

/* ---- PhyloTree-shaped driver --------------------------------------------- */
typedef struct plkh_tree plkh_tree;
/* tree over the workload's topology; ctx may be NULL (planning only) */
plkh_tree* plkh_tree_create(plk_context* ctx, const plkh_workload* w, int precision,
                            int safe_numeric, int* status);
void plkh_tree_destroy(plkh_tree* t);
int plkh_tree_instance(const plkh_tree* t);
int plkh_tree_num_branches(const plkh_tree* t);
/* branches in pre-order along the tree diameter (computeBestTraversal,
 * tree/phylotree.cpp:2689-2696); arrays of length num_branches */
int plkh_tree_branches(plkh_tree* t, int32_t* a, int32_t* b);
int plkh_tree_get_length(const plkh_tree* t, int a, int b, double* len);
int plkh_tree_set_length(plkh_tree* t, int a, int b, double len);  /* invalidates partials */
int plkh_tree_clear_all_partial(plkh_tree* t);
/* restore the branch lengths the tree was created with and clear all partials */
int plkh_tree_reset(plkh_tree* t);
/* computeTraversalInfo for branch (a, b) (tree/phylotree.cpp:5918-6028):
 * writes the post-order op list of stale partials (marking them computed)
 * and the branch ends as the engine expects them (node = partial slot,
 * dad = slot or tip). */
int plkh_tree_plan(plkh_tree* t, int a, int b, plk_op* ops, int cap, int* nops, plk_end* node,
                   plk_end* dad, double* length);
/* PhyloTree::computeLikelihood (tree/phylotree.cpp:1233-1270) */
int plkh_tree_compute_likelihood(plkh_tree* t, double* lnl, double* pattern_lh);
/* computeLikelihoodBranch on branch (a, b) */
int plkh_tree_likelihood_branch(plkh_tree* t, int a, int b, double* lnl, double* pattern_lh);
/* computeLikelihoodDerv at the current length of (a, b): theta is rebuilt
 * when the branch changes (PhyloTree::theta_computed) */
int plkh_tree_likelihood_derv(plkh_tree* t, int a, int b, double* df, double* ddf);
/* optimizeOneBranch(a, b, clearLH = true, max_nr) */
int plkh_tree_optimize_branch(plkh_tree* t, int a, int b, int max_nr, double* new_len,
                              int* n_derv);
/* optimizeAllBranches(iterations, tolerance, max_nr) (tree/phylotree.cpp:2698-2770);
 * lnl = final tree lnL, n_derv = derivative evaluations */
int plkh_tree_optimize_all_branches(plkh_tree* t, int iterations, double tolerance, int max_nr,
                                    double* lnl, int* n_derv);
/* the full post-order op list for branch (a, b), every internal node
 * included, no likelihood partial marked (for parsimony) */
int plkh_tree_plan_full(plkh_tree* t, int a, int b, plk_op* ops, int cap, int* nops, plk_end* node,
                        plk_end* dad);
/* computeParsimonyBranchFast at branch (a, b) (tree/phylotreepars.cpp:240-300):
 * the tree's frequency-weighted Fitch changes; branch_subst (nullable) the
 * changes on the branch itself. Only slots not holding the sets for the
 * needed direction are recomputed (the reference's partial_lh_computed & 2);
 * an NNI swap drops them all. */
int plkh_tree_parsimony_branch(plkh_tree* t, int a, int b, double* score, double* branch_subst);
/* RateGamma::setGammaShape + computeRates (model/rategamma.cpp:98-171): the
 * category rates of shape alpha (p_inv kept) into the tree's instance; every
 * partial becomes stale */
int plkh_tree_set_gamma_shape(plkh_tree* t, double alpha);
/* RateGamma::optimizeParameters (model/rategamma.cpp:232-246): Brent over the
 * shape in [0.02, 1000] from the current one, tolerance max(eps, 0.001), each
 * evaluation a full traversal on the device; the tree keeps the optimum.
 * *alpha the shape, *lnl the lnL there, *nevals the likelihood evaluations */
int plkh_tree_optimize_gamma_shape(plkh_tree* t, double eps, double* alpha, double* lnl, int* nevals);
/* optimize_gamma_shape of every tree (ModelFinder's candidate models of one
 * alignment, ModelFactory::optimizeParameters, model/modelfactory.cpp:
 * 1053-1135, or partitions; all on one context) advanced together: one
 * batched traversal launch per round for the trees that need a likelihood.
 * Each tree gets the shape, lnL and evaluation count it gets alone;
 * status[i] per tree, the return value the first error. */
int plkh_models_optimize_gamma_shape(plkh_tree* const* trees, int ntrees, double eps, double* alpha,
                                     double* lnl, int32_t* nevals, int32_t* status);
/* PhyloSuperTree::optimizeAllBranches (tree/phylosupertree.cpp:773-792):
 * optimizeAllBranches(iterations, tolerance / min(ntrees, 10), max_nr) on every
 * tree (all on one context), with the host Newton loop of each partition
 * advanced together: one batched traversal launch (plk_batch_evaluate) and
 * one batched derivative launch (plk_likelihood_derv_batch) per round for all
 * partitions. Each tree takes the Newton steps it takes alone. lnl, n_derv and
 * status (may be NULL) per tree; returns the first failing status. */
int plkh_supertree_optimize_all_branches(plkh_tree* const* trees, int ntrees, int iterations,
                                         double tolerance, int max_nr, double* lnl,
                                         int32_t* n_derv, int32_t* status);
/* derivative evaluations of each branch in the last optimize_all_branches
 * pass, in plkh_tree_branches order; *count = number of branches (values
 * beyond cap are not written) */
int plkh_tree_pass_counts(const plkh_tree* t, int32_t* n_derv, int cap, int* count);
/* neighbours of a node in adjacency order (deg <= 3) */
int plkh_tree_neighbors(const plkh_tree* t, int node, int32_t* nb, int* deg);
/* the length of (a, b) alone, no partial invalidated
 * (optimizeOneBranch(clearLH = false)) */
int plkh_tree_set_length_keep(plkh_tree* t, int a, int b, double len);

/* ---- NNI evaluation (getBestNNIForBran, tree/phylotree.cpp:4069-4300) ---- */
/* node1 and node2 (adjacent internal nodes) move to the instance's two NNI
 * buffers (slots nnodes - ntaxa and nnodes - ntaxa + 1, in the order the
 * reference hands out nni_partial_lh), every orientation of them stale */
int plkh_tree_nni_begin(plkh_tree* t, int node1, int node2);
/* exchange neighbour x of node1 with neighbour y of node2 (edges keep their
 * lengths; Node::updateNeighbor) */
int plkh_tree_nni_swap(plkh_tree* t, int node1, int x, int node2, int y);
/* PhyloNeighbor::clearPartialLh: the partial of node seen from `from` */
int plkh_tree_clear_partial(plkh_tree* t, int node, int from);
/* restores topology, lengths and the two nodes' buffers */
int plkh_tree_nni_end(plkh_tree* t);
/* inner branches (both ends internal) by edge id; *count = their number */
int plkh_tree_inner_branches(plkh_tree* t, int32_t* a, int32_t* b, int cap, int* count);
/* getBestNNIForBran(node1, node2) with nni5 and nni5_num_eval = 1: for both
 * moves (node1's first neighbour exchanged with node2's first, then second
 * neighbour) the tree lnL from the buffer lnl[2] and the five optimised
 * lengths lens[5 * move + k] (k = 0 the central branch, 1-2 node1's other
 * branches, 3-4 node2's, in adjacency order) with n_derv (nullable) the
 * derivative evaluations of each; topology and lengths are restored.
 * max_nr = NNI_MAX_NR_STEP (10 in the reference). One engine call per
 * derivative evaluation. */
int plkh_tree_nni_evaluate(plkh_tree* t, int node1, int node2, int max_nr, double* lnl,
                           double* lens, int32_t* n_derv);
/* plkh_tree_nni_evaluate for every inner branch, in plkh_tree_inner_branches
 * order, as one device-resident sweep (plk_optimize_sweep_ex): lnl [2m],
 * lens [10m], n_derv [10m] (nullable) for m inner branches; 4 and 20 states */
int plkh_tree_nni_evaluate_all(plkh_tree* t, int max_nr, double* lnl, double* lens,
                               int32_t* n_derv, int cap);
/* 1 (default): optimize_all_branches runs each pass as one device-resident
 * Newton sweep (plk_optimize_sweep); 0: minimizeNewton on the host, one
 * engine call per derivative evaluation */
int plkh_tree_set_device_newton(plkh_tree* t, int on);
/* min/max branch length used by optimizeOneBranch (defaults 1e-6, 10) */
int plkh_tree_set_branch_bounds(plkh_tree* t, double min_len, double max_len);

#ifdef __cplusplus
}
#endif

#endif /* PLK_HOST_H */
