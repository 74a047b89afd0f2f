/*
 * oracle.h - FP64 CPU restatement of IQ-TREE 2's reversible likelihood
 * kernels (tree/phylokernelnew.h) used as the parity oracle and as the
 * host-CPU baseline.
 *
 * TEST INFRASTRUCTURE ONLY: the product library never links or calls this.
 *
 * Layout follows the reference FMA kernel: patterns interleaved by V = 4
 * (element (ptn, c, x) at ptn/4*block*4 + (c*n + x)*4 + ptn%4,
 * tree/phylokernelnew.h:1606,1914-1916), uint16 scale counters per pattern or
 * per pattern and category (safe mode), OpenMP over 2*threads packets bounded
 * by computeBounds (tree/phylokernelnew.h:1112-1143).
 * The op / end structs are layout-identical to plk_op / plk_end (plk.h).
 */
#ifndef PLK_ORACLE_H
#define PLK_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int is_tip;
    int id;
    double length;
} orc_child;

typedef struct {
    int dad_slot;
    int nchildren;
    orc_child child[2];
} orc_op;

typedef struct {
    int is_tip;
    int id;
} orc_end;

typedef struct orc_tree orc_tree;

orc_tree* orc_create(int ntaxa, int nstates, int ncat, int npatterns, int nslots,
                     int state_unknown, int safe_numeric, int nthreads);
void orc_destroy(orc_tree* t);
void orc_set_threads(orc_tree* t, int nthreads);
void orc_set_patterns(orc_tree* t, const uint8_t* states, const double* freq,
                      const double* invar);
void orc_set_model(orc_tree* t, const double* eval, const double* evec, const double* inv_evec,
                   const double* tip_eigen, const double* rates, const double* props);
/* Lewis's ascertainment bias correction (ASC_VARIANT): the last nasc of the
 * npatterns patterns are the unobserved constant patterns (frequency 0);
 * lnL, pattern lnL and derivatives are conditioned on variable sites.
 * pattern_lh outputs then hold the npatterns - nasc observed patterns. */
void orc_set_asc(orc_tree* t, int nasc);
/* non-reversible and mixture models in the state basis (tree/phylokernelnonrev.h):
 * per mixture component m < nmix the complex eigensystem (re, im pairs: ceval
 * [nmix][n], cevec and cinv_evec [nmix][n][n]), the rate matrix [nmix][n][n]
 * and state_freq [nmix][n]; mix_weights [nmix]; the nrates site-rate
 * categories; tip [SU+1][n] the plain tip vectors (tip_partial_lh). The
 * instance's ncat must be nrates (fused_mix_rate, nmix == nrates) or
 * nrates * nmix. Unrooted trees put the frequencies at the node end of the
 * evaluated branch; rooted ones take them from the root leaf. Op children
 * and branch ends carry flags in is_tip: 1 tip, 2 toward the root (the
 * transposed matrix), 4 the root leaf. Returns 0, or -1 on a shape mismatch. */
int orc_set_model_nonrev(orc_tree* t, int nmix, int fused, const double* ceval, const double* cevec,
                         const double* cinv_evec, const double* rate_matrix, const double* state_freq,
                         const double* mix_weights, const double* site_rates, const double* site_props,
                         int nrates, const double* tip, int rooted);

void orc_update_partials(orc_tree* t, const orc_op* ops, int nops);
/* returns tree lnL; status (nullable) 0 ok, 4 non-finite fixed up */
double orc_likelihood_branch(orc_tree* t, const orc_op* ops, int nops, orc_end node, orc_end dad,
                             double length, double* pattern_lh, int* status);
void orc_likelihood_buffer(orc_tree* t, const orc_op* ops, int nops, orc_end node, orc_end dad);
/* returns 0, or 4 when df is not finite */
int orc_likelihood_derv(orc_tree* t, double length, double* df, double* ddf);
double orc_likelihood_from_buffer(orc_tree* t, double length, double* pattern_lh);
/* optimizeOneBranch / minimizeNewton; returns the new length, 0 / 7 in status */
double orc_optimize_branch(orc_tree* t, const orc_op* ops, int nops, orc_end node, orc_end dad,
                           double length, double min_len, double max_len, int max_nr,
                           double* lnl, int* n_derv, int* status);
void orc_read_partial(orc_tree* t, int slot, double* partial, int32_t* scale);

/* Fitch parsimony: ops give the state sets and subtree change counts of their
 * dad slots (kept apart from the partials), then the frequency-weighted
 * number of changes of the tree at (node, dad); ptn_score (nullable) [P]
 * each pattern's changes, branch_subst (nullable) the weighted changes on
 * the branch (node, dad) itself */
double orc_parsimony_branch(orc_tree* t, const orc_op* ops, int nops, orc_end node, orc_end dad,
                            int32_t* ptn_score, double* branch_subst);
/* the same score by the reference's bit-parallel algorithm (sites expanded by
 * frequency, planes of 32 sites, popcount; OpenMP over 256-site blocks);
 * -1 when a frequency is not an integer */
double orc_parsimony_fast(orc_tree* t, const orc_op* ops, int nops, orc_end node, orc_end dad,
                          double* branch_subst);

/* Sankoff parsimony (tree/phylotreepars.cpp:464-892): the cost matrix
 * [n][n] (cost[i][j], state i to j; seq_type 0 DNA, 1 protein, 2 other for
 * the ambiguity rules of the tip costs), then the frequency-weighted tree
 * score at (node, dad) after the ops; ptn_score (nullable) [P] each
 * pattern's score, branch_subst (nullable) the weighted cost on the branch
 * at the optimum */
void orc_set_cost_matrix(orc_tree* t, const uint32_t* cost, int seq_type);
double orc_parsimony_sankoff(orc_tree* t, const orc_op* ops, int nops, orc_end node, orc_end dad,
                             uint32_t* ptn_score, double* branch_subst);

/* minimizeNewton on a caller-supplied function (utils/optimization.cpp:422-499);
 * fderv(x, &f, &df) returns f and its derivative. status 0 ok, 7 no convergence */
typedef void (*orc_fderv)(void* user, double x, double* f, double* df);
double orc_minimize_newton(double x1, double xguess, double x2, double xacc, double* d2l,
                           int max_nr, orc_fderv fn, void* user, int* n_eval, int* status);

#ifdef __cplusplus
}
#endif

#endif
