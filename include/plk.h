/*
 * plk.h - C ABI of the MI355X Felsenstein-pruning likelihood engine.
 *
 * This is the drop-in boundary for IQ-TREE 2's likelihood kernel set: a
 * PhyloTree kernel set registered as LK_GPU (tree/phylotreesse.cpp:91-199,
 * tree/phylokernelfma.cpp:36-182) implements its four member-function
 * pointers with these calls. Plain pointers and sizes only; every function
 * returns a plk_status and never aborts the caller.
 *
 * Data contract (all host arrays FP64 unless stated):
 *   n   = number of states, C = rate categories, block = C*n
 *   P   = unique patterns (aln->size()), SU = STATE_UNKNOWN of the alignment
 *   states      uint8 [ntaxa][P]      converted tip states (0..SU)
 *   eval        [n]                   eigenvalues of Q
 *   evec        [n][n]                U (row x, column i)
 *   inv_evec    [n][n]                U^-1
 *   tip_eigen   [SU+1][n]             U^-1 * indicator(state)
 *   cat_rates   [C], cat_props [C]    site_rate->getRate / getProp
 * Partial likelihoods live in device memory in "slots" chosen by the caller
 * (a PhyloNeighbor's partial_lh maps to slot (ptr - base) / block), stored
 * pattern-major [P][C][n] in the eigen basis, with one scale counter per
 * pattern (or per pattern and category in safe mode).
 */
#ifndef PLK_H
#define PLK_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    PLK_OK = 0,
    PLK_ERR_ARG = 1,          /* invalid argument or shape */
    PLK_ERR_NOMEM = 2,        /* HBM budget or allocation exhausted */
    PLK_ERR_DEVICE = 3,       /* HIP runtime error */
    PLK_ERR_NONFINITE = 4,    /* non-finite likelihood/derivative (see error text) */
    PLK_ERR_STATE = 5,        /* call order violated (e.g. derivative without theta) */
    PLK_ERR_UNSUPPORTED = 6,  /* configuration not built into this library */
    PLK_ERR_NEWTON = 7        /* Newton-Raphson did not converge within max steps */
} plk_status;

typedef enum { PLK_PREC_F64 = 0, PLK_PREC_F32 = 1 } plk_precision;

typedef struct plk_context plk_context;

typedef struct {
    int device;           /* HIP device ordinal */
    void* stream;         /* hipStream_t to launch on; NULL = context-owned stream */
    size_t hbm_budget;    /* bytes this context may allocate; 0 = 90% of free HBM */
} plk_config;

typedef struct {
    int ntaxa;            /* leaves */
    int nstates;          /* n */
    int ncat;             /* C (1, 2, 4 or 8) */
    int npatterns;        /* P */
    int nslots;           /* partial-likelihood buffers */
    int state_unknown;    /* SU */
    int precision;        /* plk_precision */
    int safe_numeric;     /* 1: per-category scale counters (tree/phylotreesse.cpp:95-96) */
} plk_instance_desc;

/* one child of a traversal step: a tip (taxon id) or a partial slot */
typedef struct {
    int is_tip;
    int id;
    double length;        /* branch length node -> child */
} plk_child;

#define PLK_MAX_CHILDREN 2

/* TraversalInfo (tree/phylotree.h:320-331): compute slot dad_slot from its
 * children. Steps are executed in list order (post-order). A node of degree
 * k > 3 (the reference's multifurcating branch, tree/phylokernelnew.h:
 * 1405-1631: the product over all children) is planned as k - 2 steps, each
 * intermediate slot the child of the next step through a zero-length branch
 * (transition matrix = the identity), which gives the same product. */
typedef struct {
    int dad_slot;
    int nchildren;        /* 2 (binary nodes) */
    plk_child child[PLK_MAX_CHILDREN];
} plk_op;

/* end of a branch: a tip (taxon) or a partial slot */
typedef struct {
    int is_tip;
    int id;
} plk_end;

/* ---- context -------------------------------------------------------- */
plk_status plk_create(const plk_config* cfg, plk_context** out);
plk_status plk_destroy(plk_context* ctx);
/* last error text of the calling thread (static storage, never NULL) */
const char* plk_last_error(void);
/* bytes currently allocated by the context and its budget */
plk_status plk_memory(plk_context* ctx, size_t* used, size_t* budget);
/* stream the context launches on (hipStream_t) */
void* plk_stream(plk_context* ctx);
plk_status plk_synchronize(plk_context* ctx);

/* ---- instances (one PhyloTree's device state) ------------------------ */
/* HBM bytes an instance of this shape needs (admission by footprint) */
size_t plk_instance_footprint(const plk_instance_desc* desc);
plk_status plk_instance_create(plk_context* ctx, const plk_instance_desc* desc, int* instance);
plk_status plk_instance_destroy(plk_context* ctx, int instance);

/* computeTipPartialLikelihood inputs: tip states, ptn_freq, ptn_invar
 * (tree/phylotreesse.cpp:243-376, 537-617) */
plk_status plk_set_patterns(plk_context* ctx, int instance, const uint8_t* states,
                            const double* freq, const double* invar);
/* model eigensystem and rate categories (model/modelmarkov.cpp:1389-1602,
 * model/rategamma.cpp:98-171) */
plk_status plk_set_model(plk_context* ctx, int instance, const double* eval,
                         const double* evec, const double* inv_evec, const double* tip_eigen,
                         const double* cat_rates, const double* cat_props);

/* Non-reversible and mixture models: the state-basis kernel set of
 * tree/phylokernelnonrev.h (non-reversible models, and any model under
 * kernel_nonrev), selected by PhyloTree::setLikelihoodKernel when
 * !model->useRevKernel() (tree/phylotreeavx.cpp:105-125). Per mixture
 * component m < nmix: ModelMarkov's complex eigensystem as (re, im) pairs
 * (ceval [nmix][n], cevec and cinv_evec [nmix][n][n], model/modelmarkov.cpp:
 * 1300-1370), rate_matrix [nmix][n][n] and state_freq [nmix][n];
 * mix_weights [nmix] (getMixtureWeight); the nrates site-rate categories
 * (getRate, getProp); tip [SU+1][n] the plain tip vectors (tip_partial_lh,
 * tree/phylotreesse.cpp:387-420). ncat_mix is nrates with fused_mix_rate
 * (nmix == nrates), else nrates * nmix; the instance's ncat (1, 2, 4 or 8)
 * must be at least ncat_mix, the extra categories being weight-0 copies. rooted = the
 * tree's `rooted`: 0 puts the frequencies at the node end of the evaluated
 * branch, 1 takes them from the root leaf. Op children and branch ends then
 * carry flags in is_tip: 1 a tip, 2 (children) direction TOWARD_ROOT (the
 * transposed transition matrix), 4 the root leaf (with 1). 4 or 20 states;
 * plk_set_model switches back to the reversible kernels. Branch calls,
 * theta/derivative/from-buffer calls and plk_optimize_branch apply it;
 * +ASC applies as for reversible models (tree/phylokernelnonrev.h:1000-1047,
 * 1455-1515); batched traversals evaluate such items alone; the batched
 * derivative call and the device sweep refuse it (PLK_ERR_UNSUPPORTED). */
plk_status plk_set_model_nonrev(plk_context* ctx, int instance, int nmix, int fused,
                                const double* ceval, const double* cevec, const double* cinv_evec,
                                const double* rate_matrix, const double* state_freq,
                                const double* mix_weights, const double* site_rates,
                                const double* site_props, int nrates, const double* tip, int rooted);

/* Lewis's ascertainment bias correction, ASC_VARIANT and ASC_INFORMATIVE
 * (+ASC, +ASC_INF; model/modelfactory.cpp getASC, tree/phylokernelnew.h:
 * 2272-2273, 2580-2628, 3175-3221, 3385-3425): the last nasc of the instance's
 * npatterns patterns are the unobserved patterns
 * (Alignment::getUnobservedConstPatterns, alignment/alignment.cpp:4140-4196:
 * the constant patterns, or every parsimony-uninformative pattern), passed
 * with weight 0 and +I mass 0. Every lnL, pattern lnL and derivative is then
 * conditioned on the observable sites: lnL - nsite * log(1 - P(unobserved)),
 * with nsite the weight of the observed patterns. nasc = 0 turns it off. The batched
 * derivative call and the device sweep do not apply it (PLK_ERR_UNSUPPORTED);
 * batched traversals evaluate such items alone. pattern_lh outputs hold the
 * conditioned values of the observed patterns, then nasc zeros. */
plk_status plk_set_asc(plk_context* ctx, int instance, int nasc);

/* ---- kernels ---------------------------------------------------------- */
/* computePartialLikelihood over a post-order traversal list
 * (tree/phylokernelnew.h:1289-2039, with computePartialInfo 895-1110) */
plk_status plk_update_partials(plk_context* ctx, int instance, const plk_op* ops, int nops);

/* computeLikelihoodBranch (tree/phylokernelnew.h:2641-3223): run ops, then
 * the branch lnL between node (a slot) and dad (slot or tip) at `length`.
 * pattern_lh (nullable, host [P]) receives log-likelihood per pattern. */
plk_status plk_likelihood_branch(plk_context* ctx, int instance, const plk_op* ops, int nops,
                                 plk_end node, plk_end dad, double length, double* lnl,
                                 double* pattern_lh);

/* computeLikelihoodBufferSIMD (tree/phylokernelnew.h:2048-2231): run ops,
 * then theta = partial(node) * partial(dad) and buffer_scale_all. */
plk_status plk_likelihood_buffer(plk_context* ctx, int instance, const plk_op* ops, int nops,
                                 plk_end node, plk_end dad);

/* computeLikelihoodDerv on the theta buffer (tree/phylokernelnew.h:2233-2633):
 * +dlnL/dt and +d2lnL/dt2 at `length`; lnl (nullable) gets lnL at `length`. */
plk_status plk_likelihood_derv(plk_context* ctx, int instance, double length, double* df,
                               double* ddf, double* lnl);

/* computeLikelihoodFromBuffer (tree/phylokernelnew.h:3232-3428) */
plk_status plk_likelihood_from_buffer(plk_context* ctx, int instance, double length,
                                      double* lnl, double* pattern_lh);

/* optimizeOneBranch with Newton-Raphson (tree/phylotree.cpp:2594-2644,
 * utils/optimization.cpp:422-499): theta for (node, dad) after ops, then
 * minimizeNewton(min_len, length, max_len, xacc=min_len, max_nr); reports the
 * optimised length, lnL at it, and the number of derivative evaluations. */
plk_status plk_optimize_branch(plk_context* ctx, int instance, const plk_op* ops, int nops,
                               plk_end node, plk_end dad, double length, double min_len,
                               double max_len, int max_nr, double* out_length, double* out_lnl,
                               int* n_derv);

/* ---- batched evaluation (partitions, requests) -------------------------- */
/* one computeLikelihoodBranch per item, items on distinct instances */
typedef struct {
    int instance;
    const plk_op* ops;
    int nops;
    plk_end node;
    plk_end dad;
    double length;
} plk_batch_item;

/* PhyloSuperTree::computeLikelihood's partition loop
 * (tree/phylosupertree.cpp:701-735) as one persistent launch: every item's
 * traversal and branch lnL, instances of 4, 20 and 61 states with 4
 * categories and either precision, pattern tiles handed out largest-partition
 * first (4- and 20-state tiles share one launch, 61-state tiles take a launch
 * of their own in the same call); items of other category counts run one
 * traversal each in the same call. lnl[i] gets item i's tree lnL, equal to
 * plk_likelihood_branch on that item alone to rounding (1e-12 relative). */
plk_status plk_likelihood_branch_batch(plk_context* ctx, const plk_batch_item* items, int nitems,
                                       double* lnl);

/* Per-item choice of plk_batch_evaluate */
typedef enum {
    PLK_BATCH_LNL = 0,   /* computeLikelihoodBranch: the item's tree lnL */
    PLK_BATCH_THETA = 1  /* computeLikelihoodBuffer: theta of the item's branch */
} plk_batch_mode;

/* The same work list with a mode per item: traversal then branch lnL
 * (lnl[i] as plk_likelihood_branch_batch) or traversal then the theta
 * buffer (as plk_likelihood_buffer, ready for plk_likelihood_derv; lnl[i] =
 * 0), all items in one launch. modes may be NULL (all PLK_BATCH_LNL). */
plk_status plk_batch_evaluate(plk_context* ctx, const plk_batch_item* items, const int* modes,
                              int nitems, double* lnl);

/* computeLikelihoodDerv (tree/phylotreesse.cpp:220-224 ->
 * computeLikelihoodDervSIMD, tree/phylokernelnew.h:2233-2633) for n (instance,
 * length) pairs in one launch per shape, each on its instance's current
 * theta buffer; an instance may appear more than once. df/ddf/lnl[i] are
 * bit-identical to plk_likelihood_derv(instances[i], lengths[i]). Used to
 * advance the Newton iterations of many requests together. An item whose
 * derivative is not finite gets df = ddf = 0 and lnl = NaN, the others are
 * valid, and the call returns PLK_ERR_NONFINITE. */
plk_status plk_likelihood_derv_batch(plk_context* ctx, const int* instances, const double* lengths,
                                     int n, double* df, double* ddf, double* lnl);

/* plk_likelihood_derv_batch without any device allocation. No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). The batch's items, tile list, tile sums and results live in a
 * module-scope device array of 4 MiB per GPU (held until the process ends)
 * instead of the context's allocator; the call makes no hipMalloc, hipFree or
 * hipHostMalloc. Results are bit-identical to plk_likelihood_derv_batch. A
 * batch whose scratch would exceed 4 MiB (28 bytes per 256/C-pattern tile
 * of every item, plus about 200 bytes per item) is rejected with
 * PLK_ERR_NOMEM before anything runs. Every call on one GPU shares that GPU's
 * array: calls that overlap on one GPU (from different contexts or threads)
 * race; calls on one context are serialised by the context. */
plk_status plk_likelihood_derv_batch_noalloc(plk_context* ctx, const int* instances,
                                             const double* lengths, int n, double* df, double* ddf,
                                             double* lnl);

/* ---- device-resident branch optimisation ------------------------------- */
/* A traversal step whose children carry edge ids instead of lengths: the
 * lengths are read from the device copy of the branch-length array, which
 * the sweep updates as it optimises branches. */
typedef struct {
    int is_tip;
    int id;
    int edge;
} plk_sweep_child;

typedef struct {
    int dad_slot;
    int nchildren;
    plk_sweep_child child[PLK_MAX_CHILDREN];
} plk_sweep_op;

/* one branch of the sweep: ops[op_begin .. op_begin+op_count) refresh the
 * partials at its ends, then theta(node, dad) and Newton on edge `edge`;
 * lnl_index >= 0 asks for computeLikelihoodFromBuffer at the optimised
 * length right after this branch, stored in lnl_out[lnl_index]
 * (plk_optimize_sweep_ex), -1 for none */
typedef struct {
    int op_begin;
    int op_count;
    plk_end node;
    plk_end dad;
    int edge;
    int lnl_index;
} plk_sweep_branch;

/* One pass of PhyloTree::optimizeAllBranches over `branches` in order
 * (tree/phylotree.cpp:2698-2736): for each branch optimizeOneBranch's
 * minimizeNewton(min_len, len, max_len, xacc = min_len, max_nr) and the
 * divergence guard, entirely on the device. lengths[nedges] is read and
 * overwritten with the optimised lengths; n_derv (nullable, [nbranches])
 * receives derivative evaluations per branch; lnl receives
 * computeLikelihoodFromBuffer after the last branch. The op list must be the
 * one a host planner produces when every optimised branch invalidates the
 * partials that contain it (clearReversePartialLh). A pass of more than
 * PLK_SWEEP_MAX_BRANCHES branches, or one whose branch-length array does not
 * fit the kernel's LDS layout, returns PLK_ERR_UNSUPPORTED before any partial
 * is touched (the caller then runs minimizeNewton per branch with
 * plk_optimize_branch). A non-finite derivative stops the pass with
 * PLK_ERR_NONFINITE when safe_numeric is 0; with safe_numeric 1 it counts as
 * df = ddf = 0 and the pass continues, as the reference's safe kernel does.
 * After any status other than PLK_OK the partials of the instance are
 * undefined. */
#define PLK_SWEEP_MAX_BRANCHES 4096

/* The same sweep with more outputs, for callers that optimise a planned
 * sequence of branches rather than one pass over a tree - the NNI
 * evaluation of getBestNNIForBran (tree/phylotree.cpp:4069-4300): per move,
 * the five branches around the swapped edge with optimizeOneBranch(clearLH =
 * false, NNI_MAX_NR_STEP), then computeLikelihoodFromBuffer. Ops that write
 * the NNI buffers and edge ids that give each evaluation its own copy of
 * the five lengths are the caller's plan. branch_len (nullable,
 * [nbranches]) receives each branch's optimised length; lnl_out [nlnl]
 * receives the lnL of every branch with lnl_index >= 0. A non-finite lnL
 * there ends the sweep with PLK_ERR_NONFINITE. */
plk_status plk_optimize_sweep_ex(plk_context* ctx, int instance, const plk_sweep_op* ops, int nops,
                                 const plk_sweep_branch* branches, int nbranches, double* lengths,
                                 int nedges, double min_len, double max_len, int max_nr,
                                 int* n_derv, double* branch_len, double* lnl_out, int nlnl,
                                 double* lnl);
plk_status plk_optimize_sweep(plk_context* ctx, int instance, const plk_sweep_op* ops, int nops,
                              const plk_sweep_branch* branches, int nbranches, double* lengths,
                              int nedges, double min_len, double max_len, int max_nr,
                              int* n_derv, double* lnl);

/* RELL (tree/iqtree.cpp:3568-3625, tree/phylokernel.h:64-70): out[b] =
 * sum_p weights[b][p] * pattern_lh[p] over the last evaluated pattern_lh. */
plk_status plk_rell(plk_context* ctx, int instance, const float* weights, int nboot,
                    double* out);

/* ---- instrumentation (no counterpart in the reference) ------------------ */
/* Per-kernel aggregates over the launches made between start and stop:
 * launches, device time from HIP events recorded around each launch on the
 * context's stream, and the algorithmic HBM bytes and flops those launches
 * must move (per-pattern unit cost x patterns x steps). */
typedef struct {
    char name[32];
    int launches;
    double total_ms;
    double bytes;
    double flops;
} plk_kernel_stat;

/* Sankoff parsimony with a cost matrix (PhyloTree::initCostMatrix /
 * loadCostMatrixFile, computeTipPartialParsimony, computePartialParsimonySankoff
 * and computeParsimonyBranchSankoff, tree/phylotreepars.cpp:464-885):
 * cost [n][n] (state i to j, unsigned), 4, 20 or 61 states; the tip costs follow
 * the reference's rules (a state's row of the matrix, an ambiguity set 0
 * inside and the least cost into it outside, STATE_UNKNOWN 0). The branch
 * call runs the ops (every listed node's costs, kept apart from the Fitch
 * sets and the partial likelihoods), then returns the frequency-weighted
 * tree score at (node, dad) and, in branch_subst (nullable), the weighted
 * cost on that branch at the first minimum. Binary nodes. */
plk_status plk_set_cost_matrix(plk_context* ctx, int instance, const uint32_t* cost);
plk_status plk_parsimony_sankoff(plk_context* ctx, int instance, const plk_op* ops, int nops,
                                 plk_end node, plk_end dad, double* score, double* branch_subst);

/* Fitch parsimony (PhyloTree::computePartialParsimonyFast and
 * computeParsimonyBranchFast, tree/phylotreepars.cpp:32-300; the SIMD kernel
 * tree/phylokernel.h:1465-2094): ops compute the Fitch state sets and subtree
 * change counts of their dad slots (kept apart from the likelihood partials,
 * same slot numbers; a full post-order list for a first evaluation, then only
 * the slots whose direction changed), then *score = sum_p freq_p * changes_p
 * of the tree seen across the branch (node, dad) -- the reference's integer
 * score for integer frequencies -- and *branch_subst (nullable) the changes
 * on that branch alone (the reference's branch_subst). Tip sets: the observed
 * state, the ambiguity set (DNA IUPAC, protein B/Z/J) or every state for
 * STATE_UNKNOWN. DNA, protein and codon instances. */
plk_status plk_parsimony_branch(plk_context* ctx, int instance, const plk_op* ops, int nops,
                                plk_end node, plk_end dad, double* score, double* branch_subst);

/* HBM bandwidth of a plain device copy of `bytes` (read + written per launch,
 * reps launches after one warm-up), timed with HIP events on the context's
 * stream: the measured roofline denominator beside the 8 TB/s peak, and the
 * known-size kernel that calibrates the FETCH_SIZE / WRITE_SIZE counters.
 * Measurement support; no counterpart in the reference. */
plk_status plk_copy_bandwidth(plk_context* ctx, size_t bytes, int reps, double* gbps, double* avg_ms);

plk_status plk_profile_start(plk_context* ctx);
plk_status plk_profile_stop(plk_context* ctx, plk_kernel_stat* stats, int cap, int* nstats);

/* ---- inspection --------------------------------------------------------- */
/* copy one slot back as FP64 [P][C][n] and its scale counters
 * (uint16 [P] or [P][C]; the FP32 form reports its exponent counters) */
plk_status plk_read_partial(plk_context* ctx, int instance, int slot, double* partial,
                            int32_t* scale);

#ifdef __cplusplus
}
#endif

#endif /* PLK_H */
