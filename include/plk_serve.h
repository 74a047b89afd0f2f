/* Request serving on top of the likelihood engine (plk.h): continuous
 * batching of likelihood, branch-optimisation and bootstrap-reweighting
 * requests of differing taxa and pattern counts on one device.
 *
 * No counterpart in the reference: IQ-TREE evaluates its candidate trees,
 * models and bootstrap replicates one PhyloTree at a time
 * (tree/phylotree.cpp:2594-2644 optimizeOneBranch, tree/iqtree.cpp:3568-3625
 * RELL). Each request here carries the inputs one such evaluation needs
 * (alignment patterns, eigen-decomposed model, the traversal that computes
 * the partials at both ends of one branch) and gets the same numbers that
 * call would give.
 *
 * A scheduler thread admits queued requests in arrival order while the
 * number of resident requests stays below max_inflight and their summed
 * plk_instance_footprint stays within hbm_budget, keeps each admitted
 * request's alignment and partials resident until it completes, and then
 * iterates: every request waiting for a traversal (its tree lnL, or the
 * theta buffer a branch optimisation starts from) joins one fused
 * plk_batch_evaluate launch; every branch optimisation in progress advances
 * by one minimizeNewton derivative evaluation through one
 * plk_likelihood_derv_batch launch. Requests that arrive meanwhile join at
 * the next iteration. A request's result does not depend on what else was
 * resident: the batched launches reduce every item in its own fixed order.
 * A failing request (invalid op list, out of memory, non-finite
 * derivative) completes with its own status and leaves the others alone. */
#ifndef PLK_SERVE_H
#define PLK_SERVE_H

#include <stddef.h>
#include <stdint.h>

#include "plk.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct plks_server plks_server;

typedef enum {
    PLKS_LNL = 0,     /* computeLikelihoodBranch: tree lnL */
    PLKS_NEWTON = 1,  /* optimizeOneBranch: Newton-optimised length and lnL */
    PLKS_RELL = 2     /* tree lnL and RELL sums over nboot pattern-weight vectors */
} plks_kind;

typedef struct {
    int max_inflight;   /* requests resident at once (0: 256) */
    size_t hbm_budget;  /* bytes of instance state resident at once (0: 3/4 of the
                           context budget) */
} plks_config;

typedef struct {
    plk_instance_desc desc;
    const uint8_t* states;      /* [ntaxa][npatterns] */
    const double* freq;         /* [npatterns] */
    const double* invar;        /* [npatterns] */
    const double* eval;         /* model, as plk_set_model */
    const double* evec;
    const double* inv_evec;
    const double* tip_eigen;
    const double* cat_rates;
    const double* cat_props;
    const plk_op* ops;          /* traversal computing the partials of node and dad */
    int nops;
    plk_end node;
    plk_end dad;
    double length;              /* current length of branch (node, dad) */
    int kind;                   /* plks_kind */
    double min_len, max_len;    /* PLKS_NEWTON bracket; xacc = min_len */
    int max_nr;                 /* PLKS_NEWTON iterations */
    const float* weights;       /* PLKS_RELL: [nboot][npatterns] */
    int nboot;
} plks_request;

typedef struct {
    int status;          /* plk_status of this request */
    double lnl;          /* tree lnL (PLKS_NEWTON: at the optimised length) */
    double length;       /* PLKS_NEWTON: optimised length */
    int n_derv;          /* PLKS_NEWTON: derivative evaluations */
    int nboot;           /* PLKS_RELL: values written to rell_out by plks_wait */
    double submit_s;     /* server clock, seconds since plks_create */
    double admit_s;
    double finish_s;
} plks_result;

typedef struct {
    long iterations;         /* scheduler iterations that did device work */
    long traversal_batches;  /* plk_batch_evaluate launches */
    long derv_batches;       /* plk_likelihood_derv_batch launches */
    long completed;
    int peak_inflight;
    size_t peak_bytes;
    double admit_s;          /* scheduler wall time: instance set-up of admitted requests */
    double traverse_s;       /* fused traversal launches (and RELL sums) */
    double derv_s;           /* batched derivative launches */
    double retire_s;         /* releasing completed requests */
} plks_stats;

plk_status plks_create(plk_context* ctx, const plks_config* cfg, plks_server** out);
/* completes every queued and resident request, then stops the scheduler;
 * results not collected are discarded */
plk_status plks_destroy(plks_server* s);
/* copies the request (the caller may free its arrays on return) and queues
 * it; PLK_ERR_ARG for a malformed request, which is not queued */
plk_status plks_submit(plks_server* s, const plks_request* r, uint64_t* ticket);
/* waits up to timeout_s seconds (< 0: no limit) for a request; *done = 1 and
 * the result (rell_out: nboot values for PLKS_RELL, may be NULL) when it has
 * completed, after which the ticket is released; *done = 0 otherwise */
plk_status plks_wait(plks_server* s, uint64_t ticket, double timeout_s, plks_result* res,
                     double* rell_out, int* done);
plk_status plks_get_stats(plks_server* s, plks_stats* out);
/* the server clock (seconds since plks_create) */
double plks_clock(plks_server* s);

#ifdef __cplusplus
}
#endif

#endif /* PLK_SERVE_H */
