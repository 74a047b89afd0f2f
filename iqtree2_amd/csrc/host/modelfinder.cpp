// ModelFinder's per-model parameter loop for the Gamma shape
// (RateGamma::optimizeParameters, model/rategamma.cpp:232-246, called for
// every candidate model from ModelFactory::optimizeParameters,
// model/modelfactory.cpp:1053-1135) over many trees at once.
//
// The reference tests candidate models one after another (or one per
// OpenMP thread), each optimisation a sequence of full likelihood
// evaluations. Here every tree's Brent minimisation is a resumable state
// machine (brent.h) and each round evaluates every tree that asks for a
// likelihood in one plk_batch_evaluate launch. An item's arithmetic in a
// batch does not depend on the other items, so each tree ends at the shape,
// likelihood and evaluation count of optimize_gamma_shape (a batch of one).
#include <algorithm>
#include <vector>

#include "brent.h"
#include "phylotree_gpu.h"

namespace plk {

int PhyloTreeGPU::optimize_gamma_shape_batch(const std::vector<PhyloTreeGPU*>& trees, double eps, double* alpha,
                                             double* lnl, int* nevals, int* status) {
    const int n = (int)trees.size();
    if (n == 0) return PLK_OK;
    if (!(eps >= 0.0)) return PLK_ERR_ARG;
    plk_context* ctx = trees[0] ? trees[0]->ctx_ : nullptr;
    if (!ctx) return PLK_ERR_STATE;
    for (int i = 0; i < n; i++) {
        if (!trees[i] || trees[i]->ctx_ != ctx || trees[i]->inst_ < 0) return PLK_ERR_ARG;
        for (int k = 0; k < i; k++)
            if (trees[k] == trees[i]) return PLK_ERR_ARG;   // one state machine per tree
    }
    const double kMinShape = 0.02, kMaxShape = 1000.0, kTolShape = 0.001;   // utils/tools.h:600-602
    struct Run {
        BrentRun br;
        bool last = false;   // the closing evaluation at the optimum
        bool done = false;
        int n = 0;
        int status = PLK_OK;
        double lnl = 0.0;
    };
    std::vector<Run> runs(n);
    for (int i = 0; i < n; i++)
        runs[i].br.start(kMinShape, trees[i]->model_.alpha, kMaxShape, std::max(eps, kTolShape));
    std::vector<std::vector<plk_op>> ops(n);
    std::vector<plk_batch_item> items;
    std::vector<int> modes, who;
    std::vector<double> out;
    for (;;) {
        items.clear();
        modes.clear();
        who.clear();
        for (int i = 0; i < n; i++) {
            Run& r = runs[i];
            if (r.done) continue;
            PhyloTreeGPU* t = trees[i];
            // RateGamma::computeFunction: a new shape re-derives the rates
            const double x = r.last ? r.br.result() : r.br.x();
            if (x != t->model_.alpha) {
                const int st = t->set_gamma_shape(x);
                if (st != PLK_OK) {
                    r.status = st;
                    r.done = true;
                    continue;
                }
            }
            // computeLikelihood at the tree's current branch (compute_likelihood)
            if (t->cur_a_ < 0) {
                std::vector<int> height(t->tree_.nnodes, 0);
                const int leaf = t->farthest_leaf(0, -1, height);
                t->cur_a_ = leaf;
                t->cur_b_ = t->tree_.adj[3 * leaf];
            }
            plk_batch_item it{};
            it.instance = t->inst_;
            t->plan(t->cur_b_, t->cur_a_, ops[i], it.node, it.dad, it.length);
            it.ops = ops[i].data();
            it.nops = (int)ops[i].size();
            items.push_back(it);
            modes.push_back(PLK_BATCH_LNL);
            who.push_back(i);
        }
        if (items.empty()) break;
        out.assign(items.size(), 0.0);
        std::vector<int> st(items.size(), PLK_OK);
        const int s = plk_batch_evaluate(ctx, items.data(), modes.data(), (int)items.size(), out.data());
        if (s != PLK_OK && s != PLK_ERR_NONFINITE)
            // find the item the batch rejected: each on its own
            for (size_t k = 0; k < items.size(); k++) {
                const int sk = plk_batch_evaluate(ctx, &items[k], &modes[k], 1, &out[k]);
                st[k] = sk == PLK_ERR_NONFINITE ? PLK_OK : sk;
            }
        for (size_t k = 0; k < items.size(); k++) {
            Run& r = runs[who[k]];
            if (st[k] != PLK_OK) {
                r.status = st[k];
                r.done = true;
                continue;
            }
            r.n++;
            if (r.last) {
                r.lnl = out[k];
                r.done = true;
                continue;
            }
            r.br.feed(-out[k]);
            if (r.br.done()) r.last = true;
        }
    }
    int first_error = PLK_OK;
    for (int i = 0; i < n; i++) {
        if (alpha) alpha[i] = trees[i]->model_.alpha;
        if (lnl) lnl[i] = runs[i].lnl;
        if (nevals) nevals[i] = runs[i].n;
        if (status) status[i] = runs[i].status;
        if (runs[i].status != PLK_OK && first_error == PLK_OK) first_error = runs[i].status;
    }
    return first_error;
}

}  // namespace plk
