// PhyloSuperTree::optimizeAllBranches (tree/phylosupertree.cpp:773-792) with
// every partition's PhyloTree::optimizeAllBranches
// (tree/phylotree.cpp:2698-2773) advanced together on the device.
//
// The reference runs the partitions' passes side by side under OpenMP, one
// host thread each. Here every partition is a resumable state machine over
// the same steps as PhyloTreeGPU::optimize_all_branches_host (initial lnL;
// per branch in pre-order: theta, minimizeNewton, the divergence guard, the
// new length; lnL from the buffer at the end of the pass; revert or stop),
// and each round issues one engine call per kind of work for all partitions
// at once: one plk_batch_evaluate launch for every partition that needs a
// traversal (its initial lnL, or the theta buffer of its next branch) and one
// plk_likelihood_derv_batch launch for every derivative or lnL-from-buffer
// evaluation in progress. Both batched calls are bit-identical per item to
// their single-instance forms, so each partition takes the same Newton steps
// and ends with the same lengths as when optimised alone.
#include <cmath>
#include <vector>

#include "newton.h"
#include "phylotree_gpu.h"

namespace plk {

namespace {

struct Part {
    PhyloTreeGPU* t = nullptr;
    std::vector<int> nodes, nodes2;
    enum Phase { INIT, THETA, NEWTON, GUARD, PASS_END, DONE } phase = INIT;
    int it = 0;
    size_t j = 0;
    double tree_lh = 0.0, result = 0.0;
    std::vector<double> saved;
    NewtonRun nr;
    double len0 = 0.0, optx = 0.0;
    int total_derv = 0;
    int status = PLK_OK;
};

}  // namespace

int PhyloTreeGPU::optimize_all_branches_batch(const std::vector<PhyloTreeGPU*>& trees, int iterations,
                                              double tolerance, int max_nr, double* lnl, int* n_derv,
                                              int* status) {
    const int n = (int)trees.size();
    if (n == 0) return PLK_OK;
    if (iterations < 1 || max_nr < 1) return PLK_ERR_ARG;
    plk_context* ctx = trees[0] ? trees[0]->ctx_ : nullptr;
    if (!ctx) return PLK_ERR_STATE;
    for (int i = 0; i < n; i++) {
        if (!trees[i] || trees[i]->ctx_ != ctx || trees[i]->inst_ < 0) return PLK_ERR_ARG;
        for (int k = 0; k < i; k++)
            if (trees[k] == trees[i]) return PLK_ERR_ARG;   // one state machine per tree
    }
    // each partition's tolerance, as PhyloSuperTree passes it (phylosupertree.cpp:781)
    const double tol = tolerance / (n < 10 ? n : 10);
    std::vector<Part> parts(n);
    for (int i = 0; i < n; i++) {
        parts[i].t = trees[i];
        trees[i]->best_traversal(parts[i].nodes, parts[i].nodes2);
    }
    auto fail = [](Part& p, int st) {
        p.status = st;
        p.phase = Part::DONE;
    };
    auto start_pass = [](Part& p) {
        p.saved = p.t->tree_.len;
        p.t->pass_nd_.assign(p.nodes.size(), 0);
        p.j = 0;
        p.phase = Part::THETA;
    };
    // optimizeOneBranch's tail (tree/phylotree.cpp:2626-2644)
    auto finish_branch = [](Part& p) {
        PhyloTreeGPU* t = p.t;
        const int a = p.nodes[p.j], b = p.nodes2[p.j];
        if (p.optx != p.len0) {
            t->tree_.set_length(a, b, p.optx);
            t->clear_reverse(a, b);
            t->clear_reverse(b, a);
        }
        p.total_derv += p.nr.evals();
        t->pass_nd_[p.j] = p.nr.evals();
        p.j++;
        p.phase = p.j < p.nodes.size() ? Part::THETA : Part::PASS_END;
    };
    // the end of a pass of optimizeAllBranches (tree/phylotree.cpp:2729-2772)
    auto pass_end = [&](Part& p, double new_lh) {
        if (new_lh < p.tree_lh - tol * 0.1) {
            // the likelihood decreased: restore the lengths and stop
            p.t->clear_all_partial();
            p.t->tree_.len = p.saved;
            const int st = p.t->compute_likelihood(&new_lh, nullptr);
            p.result = new_lh;
            p.status = st == PLK_ERR_NONFINITE ? PLK_OK : st;
            p.phase = Part::DONE;
            return;
        }
        if (p.tree_lh <= new_lh && new_lh <= p.tree_lh + tol) {
            p.result = new_lh;
            p.phase = Part::DONE;
            return;
        }
        p.tree_lh = new_lh;
        if (++p.it < iterations) {
            start_pass(p);
        } else {
            p.result = p.tree_lh;
            p.phase = Part::DONE;
        }
    };

    std::vector<std::vector<plk_op>> ops(n);
    std::vector<plk_batch_item> items;
    std::vector<int> modes, who, insts, what;
    std::vector<double> out, lens, df, ddf, ol;
    for (;;) {
        // ---- traversals: initial lnL (computeLikelihoodBranch) or theta of the next branch
        items.clear();
        modes.clear();
        who.clear();
        for (int i = 0; i < n; i++) {
            Part& p = parts[i];
            if (p.phase != Part::INIT && p.phase != Part::THETA) continue;
            plk_batch_item it{};
            it.instance = p.t->inst_;
            if (p.phase == Part::INIT) {
                p.t->plan(p.nodes2[0], p.nodes[0], ops[i], it.node, it.dad, it.length);
                modes.push_back(PLK_BATCH_LNL);
            } else {
                const int a = p.nodes[p.j], b = p.nodes2[p.j];
                p.t->cur_a_ = a;
                p.t->cur_b_ = b;
                p.t->plan(b, a, ops[i], it.node, it.dad, it.length);
                p.len0 = it.length;
                modes.push_back(PLK_BATCH_THETA);
            }
            it.ops = ops[i].data();
            it.nops = (int)ops[i].size();
            items.push_back(it);
            who.push_back(i);
        }
        if (!items.empty()) {
            out.assign(items.size(), 0.0);
            const int st = plk_batch_evaluate(ctx, items.data(), modes.data(), (int)items.size(), out.data());
            std::vector<int> st_item(items.size(), st == PLK_ERR_NONFINITE ? PLK_OK : st);
            if (st != PLK_OK && st != PLK_ERR_NONFINITE)
                // find the failing items: each alone (a non-finite lnL is fixed
                // up per pattern and only warned about, phylokernelnew.h:3106-3132)
                for (size_t k = 0; k < items.size(); k++) {
                    const int sk = plk_batch_evaluate(ctx, &items[k], &modes[k], 1, &out[k]);
                    st_item[k] = sk == PLK_ERR_NONFINITE ? PLK_OK : sk;
                }
            for (size_t k = 0; k < items.size(); k++) {
                Part& p = parts[who[k]];
                if (st_item[k] != PLK_OK) {
                    p.t->clear_all_partial();
                    fail(p, st_item[k]);
                    continue;
                }
                if (p.phase == Part::INIT) {
                    p.tree_lh = out[k];
                    start_pass(p);
                } else {
                    p.t->theta_a_ = p.nodes[p.j];
                    p.t->theta_b_ = p.nodes2[p.j];
                    p.nr.start(p.len0, p.t->min_len, p.t->max_len, p.t->min_len, max_nr);
                    p.phase = Part::NEWTON;
                }
            }
        }
        // ---- derivative and lnL-from-buffer evaluations
        insts.clear();
        lens.clear();
        who.clear();
        what.clear();
        for (int i = 0; i < n; i++) {
            Part& p = parts[i];
            auto add = [&](double x, int w) {
                insts.push_back(p.t->inst_);
                lens.push_back(x);
                who.push_back(i);
                what.push_back(w);
            };
            if (p.phase == Part::NEWTON) {
                add(p.nr.x(), 0);
            } else if (p.phase == Part::GUARD) {
                // lnL at the length the last evaluation left, and at the old length
                add(p.nr.last_x(), 1);
                add(p.len0, 2);
            } else if (p.phase == Part::PASS_END) {
                PhyloTreeGPU* t = p.t;
                add(t->tree_.length(t->theta_a_, t->theta_b_), 3);
            }
        }
        if (insts.empty()) {
            // a round of traversals only (every partition between branches)
            bool active = false;
            for (const Part& p : parts) active = active || p.phase != Part::DONE;
            if (!active) break;
            continue;
        }
        const int m = (int)insts.size();
        df.assign(m, 0.0);
        ddf.assign(m, 0.0);
        ol.assign(m, 0.0);
        const int st = plk_likelihood_derv_batch(ctx, insts.data(), lens.data(), m, df.data(), ddf.data(), ol.data());
        if (st != PLK_OK && st != PLK_ERR_NONFINITE) {
            for (int k = 0; k < m; k++) {
                parts[who[k]].t->clear_all_partial();
                fail(parts[who[k]], st);
            }
            continue;
        }
        // lnL from the buffer alone when a batch item's derivative was not finite
        auto lnl_at = [&](int k, double& v) -> int {
            if (std::isfinite(ol[k])) {
                v = ol[k];
                return PLK_OK;
            }
            return (int)plk_likelihood_from_buffer(ctx, insts[k], lens[k], &v, nullptr);
        };
        for (int k = 0; k < m; k++) {
            Part& p = parts[who[k]];
            if (p.phase == Part::DONE) continue;
            if (what[k] == 0) {
                // computeFuncDerv: a non-finite derivative is an error of the
                // normal kernel; the safe kernel continues with df = ddf = 0
                // (tree/phylokernelnew.h:2577-2579, 2629-2632)
                if (!std::isfinite(ol[k]) && !p.t->safe_) {
                    p.t->clear_all_partial();
                    fail(p, PLK_ERR_NONFINITE);
                    continue;
                }
                p.nr.feed(-df[k], -ddf[k]);
                if (!p.nr.done()) continue;
                if (p.nr.failed()) {
                    p.t->clear_all_partial();
                    fail(p, PLK_ERR_NEWTON);
                    continue;
                }
                p.optx = p.nr.result();
                // divergence guard of optimizeOneBranch (tree/phylotree.cpp:2616-2625)
                if (p.optx > p.t->max_len * 0.95) p.phase = Part::GUARD;
                else finish_branch(p);
            } else if (what[k] == 1) {
                double opt_lh = 0.0, orig_lh = 0.0;
                int s = lnl_at(k, opt_lh);
                if (s == PLK_OK) s = lnl_at(k + 1, orig_lh);
                k++;   // the pair's second item
                if (s != PLK_OK) {
                    p.t->clear_all_partial();
                    fail(p, s);
                    continue;
                }
                if (orig_lh > opt_lh) p.optx = p.len0;
                finish_branch(p);
            } else if (what[k] == 3) {
                double new_lh = 0.0;
                const int s = lnl_at(k, new_lh);
                if (s != PLK_OK) {
                    p.t->clear_all_partial();
                    fail(p, s);
                    continue;
                }
                pass_end(p, new_lh);
            }
        }
    }
    int first = PLK_OK;
    for (int i = 0; i < n; i++) {
        if (lnl) lnl[i] = parts[i].result;
        if (n_derv) n_derv[i] = parts[i].total_derv;
        if (status) status[i] = parts[i].status;
        if (first == PLK_OK && parts[i].status != PLK_OK) first = parts[i].status;
    }
    return first;
}

}  // namespace plk
