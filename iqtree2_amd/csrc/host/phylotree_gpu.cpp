#include "phylotree_gpu.h"

#include "model.h"

#include <algorithm>
#include <cmath>

namespace plk {

PhyloTreeGPU::PhyloTreeGPU(plk_context* ctx, const Workload& w, int precision, int safe_numeric)
    : tree_(w.tree), initial_len_(w.tree.len), ctx_(ctx), model_(w.model) {
    computed_.assign((size_t)tree_.nnodes * 3, 0);
    owner_.assign(tree_.nnodes, -1);
    pars_dir_.assign(tree_.nnodes, -1);
    size_.assign((size_t)tree_.nnodes * 3, 0);
    const ModelParams& m = w.model;
    nstates_ = m.nstates;
    if (!ctx_) return;
    plk_instance_desc d{};
    d.ntaxa = tree_.ntaxa;
    d.nstates = m.nstates;
    d.ncat = m.ncat;
    d.npatterns = w.patterns.npattern;
    d.nslots = num_slots();
    d.state_unknown = seq_state_unknown(m.type);
    d.precision = precision;
    // PhyloTree::setLikelihoodKernel (tree/phylotreesse.cpp:94-95): per-category
    // scaling on request, from 2000 taxa (utils/tools.cpp:1262), and for every
    // state count other than 4 and 20
    d.safe_numeric = safe_numeric || tree_.ntaxa >= 2000 || (m.nstates != 4 && m.nstates != 20);
    safe_ = d.safe_numeric;
    status_ = plk_instance_create(ctx_, &d, &inst_);
    if (status_ != PLK_OK) return;
    status_ = plk_set_patterns(ctx_, inst_, w.patterns.states.data(), w.freq.data(), w.invar.data());
    if (status_ != PLK_OK) return;
    status_ = plk_set_model(ctx_, inst_, m.eval.data(), m.evec.data(), m.inv_evec.data(), m.tip.data(),
                            m.rates.data(), m.props.data());
}

PhyloTreeGPU::~PhyloTreeGPU() {
    if (ctx_ && inst_ >= 0) plk_instance_destroy(ctx_, inst_);
}

int PhyloTreeGPU::subtree_size(int node, int dad) {
    if (node < tree_.ntaxa) return 1;
    const int k = tree_.neighbor_index(node, dad);
    int& s = size_[(size_t)node * 3 + k];
    if (s > 0) return s;
    int total = 0;
    for (int j = 0; j < tree_.deg[node]; j++) {
        const int ch = tree_.adj[3 * node + j];
        if (ch != dad) total += subtree_size(ch, node);
    }
    s = total;
    return s;
}

// computeTraversalInfo(dad_branch = dad->node, dad) (tree/phylotree.cpp:5918-6028)
void PhyloTreeGPU::plan_rec(int node, int dad, std::vector<plk_op>& ops, std::vector<int>* edges) {
    if (node < tree_.ntaxa) return;
    const int k = tree_.neighbor_index(node, dad);
    if (computed_[(size_t)node * 3 + k]) return;
    int children[2], nc = 0;
    for (int j = 0; j < tree_.deg[node]; j++) {
        const int ch = tree_.adj[3 * node + j];
        if (ch != dad && nc < 2) children[nc++] = ch;
    }
    // larger subtree first (5931-5941): keeps the live set of a post-order small
    if (nc == 2 && subtree_size(children[0], node) < subtree_size(children[1], node))
        std::swap(children[0], children[1]);
    for (int i = 0; i < nc; i++) plan_rec(children[i], node, ops, edges);
    computed_[(size_t)node * 3 + k] = 1;
    // one buffer per internal node, taken over by the orientation in use
    // (reorientPartialLh 5898-5912, MemSlotVector::takeover)
    if (owner_[node] != k) {
        if (owner_[node] >= 0) computed_[(size_t)node * 3 + owner_[node]] = 0;
        owner_[node] = k;
    }
    plk_op op{};
    op.dad_slot = slot(node);
    op.nchildren = nc;
    for (int i = 0; i < nc; i++) {
        const int ch = children[i];
        op.child[i].is_tip = ch < tree_.ntaxa;
        op.child[i].id = ch < tree_.ntaxa ? ch : slot(ch);
        op.child[i].length = tree_.length(node, ch);
        if (edges) edges->push_back(eid(node, ch));
    }
    ops.push_back(op);
}

void PhyloTreeGPU::plan_full_rec(int node, int dad, std::vector<plk_op>& ops, bool only_invalid) {
    if (node < tree_.ntaxa || (only_invalid && pars_dir_[node] == dad)) return;
    int children[2], nc = 0;
    for (int j = 0; j < tree_.deg[node]; j++) {
        const int ch = tree_.adj[3 * node + j];
        if (ch != dad && nc < 2) children[nc++] = ch;
    }
    if (nc == 2 && subtree_size(children[0], node) < subtree_size(children[1], node))
        std::swap(children[0], children[1]);
    for (int i = 0; i < nc; i++) plan_full_rec(children[i], node, ops, only_invalid);
    plk_op op{};
    op.dad_slot = slot(node);
    op.nchildren = nc;
    for (int i = 0; i < nc; i++) {
        const int ch = children[i];
        op.child[i].is_tip = ch < tree_.ntaxa;
        op.child[i].id = ch < tree_.ntaxa ? ch : slot(ch);
        op.child[i].length = tree_.length(node, ch);
    }
    ops.push_back(op);
    if (only_invalid) pars_dir_[node] = dad;   // the op about to run fills it
}

void PhyloTreeGPU::plan_full(int a, int b, std::vector<plk_op>& ops, plk_end& node, plk_end& dad) {
    int nd = a, dd = b;
    if (nd < tree_.ntaxa) std::swap(nd, dd);
    ops.clear();
    plan_full_rec(nd, dd, ops);
    plan_full_rec(dd, nd, ops);
    node = nd < tree_.ntaxa ? plk_end{1, nd} : plk_end{0, slot(nd)};
    dad = dd < tree_.ntaxa ? plk_end{1, dd} : plk_end{0, slot(dd)};
}

int PhyloTreeGPU::parsimony_branch(int a, int b, double* score, double* branch_subst) {
    if (!ctx_) return PLK_ERR_STATE;
    int nd = a, dd = b;
    if (nd < tree_.ntaxa) std::swap(nd, dd);
    std::vector<plk_op> ops;
    plan_full_rec(nd, dd, ops, true);
    plan_full_rec(dd, nd, ops, true);
    const plk_end node = nd < tree_.ntaxa ? plk_end{1, nd} : plk_end{0, slot(nd)};
    const plk_end dad = dd < tree_.ntaxa ? plk_end{1, dd} : plk_end{0, slot(dd)};
    const int st = plk_parsimony_branch(ctx_, inst_, ops.data(), (int)ops.size(), node, dad, score,
                                        branch_subst);
    if (st != PLK_OK) std::fill(pars_dir_.begin(), pars_dir_.end(), -1);
    return st;
}

void PhyloTreeGPU::plan_sweep(const std::vector<int>& nodes, const std::vector<int>& nodes2,
                              std::vector<plk_sweep_op>& sops,
                              std::vector<plk_sweep_branch>& branches) {
    sops.clear();
    branches.clear();
    for (size_t j = 0; j < nodes.size(); j++) {
        const int n1 = nodes[j], n2 = nodes2[j];
        // computeLikelihoodDerv(current_it = n1->n2, n1): node = n2, dad = n1
        int nd = n2, dd = n1;
        if (nd < tree_.ntaxa) std::swap(nd, dd);
        std::vector<plk_op> ops;
        std::vector<int> edges;
        plan_rec(nd, dd, ops, &edges);
        plan_rec(dd, nd, ops, &edges);
        plk_sweep_branch B{};
        B.op_begin = (int)sops.size();
        B.op_count = (int)ops.size();
        B.node = plk_end{0, slot(nd)};
        B.dad = dd < tree_.ntaxa ? plk_end{1, dd} : plk_end{0, slot(dd)};
        B.edge = tree_.edge_id(n1, n2);
        B.lnl_index = -1;
        for (size_t o = 0; o < ops.size(); o++) {
            plk_sweep_op so{};
            so.dad_slot = ops[o].dad_slot;
            so.nchildren = ops[o].nchildren;
            for (int k = 0; k < ops[o].nchildren; k++) {
                so.child[k].is_tip = ops[o].child[k].is_tip;
                so.child[k].id = ops[o].child[k].id;
                so.child[k].edge = edges[2 * o + k];
            }
            sops.push_back(so);
        }
        branches.push_back(B);
        // optimizeOneBranch(clearLH = true) after a length change
        clear_reverse(n1, n2);
        clear_reverse(n2, n1);
    }
}

void PhyloTreeGPU::plan(int a, int b, std::vector<plk_op>& ops, plk_end& node, plk_end& dad,
                        double& len) {
    int nd = a, dd = b;
    if (nd < tree_.ntaxa) std::swap(nd, dd);   // the leaf goes to dad (2654-2661)
    ops.clear();
    plan_rec(nd, dd, ops);
    plan_rec(dd, nd, ops);
    node = plk_end{0, slot(nd)};
    dad = dd < tree_.ntaxa ? plk_end{1, dd} : plk_end{0, slot(dd)};
    len = tree_.length(a, b);
}

void PhyloTreeGPU::clear_reverse(int node, int dad) {
    for (int j = 0; j < tree_.deg[node]; j++) {
        const int x = tree_.adj[3 * node + j];
        if (x == dad) continue;
        computed_[(size_t)node * 3 + j] = 0;
        clear_reverse(x, node);
    }
}

void PhyloTreeGPU::clear_all_partial() {
    std::fill(computed_.begin(), computed_.end(), 0);
    theta_a_ = theta_b_ = -1;
}

void PhyloTreeGPU::reset() {
    tree_.len = initial_len_;
    clear_all_partial();
    cur_a_ = cur_b_ = -1;
}

void PhyloTreeGPU::set_length(int a, int b, double len) {
    if (tree_.length(a, b) == len) return;
    tree_.set_length(a, b, len);
    clear_reverse(a, b);
    clear_reverse(b, a);
    theta_a_ = theta_b_ = -1;
}

int PhyloTreeGPU::farthest_leaf(int node, int dad, std::vector<int>& height) {
    if (dad >= 0 && node < tree_.ntaxa) {
        height[node] = 0;
        return node;
    }
    int res = -1;
    height[node] = 0;
    for (int j = 0; j < tree_.deg[node]; j++) {
        const int ch = tree_.adj[3 * node + j];
        if (ch == dad) continue;
        const int leaf = farthest_leaf(ch, node, height);
        if (height[node] < height[ch] + 1) {
            height[node] = height[ch] + 1;
            res = leaf;
        }
    }
    return res;
}

void PhyloTreeGPU::preorder(int node, int dad, const std::vector<int>& height,
                            std::vector<int>& nodes, std::vector<int>& nodes2) {
    if (dad >= 0) {
        nodes.push_back(node);
        nodes2.push_back(dad);
    }
    std::vector<int> nb(tree_.adj.begin() + 3 * node, tree_.adj.begin() + 3 * node + tree_.deg[node]);
    std::stable_sort(nb.begin(), nb.end(), [&](int x, int y) { return height[x] < height[y]; });
    for (int ch : nb)
        if (ch != dad) preorder(ch, node, height, nodes, nodes2);
}

// MTree::findFarthestLeaf twice + getPreOrderBranches (tree/mtree.cpp:3040-3081)
void PhyloTreeGPU::best_traversal(std::vector<int>& nodes, std::vector<int>& nodes2) {
    std::vector<int> height(tree_.nnodes, 0);
    const int far1 = farthest_leaf(0, -1, height);
    const int far2 = farthest_leaf(far1, -1, height);
    (void)far2;
    nodes.clear();
    nodes2.clear();
    preorder(far1, -1, height, nodes, nodes2);
}

int PhyloTreeGPU::likelihood_branch(int a, int b, double* lnl, double* pattern_lh) {
    if (!ctx_) return PLK_ERR_STATE;
    std::vector<plk_op> ops;
    plk_end node, dad;
    double len;
    plan(a, b, ops, node, dad, len);
    return plk_likelihood_branch(ctx_, inst_, ops.data(), (int)ops.size(), node, dad, len, lnl,
                                 pattern_lh);
}

int PhyloTreeGPU::compute_likelihood(double* lnl, double* pattern_lh) {
    if (cur_a_ < 0) {
        std::vector<int> height(tree_.nnodes, 0);
        const int leaf = farthest_leaf(0, -1, height);
        cur_a_ = leaf;
        cur_b_ = tree_.adj[3 * leaf];
    }
    // computeLikelihoodBranch(current_it, current_it_back->node)
    return likelihood_branch(cur_b_, cur_a_, lnl, pattern_lh);
}

int PhyloTreeGPU::likelihood_derv(int a, int b, double* df, double* ddf) {
    if (!ctx_) return PLK_ERR_STATE;
    int st;
    if (!((theta_a_ == a && theta_b_ == b) || (theta_a_ == b && theta_b_ == a))) {
        std::vector<plk_op> ops;
        plk_end node, dad;
        double len;
        plan(b, a, ops, node, dad, len);
        st = plk_likelihood_buffer(ctx_, inst_, ops.data(), (int)ops.size(), node, dad);
        if (st != PLK_OK) return st;
        theta_a_ = a;
        theta_b_ = b;
    }
    return plk_likelihood_derv(ctx_, inst_, tree_.length(a, b), df, ddf, nullptr);
}

int PhyloTreeGPU::likelihood_from_buffer(double* lnl) {
    if (!ctx_) return PLK_ERR_STATE;
    if (theta_a_ < 0) return PLK_ERR_STATE;
    return plk_likelihood_from_buffer(ctx_, inst_, tree_.length(theta_a_, theta_b_), lnl, nullptr);
}

// PhyloTree::optimizeOneBranch (tree/phylotree.cpp:2594-2644)
int PhyloTreeGPU::optimize_branch(int n1, int n2, int max_nr, double* new_len, int* n_derv, bool clear) {
    if (!ctx_) return PLK_ERR_STATE;
    cur_a_ = n1;
    cur_b_ = n2;
    std::vector<plk_op> ops;
    plk_end node, dad;
    double len;
    plan(n2, n1, ops, node, dad, len);
    double optx = len;
    int nd = 0;
    const int st = plk_optimize_branch(ctx_, inst_, ops.data(), (int)ops.size(), node, dad, len,
                                       min_len, max_len, max_nr, &optx, nullptr, &nd);
    if (st != PLK_OK) return st;
    theta_a_ = n1;
    theta_b_ = n2;
    if (optx != len) {
        tree_.set_length(n1, n2, optx);
        if (clear) {
            clear_reverse(n1, n2);
            clear_reverse(n2, n1);
        }
    }
    if (new_len) *new_len = optx;
    if (n_derv) *n_derv = nd;
    return PLK_OK;
}

// PhyloTree::optimizeAllBranches(iterations, tolerance, maxNRStep) (2698-2773),
// every pass over the branches one device-resident Newton sweep
int PhyloTreeGPU::optimize_all_branches(int iterations, double tolerance, int max_nr, double* lnl,
                                        int* n_derv) {
    if (!ctx_) return PLK_ERR_STATE;
    // the device sweep is built for 4, 20 and 61 states (61 with at most 4
    // rate categories; other shapes report PLK_ERR_UNSUPPORTED and take the
    // host loop below)
    if (!device_newton || (nstates_ != 4 && nstates_ != 20 && nstates_ != 61))
        return optimize_all_branches_host(iterations, tolerance, max_nr, lnl, n_derv);
    // a sweep holds a pass of at most PLK_SWEEP_MAX_BRANCHES branches
    if (tree_.nbranch() > PLK_SWEEP_MAX_BRANCHES)
        return optimize_all_branches_host(iterations, tolerance, max_nr, lnl, n_derv);
    std::vector<int> nodes, nodes2;
    best_traversal(nodes, nodes2);
    double tree_lh;
    // a non-finite lnL is fixed up per pattern and only warned about
    // (tree/phylokernelnew.h:3106-3132)
    int st = likelihood_branch(nodes2[0], nodes[0], &tree_lh, nullptr);
    if (st != PLK_OK && st != PLK_ERR_NONFINITE) return st;
    int total_derv = 0;
    const int ne = tree_.nbranch();
    for (int it = 0; it < iterations; it++) {
        std::vector<double> saved(tree_.len);
        std::vector<plk_sweep_op> sops;
        std::vector<plk_sweep_branch> br;
        plan_sweep(nodes, nodes2, sops, br);
        std::vector<double> lens(ne);
        for (int e = 0; e < ne; e++) lens[e] = tree_.length(tree_.eu[e], tree_.ev[e]);
        std::vector<int> nd(br.size());
        double new_lh = 0.0;
        st = plk_optimize_sweep(ctx_, inst_, sops.data(), (int)sops.size(), br.data(), (int)br.size(),
                                lens.data(), ne, min_len, max_len, max_nr, nd.data(), &new_lh);
        if (st != PLK_OK) {
            // plan_sweep marked the planned partials computed; a sweep that did
            // not run, or stopped early, leaves some of them stale or rebuilt
            // with lengths the host does not hold
            clear_all_partial();
            cur_a_ = cur_b_ = -1;
            if (st == PLK_ERR_UNSUPPORTED && it == 0) {
                // shapes the sweep is not built for (its LDS layout): the host loop
                return optimize_all_branches_host(iterations, tolerance, max_nr, lnl, n_derv);
            }
            return st;
        }
        for (int e = 0; e < ne; e++) tree_.set_length(tree_.eu[e], tree_.ev[e], lens[e]);
        for (int x : nd) total_derv += x;
        pass_nd_ = nd;
        theta_a_ = nodes.back();
        theta_b_ = nodes2.back();
        cur_a_ = theta_a_;
        cur_b_ = theta_b_;
        if (new_lh < tree_lh - tolerance * 0.1) {
            clear_all_partial();
            tree_.len = saved;
            st = compute_likelihood(&new_lh, nullptr);
            if (lnl) *lnl = new_lh;
            if (n_derv) *n_derv = total_derv;
            return st;
        }
        if (tree_lh <= new_lh && new_lh <= tree_lh + tolerance) {
            tree_lh = new_lh;
            break;
        }
        tree_lh = new_lh;
    }
    if (lnl) *lnl = tree_lh;
    if (n_derv) *n_derv = total_derv;
    return PLK_OK;
}

// the same pass with minimizeNewton on the host, one engine call per branch
int PhyloTreeGPU::optimize_all_branches_host(int iterations, double tolerance, int max_nr,
                                             double* lnl, int* n_derv) {
    if (!ctx_) return PLK_ERR_STATE;
    std::vector<int> nodes, nodes2;
    best_traversal(nodes, nodes2);
    double tree_lh;
    // a non-finite lnL is fixed up per pattern and only warned about
    // (tree/phylokernelnew.h:3106-3132)
    int st = likelihood_branch(nodes2[0], nodes[0], &tree_lh, nullptr);
    if (st != PLK_OK && st != PLK_ERR_NONFINITE) return st;
    int total_derv = 0;
    for (int it = 0; it < iterations; it++) {
        std::vector<double> saved(tree_.len);
        pass_nd_.assign(nodes.size(), 0);
        for (size_t j = 0; j < nodes.size(); j++) {
            int nd = 0;
            st = optimize_branch(nodes[j], nodes2[j], max_nr, nullptr, &nd);
            total_derv += nd;
            pass_nd_[j] = nd;
            if (st != PLK_OK) return st;
        }
        double new_lh;
        st = likelihood_from_buffer(&new_lh);
        if (st != PLK_OK) return st;
        if (new_lh < tree_lh - tolerance * 0.1) {
            // revert branch lengths when the likelihood decreased (2737-2760)
            clear_all_partial();
            tree_.len = saved;
            st = compute_likelihood(&new_lh, nullptr);
            if (lnl) *lnl = new_lh;
            if (n_derv) *n_derv = total_derv;
            return st;
        }
        if (tree_lh <= new_lh && new_lh <= tree_lh + tolerance) {
            tree_lh = new_lh;
            break;
        }
        tree_lh = new_lh;
    }
    if (lnl) *lnl = tree_lh;
    if (n_derv) *n_derv = total_derv;
    return PLK_OK;
}

// ---- NNI evaluation ----------------------------------------------------------

int PhyloTreeGPU::nni_begin(int node1, int node2) {
    const Tree& t = tree_;
    if (nni_active() || node1 < t.ntaxa || node2 < t.ntaxa || node1 >= t.nnodes || node2 >= t.nnodes ||
        t.deg[node1] != 3 || t.deg[node2] != 3 || t.neighbor_index(node1, node2) < 0)
        return PLK_ERR_ARG;
    nni_saved_tree_ = tree_;
    std::fill(pars_dir_.begin(), pars_dir_.end(), -1);   // slot mapping changes
    for (int k = 0; k < 3; k++) {
        nni_saved_comp_[0][k] = computed_[(size_t)node1 * 3 + k];
        nni_saved_comp_[1][k] = computed_[(size_t)node2 * 3 + k];
    }
    nni_saved_owner_[0] = owner_[node1];
    nni_saved_owner_[1] = owner_[node2];
    // NNI buffers in the order getBestNNIForBran hands them out: the saved
    // neighbours are node1->node2, node2->node1, then node1's and node2's
    // other neighbours' back pointers, and the first that owns a buffer gets
    // the first NNI buffer; node2's buffer is first only when the
    // orientation it holds is the one seen from node1
    const bool node2_first = owner_[node2] >= 0 && owner_[node2] == t.neighbor_index(node2, node1);
    nni_node_[0] = node1;
    nni_node_[1] = node2;
    nni_slot_[0] = node2_first ? 1 : 0;
    nni_slot_[1] = node2_first ? 0 : 1;
    for (int k = 0; k < 3; k++) {
        computed_[(size_t)node1 * 3 + k] = 0;
        computed_[(size_t)node2 * 3 + k] = 0;
    }
    owner_[node1] = owner_[node2] = -1;
    theta_a_ = theta_b_ = -1;
    return PLK_OK;
}

int PhyloTreeGPU::nni_swap(int node1, int x, int node2, int y) {
    Tree& t = tree_;
    const int k1 = t.neighbor_index(node1, x), k2 = t.neighbor_index(node2, y);
    if (!nni_active() || k1 < 0 || k2 < 0 || x == node2 || y == node1) return PLK_ERR_ARG;
    std::fill(pars_dir_.begin(), pars_dir_.end(), -1);   // topology changes
    const int kx = t.neighbor_index(x, node1), ky = t.neighbor_index(y, node2);
    const int e1 = t.edge[3 * node1 + k1], e2 = t.edge[3 * node2 + k2];
    const double l1 = t.len[3 * node1 + k1], l2 = t.len[3 * node2 + k2];
    t.adj[3 * node1 + k1] = y;
    t.adj[3 * node2 + k2] = x;
    t.adj[3 * x + kx] = node2;
    t.adj[3 * y + ky] = node1;
    t.edge[3 * node1 + k1] = e2;
    t.len[3 * node1 + k1] = l2;
    t.edge[3 * node2 + k2] = e1;
    t.len[3 * node2 + k2] = l1;
    t.eu[e1] = x;
    t.ev[e1] = node2;
    t.eu[e2] = y;
    t.ev[e2] = node1;
    // the subtree sizes seen from the swapped nodes are stale
    std::fill(size_.begin(), size_.end(), 0);
    // every orientation of node1 and node2 contains the exchanged subtrees;
    // x's and y's partials seen from their new neighbour are their own
    // subtrees, unchanged
    for (int k = 0; k < 3; k++) {
        computed_[(size_t)node1 * 3 + k] = 0;
        computed_[(size_t)node2 * 3 + k] = 0;
    }
    owner_[node1] = owner_[node2] = -1;
    theta_a_ = theta_b_ = -1;
    return PLK_OK;
}

void PhyloTreeGPU::clear_partial(int node, int from) {
    if (node < tree_.ntaxa) return;
    const int k = tree_.neighbor_index(node, from);
    if (k < 0) return;
    computed_[(size_t)node * 3 + k] = 0;
    if (owner_[node] == k) owner_[node] = -1;
}

void PhyloTreeGPU::nni_end() {
    if (!nni_active()) return;
    const int node1 = nni_node_[0], node2 = nni_node_[1];
    tree_ = nni_saved_tree_;
    std::fill(size_.begin(), size_.end(), 0);
    std::fill(pars_dir_.begin(), pars_dir_.end(), -1);
    for (int k = 0; k < 3; k++) {
        computed_[(size_t)node1 * 3 + k] = nni_saved_comp_[0][k];
        computed_[(size_t)node2 * 3 + k] = nni_saved_comp_[1][k];
    }
    owner_[node1] = nni_saved_owner_[0];
    owner_[node2] = nni_saved_owner_[1];
    nni_node_[0] = nni_node_[1] = -1;
    theta_a_ = theta_b_ = -1;
    cur_a_ = cur_b_ = -1;
}

void PhyloTreeGPU::inner_branches(std::vector<int>& a, std::vector<int>& b) {
    a.clear();
    b.clear();
    for (int e = 0; e < tree_.nbranch(); e++)
        if (tree_.eu[e] >= tree_.ntaxa && tree_.ev[e] >= tree_.ntaxa) {
            a.push_back(tree_.eu[e]);
            b.push_back(tree_.ev[e]);
        }
}

namespace {
// node's neighbours other than `except`, in adjacency order
std::vector<int> others(const Tree& t, int node, int except) {
    std::vector<int> v;
    for (int k = 0; k < t.deg[node]; k++)
        if (t.adj[3 * node + k] != except) v.push_back(t.adj[3 * node + k]);
    return v;
}
}  // namespace

int PhyloTreeGPU::nni_evaluate(int node1, int node2, int max_nr, double lnl[2], double lens[10],
                               int nd[10]) {
    if (!ctx_) return PLK_ERR_STATE;
    int st = nni_begin(node1, node2);
    if (st != PLK_OK) return st;
    const std::vector<int> n1 = others(tree_, node1, node2), n2 = others(tree_, node2, node1);
    for (int cnt = 0; cnt < 2 && st == PLK_OK; cnt++) {
        // the move exchanges node1's first neighbour with node2's cnt-th
        const int a = n1[0], c = n2[cnt];
        nni_swap(node1, a, node2, c);
        clear_partial(node2, node1);   // node12_it
        clear_partial(node1, node2);   // node21_it
        double* L = lens + 5 * cnt;
        int* D = nd + 5 * cnt;
        int i = 1;
        for (int x : others(tree_, node1, node2)) {
            clear_partial(node1, x);
            if ((st = optimize_branch(node1, x, max_nr, &L[i], &D[i], false)) != PLK_OK) break;
            i++;
        }
        if (st != PLK_OK) break;
        clear_partial(node1, node2);
        if ((st = optimize_branch(node1, node2, max_nr, &L[0], &D[0], false)) != PLK_OK) break;
        for (int y : others(tree_, node2, node1)) {
            clear_partial(node2, y);
            if ((st = optimize_branch(node2, y, max_nr, &L[i], &D[i], false)) != PLK_OK) break;
            i++;
        }
        if (st != PLK_OK) break;
        clear_partial(node2, node1);
        st = likelihood_from_buffer(&lnl[cnt]);
        if (st == PLK_ERR_NONFINITE && safe_) st = PLK_OK;   // fixed up per pattern
        nni_swap(node1, c, node2, a);
    }
    nni_end();
    return st;
}

int PhyloTreeGPU::nni_evaluate_all(int max_nr, std::vector<double>& lnl, std::vector<double>& lens,
                                   std::vector<int>& nd) {
    if (!ctx_) return PLK_ERR_STATE;
    if (nstates_ != 4 && nstates_ != 20 && nstates_ != 61) return PLK_ERR_UNSUPPORTED;
    std::vector<int> I1, I2;
    inner_branches(I1, I2);
    const int m = (int)I1.size(), ne = tree_.nbranch();
    // every inner branch's five edges get lengths of their own (ids ne + 5j ..),
    // initialised with the current lengths, so each evaluation starts from the
    // tree's lengths and the moves of one evaluation carry them over
    std::vector<double> L(ne + 5 * (size_t)m);
    for (int e = 0; e < ne; e++) L[e] = tree_.length(tree_.eu[e], tree_.ev[e]);
    std::vector<plk_sweep_op> sops;
    std::vector<plk_sweep_branch> br;
    edge_map_.assign(ne, -1);
    int st = PLK_OK;
    for (int j = 0; j < m && st == PLK_OK; j++) {
        const int node1 = I1[j], node2 = I2[j];
        if ((st = nni_begin(node1, node2)) != PLK_OK) break;
        const std::vector<int> n1 = others(tree_, node1, node2), n2 = others(tree_, node2, node1);
        const int own[5][2] = {{node1, node2}, {node1, n1[0]}, {node1, n1[1]}, {node2, n2[0]}, {node2, n2[1]}};
        for (int q = 0; q < 5; q++) {
            const int e = tree_.edge_id(own[q][0], own[q][1]);
            edge_map_[e] = ne + 5 * j + q;
            L[ne + 5 * j + q] = tree_.length(own[q][0], own[q][1]);
        }
        auto add = [&](int n_a, int n_b, int lnl_index) {
            // optimizeOneBranch(n_a, n_b): theta of (node = n_b side, dad = n_a side)
            int nd_ = n_b, dd = n_a;
            if (nd_ < tree_.ntaxa) std::swap(nd_, dd);
            std::vector<plk_op> ops;
            std::vector<int> edges;
            plan_rec(nd_, dd, ops, &edges);
            plan_rec(dd, nd_, ops, &edges);
            plk_sweep_branch B{};
            B.op_begin = (int)sops.size();
            B.op_count = (int)ops.size();
            B.node = plk_end{0, slot(nd_)};
            B.dad = dd < tree_.ntaxa ? plk_end{1, dd} : plk_end{0, slot(dd)};
            B.edge = eid(n_a, n_b);
            B.lnl_index = lnl_index;
            for (size_t o = 0; o < ops.size(); o++) {
                plk_sweep_op so{};
                so.dad_slot = ops[o].dad_slot;
                so.nchildren = ops[o].nchildren;
                for (int k = 0; k < ops[o].nchildren; k++) {
                    so.child[k].is_tip = ops[o].child[k].is_tip;
                    so.child[k].id = ops[o].child[k].id;
                    so.child[k].edge = edges[2 * o + k];
                }
                sops.push_back(so);
            }
            br.push_back(B);
        };
        for (int cnt = 0; cnt < 2; cnt++) {
            const int a = n1[0], c = n2[cnt];
            nni_swap(node1, a, node2, c);
            clear_partial(node2, node1);
            clear_partial(node1, node2);
            for (int x : others(tree_, node1, node2)) {
                clear_partial(node1, x);
                add(node1, x, -1);
            }
            clear_partial(node1, node2);
            add(node1, node2, -1);
            const std::vector<int> ys = others(tree_, node2, node1);
            for (size_t q = 0; q < ys.size(); q++) {
                clear_partial(node2, ys[q]);
                add(node2, ys[q], q + 1 == ys.size() ? 2 * j + cnt : -1);
            }
            clear_partial(node2, node1);
            nni_swap(node1, c, node2, a);
        }
        nni_end();
        for (int q = 0; q < 5; q++) edge_map_[tree_.edge_id(own[q][0], own[q][1])] = -1;
    }
    edge_map_.clear();
    if (st != PLK_OK) {
        clear_all_partial();
        return st;
    }
    std::vector<int> ndv(br.size());
    std::vector<double> brlen(br.size());
    lnl.assign(2 * (size_t)m, 0.0);
    double last = 0.0;
    st = plk_optimize_sweep_ex(ctx_, inst_, sops.data(), (int)sops.size(), br.data(), (int)br.size(),
                               L.data(), (int)L.size(), min_len, max_len, max_nr, ndv.data(),
                               brlen.data(), lnl.data(), 2 * m, &last);
    if (st != PLK_OK) {
        clear_all_partial();
        return st;
    }
    // per inner branch and move, the five branches in the order of
    // nni_evaluate's output: central, node1's, node2's
    lens.assign(10 * (size_t)m, 0.0);
    nd.assign(10 * (size_t)m, 0);
    for (int j = 0; j < m; j++)
        for (int cnt = 0; cnt < 2; cnt++) {
            const int b0 = 10 * j + 5 * cnt;   // sweep order: node1 x2, central, node2 x2
            const int order[5] = {2, 0, 1, 3, 4};
            for (int q = 0; q < 5; q++) {
                lens[b0 + q] = brlen[b0 + order[q]];
                nd[b0 + q] = ndv[b0 + order[q]];
            }
        }
    theta_a_ = theta_b_ = -1;
    return PLK_OK;
}


int PhyloTreeGPU::set_gamma_shape(double alpha) {
    if (!ctx_) return PLK_ERR_STATE;
    if (!(alpha > 0.0) || !std::isfinite(alpha)) return PLK_ERR_ARG;
    model_.alpha = alpha;
    gamma_mean_rates(model_.ncat, alpha, model_.p_inv, model_.rates.data(), model_.props.data());
    const int st = plk_set_model(ctx_, inst_, model_.eval.data(), model_.evec.data(), model_.inv_evec.data(),
                                 model_.tip.data(), model_.rates.data(), model_.props.data());
    clear_all_partial();
    return st;
}

int PhyloTreeGPU::optimize_gamma_shape(double eps, double* alpha, double* lnl, int* nevals) {
    // the batched driver with one tree: each evaluation is the same
    // plk_batch_evaluate call as in a batch of candidates, so a tree ends
    // where it ends in any batch, bit for bit
    int st = PLK_OK;
    const int r = optimize_gamma_shape_batch({this}, eps, alpha, lnl, nevals, &st);
    return r != PLK_OK ? r : st;
}

}  // namespace plk
