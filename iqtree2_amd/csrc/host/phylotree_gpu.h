// PhyloTree-shaped host driver over the device engine: the part of
// PhyloTree that owns partial-likelihood buffers, plans traversals and runs
// branch-length Newton-Raphson (tree/phylotree.cpp), calling plk.h.
#pragma once
#include <vector>

#include "plk.h"
#include "synth.h"

namespace plk {

class PhyloTreeGPU {
public:
    // ctx may be null: the object then only plans traversals
    PhyloTreeGPU(plk_context* ctx, const Workload& w, int precision, int safe_numeric);
    ~PhyloTreeGPU();
    int status() const { return status_; }
    int instance() const { return inst_; }
    const Tree& tree() const { return tree_; }

    // LM_PER_NODE: one buffer per internal node, then the two NNI buffers
    // (tree/phylotree.cpp:1070-1076); during an NNI evaluation its two nodes
    // live in the NNI buffers (4127-4136)
    int slot(int v) const {
        if (v == nni_node_[0]) return num_central_slots() + nni_slot_[0];
        if (v == nni_node_[1]) return num_central_slots() + nni_slot_[1];
        return v - tree_.ntaxa;
    }
    int num_central_slots() const { return tree_.nnodes - tree_.ntaxa; }
    int num_slots() const { return num_central_slots() + 2; }

    // computeBestTraversal (tree/phylotree.cpp:2689-2696)
    void best_traversal(std::vector<int>& nodes, std::vector<int>& nodes2);
    // computeTraversalInfo for branch (a, b)
    void plan(int a, int b, std::vector<plk_op>& ops, plk_end& node, plk_end& dad, double& len);
    // the whole optimizeAllBranches pass as one device sweep: ops of every
    // branch in order, assuming each optimised branch changes its length
    void plan_sweep(const std::vector<int>& nodes, const std::vector<int>& nodes2,
                    std::vector<plk_sweep_op>& sops, std::vector<plk_sweep_branch>& branches);
    int optimize_all_branches_host(int iterations, double tolerance, int max_nr, double* lnl,
                                   int* n_derv);
    // PhyloSuperTree::optimizeAllBranches (tree/phylosupertree.cpp:773-792):
    // optimize_all_branches_host(iterations, tolerance / min(n, 10), max_nr)
    // on every tree of one context, all advanced together (supertree.cpp);
    // lnl, n_derv and status per tree; returns the first failing status
    static int optimize_all_branches_batch(const std::vector<PhyloTreeGPU*>& trees, int iterations,
                                           double tolerance, int max_nr, double* lnl, int* n_derv,
                                           int* status);
    bool device_newton = true;
    void set_length(int a, int b, double len);
    // the length alone, no partial invalidated (optimizeOneBranch with clearLH = false)
    void set_length_keep(int a, int b, double len) { tree_.set_length(a, b, len); }
    void clear_all_partial();
    void reset();
    // PhyloNode::clearReversePartialLh (tree/phylonode.cpp:22-33)
    void clear_reverse(int node, int dad);

    int compute_likelihood(double* lnl, double* pattern_lh);
    // ---- rate heterogeneity (model/rategamma.cpp) ----
    // RateGamma::setGammaShape + computeRates: the category rates of the new
    // shape (p_inv unchanged) into the instance's model, every partial stale
    int set_gamma_shape(double alpha);
    double gamma_shape() const { return model_.alpha; }
    // RateGamma::optimizeParameters (model/rategamma.cpp:232-246): Brent's
    // minimizeOneDimen over [0.02, 1000] from the current shape, tolerance
    // max(eps, 0.001), one full traversal per evaluation (a batch of one in
    // optimize_gamma_shape_batch); the shape is left at the optimum, *lnl is
    // the lnL there, *nevals the evaluations
    int optimize_gamma_shape(double eps, double* alpha, double* lnl, int* nevals);
    // optimize_gamma_shape of many trees on one context (ModelFinder's
    // candidate models of one alignment, or partitions) advanced together:
    // one batched traversal launch per round for every tree that needs a
    // likelihood, so each tree takes the same steps as when run alone
    static int optimize_gamma_shape_batch(const std::vector<PhyloTreeGPU*>& trees, double eps, double* alpha,
                                          double* lnl, int* nevals, int* status);
    int likelihood_branch(int a, int b, double* lnl, double* pattern_lh);
    int likelihood_derv(int a, int b, double* df, double* ddf);
    // optimizeOneBranch(a, b, clearLH, max_nr)
    int optimize_branch(int a, int b, int max_nr, double* new_len, int* n_derv, bool clear = true);
    int optimize_all_branches(int iterations, double tolerance, int max_nr, double* lnl, int* n_derv);
    int likelihood_from_buffer(double* lnl);
    // derivative evaluations per branch of the last optimizeAllBranches pass,
    // in best_traversal order
    const std::vector<int>& pass_counts() const { return pass_nd_; }

    // ---- NNI evaluation (getBestNNIForBran, tree/phylotree.cpp:4069-4300) ----
    // node1 and node2 (adjacent internal nodes) move to the NNI buffers, with
    // every orientation of them stale
    int nni_begin(int node1, int node2);
    // exchange neighbour x of node1 with neighbour y of node2; the edges keep
    // their lengths and ids (Node::updateNeighbor)
    int nni_swap(int node1, int x, int node2, int y);
    // PhyloNeighbor::clearPartialLh of the partial of `node` seen from `from`
    void clear_partial(int node, int from);
    // restores the topology, the lengths and node1 and node2's buffers
    void nni_end();
    bool nni_active() const { return nni_node_[0] >= 0; }
    // the two moves of getBestNNIForBran (nni5, nni5_num_eval 1): per move the
    // tree lnL from the buffer and the five optimised lengths (central branch
    // first, then node1's and node2's other branches); one engine call per
    // derivative evaluation
    int nni_evaluate(int node1, int node2, int max_nr, double lnl[2], double lens[10], int nd[10]);
    // inner branches (both ends internal), by edge id
    void inner_branches(std::vector<int>& a, std::vector<int>& b);
    // nni_evaluate for every inner branch in one device-resident sweep;
    // lnl [2m], lens [10m], nd [10m] for the m inner branches
    int nni_evaluate_all(int max_nr, std::vector<double>& lnl, std::vector<double>& lens,
                         std::vector<int>& nd);

    // ---- Fitch parsimony (tree/phylotreepars.cpp:32-300) ----
    // the full post-order op list for branch (a, b), every partial included
    // and no likelihood partial marked (parsimony sets are kept apart)
    void plan_full(int a, int b, std::vector<plk_op>& ops, plk_end& node, plk_end& dad);
    // computeParsimonyBranchFast at (a, b): frequency-weighted changes of
    // the tree and of the branch itself; only the slots not yet holding the
    // sets for the needed direction are recomputed (partial_lh_computed & 2)
    int parsimony_branch(int a, int b, double* score, double* branch_subst);

    double min_len = 1e-6, max_len = 10.0;

private:
    void plan_full_rec(int node, int dad, std::vector<plk_op>& ops, bool only_invalid = false);
    void plan_rec(int node, int dad, std::vector<plk_op>& ops, std::vector<int>* edges = nullptr);
    int subtree_size(int node, int dad);
    int farthest_leaf(int node, int dad, std::vector<int>& height);
    void preorder(int node, int dad, const std::vector<int>& height, std::vector<int>& nodes,
                  std::vector<int>& nodes2);

    Tree tree_;
    std::vector<double> initial_len_;
    plk_context* ctx_ = nullptr;
    int inst_ = -1;
    int status_ = 0;
    int nstates_ = 0;
    bool safe_ = false;
    std::vector<char> computed_;   // [nnodes*3]: partial of v seen from adj[v][k]
    std::vector<int> owner_;       // [nnodes]: orientation holding v's buffer
    std::vector<int> pars_dir_;    // [nnodes]: neighbour v's parsimony sets point away from; -1 none
    ModelParams model_;            // the instance's model (rates follow set_gamma_shape)
    std::vector<int> size_;        // [nnodes*3] leaves below v away from adj[v][k]; 0 = unknown
    int theta_a_ = -1, theta_b_ = -1;   // branch whose theta buffer is current
    int cur_a_ = -1, cur_b_ = -1;       // current_it of PhyloTree
    std::vector<int> pass_nd_;
    // NNI evaluation in progress
    int nni_node_[2] = {-1, -1}, nni_slot_[2] = {0, 1};
    Tree nni_saved_tree_;
    char nni_saved_comp_[2][3] = {{0, 0, 0}, {0, 0, 0}};
    int nni_saved_owner_[2] = {-1, -1};
    // sweep planning: edge id -> the id the sweep's ops use (-1: itself)
    std::vector<int> edge_map_;
    int eid(int u, int v) const {
        const int e = tree_.edge_id(u, v);
        return edge_map_.empty() || edge_map_[e] < 0 ? e : edge_map_[e];
    }
};

}  // namespace plk
