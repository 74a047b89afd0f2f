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

    int slot(int v) const { return v - tree_.ntaxa; }
    int num_slots() const { return tree_.nnodes - tree_.ntaxa; }

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
    bool device_newton = true;
    void set_length(int a, int b, double len);
    void clear_all_partial();
    void reset();
    // PhyloNode::clearReversePartialLh (tree/phylonode.cpp:22-33)
    void clear_reverse(int node, int dad);

    int compute_likelihood(double* lnl, double* pattern_lh);
    int likelihood_branch(int a, int b, double* lnl, double* pattern_lh);
    int likelihood_derv(int a, int b, double* df, double* ddf);
    int optimize_branch(int a, int b, int max_nr, double* new_len, int* n_derv);
    int optimize_all_branches(int iterations, double tolerance, int max_nr, double* lnl, int* n_derv);
    int likelihood_from_buffer(double* lnl);

    double min_len = 1e-6, max_len = 10.0;

private:
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
    std::vector<int> size_;        // [nnodes*3] leaves below v away from adj[v][k]; 0 = unknown
    int theta_a_ = -1, theta_b_ = -1;   // branch whose theta buffer is current
    int cur_a_ = -1, cur_b_ = -1;       // current_it of PhyloTree
};

}  // namespace plk
