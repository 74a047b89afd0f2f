// Seeded synthetic workloads: random unrooted binary trees, random model
// parameters, sequences simulated down the tree in FP64, then compressed to
// patterns. This is the shared input of the device engine and of the CPU
// restatement it is checked against.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "alignment.h"
#include "model.h"
#include "rng.h"

namespace plk {

// Unrooted tree: leaves 0..ntaxa-1, internal nodes ntaxa..nnodes-1.
struct Tree {
    int ntaxa = 0;
    int nnodes = 0;
    std::vector<int> deg;                    // [nnodes]
    std::vector<int> adj;                    // [nnodes][3]
    std::vector<double> len;                 // [nnodes][3] (symmetric)
    std::vector<int> edge;                   // [nnodes][3] edge id of each adjacency slot
    std::vector<int> eu, ev;                 // [nnodes-1] endpoints of each edge id
    int nbranch() const { return nnodes - 1; }
    int edge_id(int u, int v) const { return edge[3 * u + neighbor_index(u, v)]; }
    int neighbor_index(int v, int u) const {
        for (int k = 0; k < deg[v]; k++)
            if (adj[3 * v + k] == u) return k;
        return -1;
    }
    double length(int u, int v) const { return len[3 * u + neighbor_index(u, v)]; }
    void set_length(int u, int v, double t) {
        len[3 * u + neighbor_index(u, v)] = t;
        len[3 * v + neighbor_index(v, u)] = t;
    }
    std::string newick() const;
};

// Random taxon addition onto uniformly chosen branches; Exp(mean) lengths.
Tree random_tree(int ntaxa, double mean_len, Rng& rng);
// Build from an edge list (u[i], v[i], t[i]) of a binary or star tree.
Tree tree_from_edges(int ntaxa, int nedges, const int* u, const int* v, const double* t);

struct ModelParams {
    SeqType type = SEQ_DNA;
    int nstates = 4;
    int ncat = 4;
    double alpha = 0.5;
    double p_inv = 0.0;
    double kappa = 1.0, omega = 1.0;    // codon only
    std::vector<double> exch;           // upper triangle
    std::vector<double> pi;             // [n]
    // derived
    std::vector<double> Q, eval, evec, inv_evec, rates, props, tip;
    void derive();                      // fills the derived arrays
};

// Random parameters: DNA GTR exchangeabilities U(0.5,2), pi ~ Dir(5^4);
// protein exchangeabilities Gamma(1,1), pi ~ Dir(2^20); codon GY94 with
// kappa ~ U(1,4), omega ~ Gamma(2, 0.25) and equal codon frequencies.
ModelParams random_model(SeqType t, int ncat, double alpha, double p_inv, Rng& rng);

// Simulate nsite columns down the tree; returns nseq rows of characters
// (3 per site for codons). gap_frac of DNA tip cells become '-', 'N' or a
// two/three-state IUPAC code.
std::vector<char> simulate(const Tree& tree, const ModelParams& m, int nsite, double gap_frac,
                           Rng& rng);

// A complete likelihood input: tree, model, patterns, padded frequency and
// invariant-site vectors.
struct Workload {
    Tree tree;
    ModelParams model;
    PatternSet patterns;
    std::vector<double> freq;   // [npattern] as doubles
    std::vector<double> invar;  // [npattern]
};

struct WorkloadSpec {
    int type = 0;          // SeqType
    int ntaxa = 16;
    int nsite = 1000;
    int ncat = 4;
    double mean_len = 0.1;
    double alpha = 0.5;    // <= 0: draw U(0.2, 2.0)
    double p_inv = 0.0;
    double gap_frac = 0.0;
    uint64_t seed = 1;
};

int make_workload(const WorkloadSpec& spec, Workload& w);
// Same but on a caller-given tree (partitions share one tree).
int make_workload_on_tree(const WorkloadSpec& spec, const Tree& tree, Workload& w);

}  // namespace plk
