#include "synth.h"

#include <cmath>
#include <cstdio>
#include <functional>
#include <sstream>
#include <utility>

namespace plk {

static void add_edge(Tree& t, int u, int v, double len) {
    const int e = (int)t.eu.size();
    t.eu.push_back(u);
    t.ev.push_back(v);
    t.adj[3 * u + t.deg[u]] = v;
    t.len[3 * u + t.deg[u]] = len;
    t.edge[3 * u + t.deg[u]] = e;
    t.deg[u]++;
    t.adj[3 * v + t.deg[v]] = u;
    t.len[3 * v + t.deg[v]] = len;
    t.edge[3 * v + t.deg[v]] = e;
    t.deg[v]++;
}

Tree tree_from_edges(int ntaxa, int nedges, const int* u, const int* v, const double* tl) {
    Tree t;
    t.ntaxa = ntaxa;
    t.nnodes = nedges + 1;
    t.deg.assign(t.nnodes, 0);
    t.adj.assign(3 * t.nnodes, -1);
    t.len.assign(3 * t.nnodes, 0.0);
    t.edge.assign(3 * t.nnodes, -1);
    for (int e = 0; e < nedges; e++) add_edge(t, u[e], v[e], tl[e]);
    return t;
}

Tree random_tree(int ntaxa, double mean_len, Rng& rng) {
    // edges as (u, v) pairs; start from the 3-taxon star around node ntaxa
    std::vector<std::pair<int, int>> edges;
    int next_internal = ntaxa;
    const int c0 = next_internal++;
    for (int k = 0; k < 3 && k < ntaxa; k++) edges.push_back({c0, k});
    for (int k = 3; k < ntaxa; k++) {
        const size_t e = (size_t)rng.below(edges.size());
        const int a = edges[e].first, b = edges[e].second;
        const int w = next_internal++;
        edges[e] = {a, w};
        edges.push_back({w, b});
        edges.push_back({w, k});
    }
    std::vector<int> u(edges.size()), v(edges.size());
    std::vector<double> tl(edges.size());
    for (size_t e = 0; e < edges.size(); e++) {
        u[e] = edges[e].first;
        v[e] = edges[e].second;
        tl[e] = rng.exponential(mean_len);
    }
    return tree_from_edges(ntaxa, (int)edges.size(), u.data(), v.data(), tl.data());
}

std::string Tree::newick() const {
    std::ostringstream os;
    os.precision(17);
    std::function<void(int, int)> rec = [&](int v, int dad) {
        if (v < ntaxa) {
            os << "T" << v;
            return;
        }
        os << "(";
        bool first = true;
        for (int k = 0; k < deg[v]; k++) {
            const int c = adj[3 * v + k];
            if (c == dad) continue;
            if (!first) os << ",";
            first = false;
            rec(c, v);
            os << ":" << len[3 * v + k];
        }
        os << ")";
    };
    rec(ntaxa, -1);
    os << ";";
    return os.str();
}

void ModelParams::derive() {
    const int n = nstates;
    Q.assign((size_t)n * n, 0.0);
    eval.assign(n, 0.0);
    evec.assign((size_t)n * n, 0.0);
    inv_evec.assign((size_t)n * n, 0.0);
    build_reversible_q(n, exch.data(), pi.data(), Q.data());
    decompose_reversible(n, Q.data(), pi.data(), eval.data(), evec.data(), inv_evec.data());
    rates.assign(ncat, 0.0);
    props.assign(ncat, 0.0);
    gamma_mean_rates(ncat, alpha, p_inv, rates.data(), props.data());
    tip.assign((size_t)(seq_state_unknown(type) + 1) * n, 0.0);
    tip_eigen_vectors(type, inv_evec.data(), tip.data());
}

ModelParams random_model(SeqType t, int ncat, double alpha, double p_inv, Rng& rng) {
    ModelParams m;
    m.type = t;
    m.nstates = seq_num_states(t);
    m.ncat = ncat;
    m.alpha = alpha;
    m.p_inv = p_inv;
    const int n = m.nstates;
    m.exch.assign((size_t)n * (n - 1) / 2, 0.0);
    m.pi.assign(n, 0.0);
    if (t == SEQ_DNA) {
        for (auto& r : m.exch) r = rng.uniform(0.5, 2.0);
        double a[4] = {5, 5, 5, 5};
        rng.dirichlet(4, a, m.pi.data());
    } else if (t == SEQ_PROTEIN) {
        for (auto& r : m.exch) r = rng.gamma(1.0, 1.0);
        std::vector<double> a(20, 2.0);
        rng.dirichlet(20, a.data(), m.pi.data());
    } else {
        m.kappa = rng.uniform(1.0, 4.0);
        m.omega = rng.gamma(2.0, 0.25);
        gy94_exchangeabilities(m.kappa, m.omega, m.exch.data());
        for (auto& p : m.pi) p = 1.0 / 61.0;
    }
    m.derive();
    return m;
}

static const char kDnaChars[] = "ACGT";
static const char kProtChars[] = "ARNDCQEGHILKMFPSTWYV";
static const char kGapChars[] = "-NRYWSMKBHDV";

std::vector<char> simulate(const Tree& tree, const ModelParams& m, int nsite, double gap_frac,
                           Rng& rng) {
    const int n = m.nstates, C = m.ncat;
    const int nn = tree.nnodes;
    // pre-order from the first internal node (or leaf 0 for tiny trees)
    const int root = tree.nnodes > tree.ntaxa ? tree.ntaxa : 0;
    std::vector<int> order, parent(nn, -1), pidx(nn, -1);
    std::vector<int> stack = {root};
    std::vector<char> seen(nn, 0);
    seen[root] = 1;
    while (!stack.empty()) {
        const int v = stack.back();
        stack.pop_back();
        order.push_back(v);
        for (int k = 0; k < tree.deg[v]; k++) {
            const int c = tree.adj[3 * v + k];
            if (seen[c]) continue;
            seen[c] = 1;
            parent[c] = v;
            pidx[c] = k;
            stack.push_back(c);
        }
    }
    // CDF of P(r_c t) rows for every (node, category)
    std::vector<double> cdf((size_t)nn * C * n * n, 0.0);
    std::vector<double> P((size_t)n * n);
    for (int v = 0; v < nn; v++) {
        if (parent[v] < 0) continue;
        const double t = tree.len[3 * parent[v] + pidx[v]];
        for (int c = 0; c < C; c++) {
            for (int x = 0; x < n; x++)
                for (int y = 0; y < n; y++) {
                    double acc = 0.0;
                    for (int i = 0; i < n; i++)
                        acc += m.evec[x * n + i] * std::exp(m.eval[i] * m.rates[c] * t) *
                               m.inv_evec[i * n + y];
                    P[x * n + y] = acc < 0.0 ? 0.0 : acc;
                }
            double* out = &cdf[(((size_t)v * C + c) * n) * n];
            for (int x = 0; x < n; x++) {
                double run = 0.0;
                for (int y = 0; y < n; y++) {
                    run += P[x * n + y];
                    out[x * n + y] = run;
                }
            }
        }
    }
    std::vector<double> pi_cdf(n);
    {
        double run = 0.0;
        for (int x = 0; x < n; x++) { run += m.pi[x]; pi_cdf[x] = run; }
    }
    const int width = (m.type == SEQ_CODON) ? 3 : 1;
    const int nchar = nsite * width;
    std::vector<char> out((size_t)tree.ntaxa * nchar);
    std::vector<int> st(nn);
    const int* s2c = sense_to_codon();
    for (int site = 0; site < nsite; site++) {
        const bool invariant = m.p_inv > 0.0 && rng.uniform() < m.p_inv;
        const int c = invariant ? -1 : (int)rng.below(C);
        st[root] = rng.categorical_cdf(n, pi_cdf.data());
        for (size_t k = 1; k < order.size(); k++) {
            const int v = order[k];
            const int ps = st[parent[v]];
            if (invariant) {
                st[v] = ps;
            } else {
                st[v] = rng.categorical_cdf(n, &cdf[((((size_t)v * C + c) * n) + ps) * n]);
            }
        }
        for (int taxon = 0; taxon < tree.ntaxa; taxon++) {
            char* cell = &out[(size_t)taxon * nchar + (size_t)site * width];
            const int s = st[taxon];
            if (m.type == SEQ_DNA) {
                cell[0] = kDnaChars[s];
                if (gap_frac > 0.0 && rng.uniform() < gap_frac)
                    cell[0] = kGapChars[rng.below(sizeof(kGapChars) - 1)];
            } else if (m.type == SEQ_PROTEIN) {
                cell[0] = kProtChars[s];
            } else {
                const int codon = s2c[s];
                cell[0] = kDnaChars[(codon >> 4) & 3];
                cell[1] = kDnaChars[(codon >> 2) & 3];
                cell[2] = kDnaChars[codon & 3];
            }
        }
    }
    return out;
}

int make_workload_on_tree(const WorkloadSpec& spec, const Tree& tree, Workload& w) {
    Rng rng(spec.seed);
    const SeqType t = (SeqType)spec.type;
    double alpha = spec.alpha;
    if (alpha <= 0.0) alpha = rng.uniform(0.2, 2.0);
    w.tree = tree;
    w.model = random_model(t, spec.ncat, alpha, spec.p_inv, rng);
    std::vector<char> chars = simulate(tree, w.model, spec.nsite, spec.gap_frac, rng);
    const int width = (t == SEQ_CODON) ? 3 : 1;
    int rc = build_patterns(t, tree.ntaxa, spec.nsite * width, chars.data(), w.patterns);
    if (rc != 0) return rc;
    const int np = w.patterns.npattern;
    w.freq.assign(np, 0.0);
    for (int p = 0; p < np; p++) w.freq[p] = (double)w.patterns.freq[p];
    w.invar.assign(np, 0.0);
    compute_ptn_invar(t, w.patterns, w.model.pi.data(), spec.p_inv, np, w.invar.data());
    return 0;
}

int make_workload(const WorkloadSpec& spec, Workload& w) {
    // the tree comes from its own stream so that model/sequence draws do not
    // depend on the topology generator
    Rng trng(spec.seed ^ 0x7265657472656574ULL);
    Tree tree = random_tree(spec.ntaxa, spec.mean_len, trng);
    return make_workload_on_tree(spec, tree, w);
}

}  // namespace plk
