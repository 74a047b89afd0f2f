// extern "C" surface of the host library (include/plk_host.h).
#include <cstring>
#include <string>

#include "alignment.h"
#include "brent.h"
#include "model.h"
#include "phylotree_gpu.h"
#include "plk_host.h"
#include "synth.h"

using namespace plk;

struct plkh_patterns {
    PatternSet ps;
};
struct plkh_workload {
    Workload w;
};
struct plkh_tree {
    PhyloTreeGPU* t;
};

template <typename T>
static void put(void* out, const std::vector<T>& v) {
    std::memcpy(out, v.data(), v.size() * sizeof(T));
}

static bool valid_type(int t) { return t == PLKH_DNA || t == PLKH_PROTEIN || t == PLKH_CODON; }

extern "C" {

int plkh_num_states(int seq_type) { return valid_type(seq_type) ? seq_num_states((SeqType)seq_type) : -1; }
int plkh_state_unknown(int seq_type) {
    return valid_type(seq_type) ? seq_state_unknown((SeqType)seq_type) : -1;
}

int plkh_rate_matrix(int n, const double* exch_upper, const double* pi, double* Q) {
    if (n < 2 || !exch_upper || !pi || !Q) return PLK_ERR_ARG;
    build_reversible_q(n, exch_upper, pi, Q);
    return PLK_OK;
}

int plkh_decompose(int n, const double* Q, const double* pi, double* eval, double* evec,
                   double* inv_evec) {
    if (n < 2 || !Q || !pi || !eval || !evec || !inv_evec) return PLK_ERR_ARG;
    return decompose_reversible(n, Q, pi, eval, evec, inv_evec) < 0 ? PLK_ERR_NONFINITE : PLK_OK;
}

int plkh_gamma_rates(int ncat, double alpha, double p_inv, double* rates, double* props) {
    if (ncat < 1 || !(alpha > 0.0) || !(p_inv >= 0.0 && p_inv < 1.0) || !rates || !props)
        return PLK_ERR_ARG;
    gamma_mean_rates(ncat, alpha, p_inv, rates, props);
    return PLK_OK;
}

int plkh_uninformative_patterns(int nstates, int nseq, long max_patterns, uint8_t* out, long cap, long* count) {
    if (!count || nstates < 2 || nseq < 2 || max_patterns < 0 || cap < 0) return PLK_ERR_ARG;
    std::vector<uint8_t> rows;
    const long n = uninformative_patterns(nstates, nseq, max_patterns, rows);
    if (n < 0) return PLK_ERR_UNSUPPORTED;
    *count = n;
    if (out) std::memcpy(out, rows.data(), (size_t)std::min(n, cap) * nseq);
    return PLK_OK;
}

int plkh_minimize_one_dimen(double (*f)(double, void*), void* arg, double xmin, double xguess, double xmax,
                            double tol, double* x, double* fx, int* nevals) {
    if (!f || !x || !(xmin <= xmax) || !(tol > 0.0)) return PLK_ERR_ARG;
    BrentRun br;
    br.start(xmin, xguess, xmax, tol);
    while (!br.done()) br.feed(f(br.x(), arg));
    *x = br.result();
    if (fx) *fx = br.fx();
    if (nevals) *nevals = br.evals();
    return PLK_OK;
}

int plkh_tip_vectors(int seq_type, const double* inv_evec, double* tip) {
    if (!valid_type(seq_type) || !inv_evec || !tip) return PLK_ERR_ARG;
    tip_eigen_vectors((SeqType)seq_type, inv_evec, tip);
    return PLK_OK;
}

int plkh_codon_exchangeabilities(double kappa, double omega, double* exch_upper) {
    if (!exch_upper) return PLK_ERR_ARG;
    gy94_exchangeabilities(kappa, omega, exch_upper);
    return PLK_OK;
}

plkh_patterns* plkh_patterns_build(int seq_type, int nseq, int nchar, const char* chars, int* status) {
    if (!valid_type(seq_type) || nseq < 1 || nchar < 1 || !chars) {
        if (status) *status = PLK_ERR_ARG;
        return nullptr;
    }
    plkh_patterns* p = new plkh_patterns();
    const int rc = build_patterns((SeqType)seq_type, nseq, nchar, chars, p->ps);
    if (rc != 0) {
        delete p;
        if (status) *status = PLK_ERR_ARG;
        return nullptr;
    }
    if (status) *status = PLK_OK;
    return p;
}

void plkh_patterns_destroy(plkh_patterns* p) { delete p; }

int plkh_patterns_dims(const plkh_patterns* p, int* nseq, int* nsite, int* npattern) {
    if (!p) return PLK_ERR_ARG;
    if (nseq) *nseq = p->ps.nseq;
    if (nsite) *nsite = p->ps.nsite;
    if (npattern) *npattern = p->ps.npattern;
    return PLK_OK;
}

int plkh_patterns_get(const plkh_patterns* p, uint8_t* states, int32_t* freq, int32_t* const_char,
                      int32_t* site_pattern) {
    if (!p) return PLK_ERR_ARG;
    const PatternSet& ps = p->ps;
    if (states) std::memcpy(states, ps.states.data(), ps.states.size());
    if (freq) std::memcpy(freq, ps.freq.data(), ps.freq.size() * sizeof(int32_t));
    if (const_char) std::memcpy(const_char, ps.const_char.data(), ps.const_char.size() * sizeof(int32_t));
    if (site_pattern)
        std::memcpy(site_pattern, ps.site_pattern.data(), ps.site_pattern.size() * sizeof(int32_t));
    return PLK_OK;
}

int plkh_patterns_invar(const plkh_patterns* p, const double* pi, double p_inv, double* invar) {
    if (!p || !pi || !invar) return PLK_ERR_ARG;
    compute_ptn_invar(p->ps.type, p->ps, pi, p_inv, p->ps.npattern, invar);
    return PLK_OK;
}

static WorkloadSpec to_spec(const plkh_workload_spec* s) {
    WorkloadSpec ws;
    ws.type = s->seq_type;
    ws.ntaxa = s->ntaxa;
    ws.nsite = s->nsite;
    ws.ncat = s->ncat;
    ws.mean_len = s->mean_len;
    ws.alpha = s->alpha;
    ws.p_inv = s->p_inv;
    ws.gap_frac = s->gap_frac;
    ws.seed = s->seed;
    return ws;
}

static bool valid_spec(const plkh_workload_spec* s) {
    return s && valid_type(s->seq_type) && s->ntaxa >= 3 && s->ntaxa <= 100000 && s->nsite >= 1 &&
           s->ncat >= 1 && s->ncat <= 16 && s->mean_len > 0.0 && s->p_inv >= 0.0 && s->p_inv < 1.0 &&
           s->gap_frac >= 0.0 && s->gap_frac <= 1.0;
}

plkh_workload* plkh_workload_create(const plkh_workload_spec* spec, int* status) {
    if (!valid_spec(spec)) {
        if (status) *status = PLK_ERR_ARG;
        return nullptr;
    }
    plkh_workload* w = new plkh_workload();
    const int rc = make_workload(to_spec(spec), w->w);
    if (rc != 0) {
        delete w;
        if (status) *status = PLK_ERR_ARG;
        return nullptr;
    }
    if (status) *status = PLK_OK;
    return w;
}

plkh_workload* plkh_workload_create_on_tree(const plkh_workload_spec* spec, int nedges,
                                            const int* u, const int* v, const double* len,
                                            int* status) {
    if (!valid_spec(spec) || !u || !v || !len || nedges != 2 * spec->ntaxa - 3) {
        if (status) *status = PLK_ERR_ARG;
        return nullptr;
    }
    for (int e = 0; e < nedges; e++)
        if (u[e] < 0 || v[e] < 0 || u[e] > nedges || v[e] > nedges || !(len[e] >= 0.0)) {
            if (status) *status = PLK_ERR_ARG;
            return nullptr;
        }
    Tree tree = tree_from_edges(spec->ntaxa, nedges, u, v, len);
    for (int x = 0; x < tree.nnodes; x++)
        if ((x < tree.ntaxa && tree.deg[x] != 1) || (x >= tree.ntaxa && tree.deg[x] != 3)) {
            if (status) *status = PLK_ERR_ARG;
            return nullptr;
        }
    plkh_workload* w = new plkh_workload();
    const int rc = make_workload_on_tree(to_spec(spec), tree, w->w);
    if (rc != 0) {
        delete w;
        if (status) *status = PLK_ERR_ARG;
        return nullptr;
    }
    if (status) *status = PLK_OK;
    return w;
}

void plkh_workload_destroy(plkh_workload* w) { delete w; }

int plkh_workload_dims(const plkh_workload* w, int* ntaxa, int* nstates, int* ncat, int* npattern,
                       int* state_unknown, int* nsite) {
    if (!w) return PLK_ERR_ARG;
    if (ntaxa) *ntaxa = w->w.tree.ntaxa;
    if (nstates) *nstates = w->w.model.nstates;
    if (ncat) *ncat = w->w.model.ncat;
    if (npattern) *npattern = w->w.patterns.npattern;
    if (state_unknown) *state_unknown = seq_state_unknown(w->w.model.type);
    if (nsite) *nsite = w->w.patterns.nsite;
    return PLK_OK;
}

int plkh_workload_get(const plkh_workload* wl, const char* name, void* out) {
    if (!wl || !name || !out) return PLK_ERR_ARG;
    const Workload& w = wl->w;
    const std::string n(name);
    if (n == "states") put(out, w.patterns.states);
    else if (n == "freq") put(out, w.freq);
    else if (n == "invar") put(out, w.invar);
    else if (n == "eval") put(out, w.model.eval);
    else if (n == "evec") put(out, w.model.evec);
    else if (n == "inv_evec") put(out, w.model.inv_evec);
    else if (n == "Q") put(out, w.model.Q);
    else if (n == "pi") put(out, w.model.pi);
    else if (n == "tip") put(out, w.model.tip);
    else if (n == "rates") put(out, w.model.rates);
    else if (n == "props") put(out, w.model.props);
    else if (n == "exch") put(out, w.model.exch);
    else if (n == "site_pattern") put(out, w.patterns.site_pattern);
    else if (n == "const_char") put(out, w.patterns.const_char);
    else if (n == "params") {
        const double p[4] = {w.model.alpha, w.model.p_inv, w.model.kappa, w.model.omega};
        std::memcpy(out, p, sizeof(p));
    } else if (n == "edges_u" || n == "edges_v" || n == "edges_len") {
        // in edge-id order
        const Tree& t = w.tree;
        for (int e = 0; e < t.nbranch(); e++) {
            if (n == "edges_u") static_cast<int32_t*>(out)[e] = t.eu[e];
            else if (n == "edges_v") static_cast<int32_t*>(out)[e] = t.ev[e];
            else static_cast<double*>(out)[e] = t.length(t.eu[e], t.ev[e]);
        }
    } else {
        return PLK_ERR_ARG;
    }
    return PLK_OK;
}

plkh_tree* plkh_tree_create(plk_context* ctx, const plkh_workload* w, int precision,
                            int safe_numeric, int* status) {
    if (!w) {
        if (status) *status = PLK_ERR_ARG;
        return nullptr;
    }
    plkh_tree* t = new plkh_tree();
    t->t = new PhyloTreeGPU(ctx, w->w, precision, safe_numeric);
    if (t->t->status() != PLK_OK) {
        const int st = t->t->status();
        delete t->t;
        delete t;
        if (status) *status = st;
        return nullptr;
    }
    if (status) *status = PLK_OK;
    return t;
}

void plkh_tree_destroy(plkh_tree* t) {
    if (!t) return;
    delete t->t;
    delete t;
}

int plkh_tree_instance(const plkh_tree* t) { return t ? t->t->instance() : -1; }
int plkh_tree_num_branches(const plkh_tree* t) { return t ? t->t->tree().nbranch() : -1; }

int plkh_tree_branches(plkh_tree* t, int32_t* a, int32_t* b) {
    if (!t || !a || !b) return PLK_ERR_ARG;
    std::vector<int> nodes, nodes2;
    t->t->best_traversal(nodes, nodes2);
    for (size_t i = 0; i < nodes.size(); i++) {
        a[i] = nodes[i];
        b[i] = nodes2[i];
    }
    return PLK_OK;
}

static bool is_branch(const PhyloTreeGPU* t, int a, int b) {
    const Tree& tr = t->tree();
    return a >= 0 && b >= 0 && a < tr.nnodes && b < tr.nnodes && tr.neighbor_index(a, b) >= 0;
}

int plkh_tree_get_length(const plkh_tree* t, int a, int b, double* len) {
    if (!t || !len || !is_branch(t->t, a, b)) return PLK_ERR_ARG;
    *len = t->t->tree().length(a, b);
    return PLK_OK;
}

int plkh_tree_set_length(plkh_tree* t, int a, int b, double len) {
    if (!t || !is_branch(t->t, a, b) || !(len >= 0.0)) return PLK_ERR_ARG;
    t->t->set_length(a, b, len);
    return PLK_OK;
}

int plkh_tree_clear_all_partial(plkh_tree* t) {
    if (!t) return PLK_ERR_ARG;
    t->t->clear_all_partial();
    return PLK_OK;
}

int plkh_tree_reset(plkh_tree* t) {
    if (!t) return PLK_ERR_ARG;
    t->t->reset();
    return PLK_OK;
}

int plkh_tree_plan(plkh_tree* t, int a, int b, plk_op* ops, int cap, int* nops, plk_end* node,
                   plk_end* dad, double* length) {
    if (!t || !ops || !nops || !node || !dad || !is_branch(t->t, a, b)) return PLK_ERR_ARG;
    std::vector<plk_op> v;
    double len;
    t->t->plan(a, b, v, *node, *dad, len);
    if ((int)v.size() > cap) return PLK_ERR_ARG;
    std::memcpy(ops, v.data(), v.size() * sizeof(plk_op));
    *nops = (int)v.size();
    if (length) *length = len;
    return PLK_OK;
}

int plkh_tree_compute_likelihood(plkh_tree* t, double* lnl, double* pattern_lh) {
    if (!t || !lnl) return PLK_ERR_ARG;
    return t->t->compute_likelihood(lnl, pattern_lh);
}

int plkh_tree_likelihood_branch(plkh_tree* t, int a, int b, double* lnl, double* pattern_lh) {
    if (!t || !lnl || !is_branch(t->t, a, b)) return PLK_ERR_ARG;
    return t->t->likelihood_branch(a, b, lnl, pattern_lh);
}

int plkh_tree_likelihood_derv(plkh_tree* t, int a, int b, double* df, double* ddf) {
    if (!t || !df || !ddf || !is_branch(t->t, a, b)) return PLK_ERR_ARG;
    return t->t->likelihood_derv(a, b, df, ddf);
}

int plkh_tree_optimize_branch(plkh_tree* t, int a, int b, int max_nr, double* new_len, int* n_derv) {
    if (!t || !is_branch(t->t, a, b) || max_nr < 1) return PLK_ERR_ARG;
    return t->t->optimize_branch(a, b, max_nr, new_len, n_derv);
}

int plkh_tree_optimize_all_branches(plkh_tree* t, int iterations, double tolerance, int max_nr,
                                    double* lnl, int* n_derv) {
    if (!t || iterations < 1 || max_nr < 1) return PLK_ERR_ARG;
    return t->t->optimize_all_branches(iterations, tolerance, max_nr, lnl, n_derv);
}

int plkh_tree_plan_full(plkh_tree* t, int a, int b, plk_op* ops, int cap, int* nops, plk_end* node,
                        plk_end* dad) {
    if (!t || !nops || !node || !dad || cap < 0 || (cap > 0 && !ops) || !is_branch(t->t, a, b))
        return PLK_ERR_ARG;
    std::vector<plk_op> v;
    t->t->plan_full(a, b, v, *node, *dad);
    *nops = (int)v.size();
    if ((int)v.size() > cap) return PLK_ERR_ARG;
    if (!v.empty()) std::memcpy(ops, v.data(), sizeof(plk_op) * v.size());
    return PLK_OK;
}

int plkh_tree_parsimony_branch(plkh_tree* t, int a, int b, double* score, double* branch_subst) {
    if (!t || !score || !is_branch(t->t, a, b)) return PLK_ERR_ARG;
    return t->t->parsimony_branch(a, b, score, branch_subst);
}

int plkh_tree_set_gamma_shape(plkh_tree* t, double alpha) {
    if (!t) return PLK_ERR_ARG;
    return t->t->set_gamma_shape(alpha);
}

int plkh_tree_optimize_gamma_shape(plkh_tree* t, double eps, double* alpha, double* lnl, int* nevals) {
    if (!t || !(eps >= 0.0)) return PLK_ERR_ARG;
    return t->t->optimize_gamma_shape(eps, alpha, lnl, nevals);
}

int plkh_models_optimize_gamma_shape(plkh_tree* const* trees, int ntrees, double eps, double* alpha,
                                     double* lnl, int32_t* nevals, int32_t* status) {
    if (ntrees < 0 || (ntrees > 0 && !trees) || !(eps >= 0.0)) return PLK_ERR_ARG;
    std::vector<PhyloTreeGPU*> ts(ntrees);
    for (int i = 0; i < ntrees; i++) {
        if (!trees[i]) return PLK_ERR_ARG;
        ts[i] = trees[i]->t;
    }
    std::vector<int> ne(ntrees), st(ntrees);
    const int r = PhyloTreeGPU::optimize_gamma_shape_batch(ts, eps, alpha, lnl, ne.data(), st.data());
    for (int i = 0; i < ntrees; i++) {
        if (nevals) nevals[i] = ne[i];
        if (status) status[i] = st[i];
    }
    return r;
}

int plkh_supertree_optimize_all_branches(plkh_tree* const* trees, int ntrees, int iterations,
                                         double tolerance, int max_nr, double* lnl,
                                         int32_t* n_derv, int32_t* status) {
    if (ntrees < 0 || (ntrees > 0 && !trees) || iterations < 1 || max_nr < 1) return PLK_ERR_ARG;
    std::vector<PhyloTreeGPU*> ts(ntrees);
    for (int i = 0; i < ntrees; i++) {
        if (!trees[i]) return PLK_ERR_ARG;
        ts[i] = trees[i]->t;
    }
    std::vector<int> nd(ntrees), st(ntrees);
    const int r = PhyloTreeGPU::optimize_all_branches_batch(ts, iterations, tolerance, max_nr, lnl,
                                                            nd.data(), st.data());
    for (int i = 0; i < ntrees; i++) {
        if (n_derv) n_derv[i] = nd[i];
        if (status) status[i] = st[i];
    }
    return r;
}

int plkh_tree_pass_counts(const plkh_tree* t, int32_t* n_derv, int cap, int* count) {
    if (!t || !count || cap < 0 || (cap > 0 && !n_derv)) return PLK_ERR_ARG;
    const std::vector<int>& v = t->t->pass_counts();
    *count = (int)v.size();
    for (int k = 0; k < (int)v.size() && k < cap; k++) n_derv[k] = v[k];
    return PLK_OK;
}

int plkh_tree_neighbors(const plkh_tree* t, int node, int32_t* nb, int* deg) {
    if (!t || !nb || !deg) return PLK_ERR_ARG;
    const Tree& tr = t->t->tree();
    if (node < 0 || node >= tr.nnodes) return PLK_ERR_ARG;
    *deg = tr.deg[node];
    for (int k = 0; k < tr.deg[node]; k++) nb[k] = tr.adj[3 * node + k];
    return PLK_OK;
}

int plkh_tree_set_length_keep(plkh_tree* t, int a, int b, double len) {
    if (!t || !is_branch(t->t, a, b) || !(len >= 0.0)) return PLK_ERR_ARG;
    t->t->set_length_keep(a, b, len);
    return PLK_OK;
}

int plkh_tree_nni_begin(plkh_tree* t, int node1, int node2) {
    if (!t) return PLK_ERR_ARG;
    return t->t->nni_begin(node1, node2);
}

int plkh_tree_nni_swap(plkh_tree* t, int node1, int x, int node2, int y) {
    if (!t) return PLK_ERR_ARG;
    return t->t->nni_swap(node1, x, node2, y);
}

int plkh_tree_clear_partial(plkh_tree* t, int node, int from) {
    if (!t || !is_branch(t->t, node, from)) return PLK_ERR_ARG;
    t->t->clear_partial(node, from);
    return PLK_OK;
}

int plkh_tree_nni_end(plkh_tree* t) {
    if (!t) return PLK_ERR_ARG;
    t->t->nni_end();
    return PLK_OK;
}

int plkh_tree_inner_branches(plkh_tree* t, int32_t* a, int32_t* b, int cap, int* count) {
    if (!t || !count || cap < 0 || (cap > 0 && (!a || !b))) return PLK_ERR_ARG;
    std::vector<int> u, v;
    t->t->inner_branches(u, v);
    *count = (int)u.size();
    for (int k = 0; k < (int)u.size() && k < cap; k++) {
        a[k] = u[k];
        b[k] = v[k];
    }
    return PLK_OK;
}

int plkh_tree_nni_evaluate(plkh_tree* t, int node1, int node2, int max_nr, double* lnl,
                           double* lens, int32_t* n_derv) {
    if (!t || !lnl || !lens || max_nr < 1) return PLK_ERR_ARG;
    int nd[10] = {0};
    const int st = t->t->nni_evaluate(node1, node2, max_nr, lnl, lens, nd);
    if (n_derv)
        for (int k = 0; k < 10; k++) n_derv[k] = nd[k];
    return st;
}

int plkh_tree_nni_evaluate_all(plkh_tree* t, int max_nr, double* lnl, double* lens, int32_t* n_derv,
                               int cap) {
    if (!t || !lnl || !lens || max_nr < 1 || cap < 0) return PLK_ERR_ARG;
    std::vector<int> u, v;
    t->t->inner_branches(u, v);
    if ((int)u.size() > cap) return PLK_ERR_ARG;
    std::vector<double> l, ln;
    std::vector<int> nd;
    const int st = t->t->nni_evaluate_all(max_nr, l, ln, nd);
    if (st != PLK_OK) return st;
    std::memcpy(lnl, l.data(), sizeof(double) * l.size());
    std::memcpy(lens, ln.data(), sizeof(double) * ln.size());
    if (n_derv)
        for (size_t k = 0; k < nd.size(); k++) n_derv[k] = nd[k];
    return PLK_OK;
}

int plkh_tree_set_device_newton(plkh_tree* t, int on) {
    if (!t) return PLK_ERR_ARG;
    t->t->device_newton = on != 0;
    return PLK_OK;
}

int plkh_tree_set_branch_bounds(plkh_tree* t, double min_len, double max_len) {
    if (!t || !(min_len > 0.0) || !(max_len > min_len)) return PLK_ERR_ARG;
    t->t->min_len = min_len;
    t->t->max_len = max_len;
    return PLK_OK;
}

}  // extern "C"
