/*
 * phylokernelgpu.cpp - the MI355X likelihood engine (libplk, include/plk.h)
 * as an IQ-TREE 2 kernel set.
 *
 * This file is compiled into IQ-TREE (next to tree/phylokernelfma.cpp) and
 * linked against libplk.so; it is not part of libplk. It implements the four
 * kernel-set members that PhyloTree::setLikelihoodKernel assigns
 * (tree/phylotreesse.cpp:91-199, tree/phylokernelfma.cpp:36-182) as
 * explicit specializations of PhyloTree's generic kernel templates
 * (tree/phylotree.h:919-922, 972-975, 1007-1010, 1306-1309) for the tag type
 * PlkDevice, so the reference headers need no change:
 *
 *   computePartialLikelihood    computePartialLikelihoodGenericSIMD<PlkDevice>
 *   computeLikelihoodBranch     computeLikelihoodBranchGenericSIMD<PlkDevice>
 *   computeLikelihoodDerv       computeLikelihoodDervGenericSIMD<PlkDevice>
 *   computeLikelihoodFromBuffer computeLikelihoodFromBufferGenericSIMD<PlkDevice>
 *
 * plkUseDeviceKernel(tree) points the tree's four member-function pointers
 * at them (setLikelihoodKernel calls it for LK_GPU). Tree search,
 * ModelFinder and UFBoot keep calling computeLikelihoodBranch,
 * computeLikelihoodDerv, computeLikelihoodFromBuffer and optimizeOneBranch
 * unchanged.
 *
 * Buffers. PhyloTree keeps owning its partial-likelihood pointers; the
 * device holds the values. A PhyloNeighbor's partial_lh is a slot id: a
 * pointer into central_partial_lh maps to slot (ptr - central_partial_lh) /
 * getPartialLhSize(), a pointer into nni_partial_lh (the two NNI buffers,
 * tree/phylotree.cpp:1070-1076, 4127-4136) to max_lh_slots + (ptr -
 * nni_partial_lh) / getPartialLhSize(). Pointer swaps (NNI, mem_slots
 * takeover) are therefore free. Slot validity follows partial_lh_computed
 * exactly as for the CPU kernels: computeTraversalInfo lists the stale
 * partials, and one engine call computes them all.
 *
 * Per tree, one engine instance holds tip states, ptn_freq, ptn_invar and
 * the eigensystem; they are uploaded again when their contents change
 * (model or rate parameters during ModelFinder, p_inv changes).
 */
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "plk.h"
#include "tree/phylotree.h"

/* tag type naming the device kernel set in PhyloTree's kernel templates */
struct PlkDevice {};

namespace {

struct PlkTreeState {
    int instance = -1;
    plk_instance_desc desc{};
    uint64_t model_sum = 0, pattern_sum = 0;
    std::vector<plk_op> ops;
};

std::mutex g_mu;
plk_context* g_ctx = nullptr;
std::unordered_map<const PhyloTree*, PlkTreeState> g_trees;

uint64_t mix_bytes(uint64_t h, const void* p, size_t n) {
    const unsigned char* b = static_cast<const unsigned char*>(p);
    for (size_t i = 0; i < n; i++) h = (h ^ b[i]) * 1099511628211ull;
    return h;
}

void plk_check(plk_status st, const char* what) {
    if (st != PLK_OK) outError(std::string(what) + ": " + plk_last_error());
}

plk_context* device() {
    if (!g_ctx) {
        plk_config cfg{};
        cfg.device = 0;
        plk_check(plk_create(&cfg, &g_ctx), "plk_create");
    }
    return g_ctx;
}

}  // namespace

/* One tree's view of what the engine needs, gathered inside the member
 * specializations (which may read PhyloTree's protected members). */
struct PlkTreeView {
    const PhyloTree* tree;
    int ntaxa, nstates, ncat, npatterns, nslots, state_unknown, safe;
    const double *central, *nni;
    size_t block;
    const double *eval, *evec, *inv_evec, *tip, *freq, *invar;
    std::vector<double> rates, props;
    Alignment* aln;

    int slot(const double* p) const {
        if (p >= central && p < central + (size_t)(nslots - 2) * block)
            return (int)((p - central) / block);
        if (p >= nni && p < nni + 2 * block) return nslots - 2 + (int)((p - nni) / block);
        outError("partial_lh outside the tree's partial-likelihood buffers");
        return -1;
    }
};

namespace {

/* the tree's engine instance, created and refreshed as its inputs change */
PlkTreeState& sync(const PlkTreeView& v) {
    std::lock_guard<std::mutex> lock(g_mu);
    plk_context* ctx = device();
    PlkTreeState& s = g_trees[v.tree];
    plk_instance_desc d{};
    d.ntaxa = v.ntaxa;
    d.nstates = v.nstates;
    d.ncat = v.ncat;
    d.npatterns = v.npatterns;
    d.nslots = v.nslots;
    d.state_unknown = v.state_unknown;
    d.precision = PLK_PREC_F64;
    d.safe_numeric = v.safe;
    if (s.instance >= 0 && std::memcmp(&d, &s.desc, sizeof(d)) != 0) {
        plk_check(plk_instance_destroy(ctx, s.instance), "plk_instance_destroy");
        s = PlkTreeState();
    }
    if (s.instance < 0) {
        plk_check(plk_instance_create(ctx, &d, &s.instance), "plk_instance_create");
        s.desc = d;
    }
    const size_t P = v.npatterns, n = v.nstates;
    uint64_t ps = mix_bytes(1469598103934665603ull, v.freq, P * sizeof(double));
    ps = mix_bytes(ps, v.invar, P * sizeof(double));
    ps = mix_bytes(ps, &v.aln, sizeof(v.aln));
    if (ps != s.pattern_sum) {
        // taxon-major converted states (alignment/alignmentsummary.h:37)
        std::vector<uint8_t> states((size_t)v.ntaxa * P);
        for (size_t p = 0; p < P; p++) {
            const Pattern& pat = v.aln->at(p);
            for (int t = 0; t < v.ntaxa; t++) states[(size_t)t * P + p] = (uint8_t)pat[t];
        }
        plk_check(plk_set_patterns(ctx, s.instance, states.data(), v.freq, v.invar),
                  "plk_set_patterns");
        s.pattern_sum = ps;
    }
    uint64_t ms = mix_bytes(1469598103934665603ull, v.eval, n * sizeof(double));
    ms = mix_bytes(ms, v.evec, n * n * sizeof(double));
    ms = mix_bytes(ms, v.inv_evec, n * n * sizeof(double));
    ms = mix_bytes(ms, v.tip, (size_t)(v.state_unknown + 1) * n * sizeof(double));
    ms = mix_bytes(ms, v.rates.data(), v.rates.size() * sizeof(double));
    ms = mix_bytes(ms, v.props.data(), v.props.size() * sizeof(double));
    if (ms != s.model_sum) {
        plk_check(plk_set_model(ctx, s.instance, v.eval, v.evec, v.inv_evec, v.tip, v.rates.data(),
                                v.props.data()),
                  "plk_set_model");
        s.model_sum = ms;
    }
    return s;
}

/* TraversalInfo list -> engine op list (tree/phylotree.h:320-331).
 * partial_of reads PhyloNeighbor::partial_lh, which only PhyloTree members
 * may (a lambda written inside one). */
template <typename PartialOf>
void to_ops(const PlkTreeView& v, const std::vector<TraversalInfo>& info, std::vector<plk_op>& ops,
            PartialOf partial_of) {
    ops.clear();
    for (const TraversalInfo& ti : info) {
        PhyloNode* node = (PhyloNode*)ti.dad_branch->node;
        plk_op op{};
        op.dad_slot = v.slot(partial_of(ti.dad_branch));
        FOR_NEIGHBOR_DECLARE(node, ti.dad, it) {
            if (op.nchildren == PLK_MAX_CHILDREN)
                outError("the device kernel set handles binary trees only");
            PhyloNeighbor* ch = (PhyloNeighbor*)*it;
            plk_child& c = op.child[op.nchildren++];
            c.is_tip = ch->node->isLeaf();
            c.id = c.is_tip ? ch->node->id : v.slot(partial_of(ch));
            c.length = ch->length;
        }
        ops.push_back(op);
    }
}

}  // namespace

#define PLK_TREE_VIEW(v)                                                                      \
    PlkTreeView v;                                                                            \
    v.tree = this;                                                                            \
    v.ntaxa = (int)aln->getNSeq();                                                            \
    v.nstates = aln->num_states;                                                              \
    v.ncat = site_rate->getNRate();                                                           \
    v.npatterns = (int)aln->size();                                                           \
    v.nslots = (int)max_lh_slots + 2;                                                         \
    v.state_unknown = aln->STATE_UNKNOWN;                                                     \
    v.safe = safe_numeric ? 1 : 0;                                                            \
    v.central = central_partial_lh;                                                           \
    v.nni = nni_partial_lh;                                                                   \
    v.block = getPartialLhSize();                                                             \
    v.eval = model->getEigenvalues();                                                         \
    v.evec = model->getEigenvectors();                                                        \
    v.inv_evec = model->getInverseEigenvectors();                                             \
    v.tip = tip_partial_lh;                                                                   \
    v.freq = ptn_freq;                                                                        \
    v.invar = ptn_invar;                                                                      \
    v.aln = aln;                                                                              \
    for (int c_ = 0; c_ < v.ncat; c_++) {                                                     \
        v.rates.push_back(site_rate->getRate(c_));                                            \
        v.props.push_back(site_rate->getProp(c_));                                            \
    }                                                                                         \
    auto partial_of = [](PhyloNeighbor* nb) -> const double* { return nb->partial_lh; };

/* computePartialLikelihood (tree/phylokernelnew.h:1289-2039): the device
 * kernel set computes a whole traversal in one engine call from
 * computeLikelihoodBranch / computeLikelihoodDerv, so this member is reached
 * only through the pointer; it updates the one partial of `info`. */
template <>
void PhyloTree::computePartialLikelihoodGenericSIMD<PlkDevice, false, false, false>(
    TraversalInfo& info, size_t ptn_lower, size_t ptn_upper, int thread_id) {
    if (ptn_lower != 0 || thread_id != 0) return;   // one call does every pattern
    (void)ptn_upper;
    PLK_TREE_VIEW(v);
    PlkTreeState& s = sync(v);
    std::vector<TraversalInfo> one(1, info);
    to_ops(v, one, s.ops, partial_of);
    plk_check(plk_update_partials(g_ctx, s.instance, s.ops.data(), (int)s.ops.size()),
              "plk_update_partials");
}

/* computeLikelihoodBranchSIMD (tree/phylokernelnew.h:2641-3223) */
template <>
double PhyloTree::computeLikelihoodBranchGenericSIMD<PlkDevice, false, false, false>(
    PhyloNeighbor* dad_branch, PhyloNode* dad, bool save_log_value) {
    PhyloNode* node = (PhyloNode*)dad_branch->node;
    PhyloNeighbor* node_branch = (PhyloNeighbor*)node->findNeighbor(dad);
    if (!central_partial_lh) initializeAllPartialLh();
    if ((tip_partial_lh_computed & 1) == 0) computeTipPartialLikelihood();
    if (node->isLeaf()) {   // the leaf goes to dad (2654-2661)
        PhyloNode* t = dad;
        dad = node;
        node = t;
        PhyloNeighbor* b = dad_branch;
        dad_branch = node_branch;
        node_branch = b;
    }
    // stale partials on both sides, as computeTraversalInfo<VectorClass>
    traversal_info.clear();
    double* buffer = buffer_partial_lh;
    computeTraversalInfo(dad_branch, dad, buffer);
    computeTraversalInfo(node_branch, node, buffer);
    PLK_TREE_VIEW(v);
    PlkTreeState& s = sync(v);
    to_ops(v, traversal_info, s.ops, partial_of);
    traversal_info.clear();
    plk_end a{0, v.slot(partial_of(dad_branch))};
    plk_end b = dad->isLeaf() ? plk_end{1, dad->id} : plk_end{0, v.slot(partial_of(node_branch))};
    double lnl = 0.0;
    const plk_status st = plk_likelihood_branch(g_ctx, s.instance, s.ops.data(), (int)s.ops.size(), a,
                                                b, dad_branch->length, &lnl,
                                                save_log_value ? _pattern_lh : nullptr);
    if (st == PLK_ERR_NONFINITE) {
        // per-pattern fix-up already applied (3106-3132)
        outWarning(safe_numeric ? "Numerical underflow for lh-branch"
                                : "Numerical underflow for lh-branch, use `-safe` option to avoid this warning");
    } else {
        plk_check(st, "plk_likelihood_branch");
    }
    return lnl;
}

/* computeLikelihoodDervSIMD (tree/phylokernelnew.h:2233-2633), theta
 * computed once per branch as computeLikelihoodBufferSIMD (2048-2231) */
template <>
void PhyloTree::computeLikelihoodDervGenericSIMD<PlkDevice, false, false, false>(
    PhyloNeighbor* dad_branch, PhyloNode* dad, double* df, double* ddf) {
    PhyloNode* node = (PhyloNode*)dad_branch->node;
    PhyloNeighbor* node_branch = (PhyloNeighbor*)node->findNeighbor(dad);
    if (!central_partial_lh) initializeAllPartialLh();
    if ((tip_partial_lh_computed & 1) == 0) computeTipPartialLikelihood();
    if (node->isLeaf()) {
        PhyloNode* t = dad;
        dad = node;
        node = t;
        PhyloNeighbor* b = dad_branch;
        dad_branch = node_branch;
        node_branch = b;
    }
    PLK_TREE_VIEW(v);
    PlkTreeState& s = sync(v);
    if (!theta_computed) {
        traversal_info.clear();
        double* buffer = buffer_partial_lh;
        computeTraversalInfo(dad_branch, dad, buffer);
        computeTraversalInfo(node_branch, node, buffer);
        to_ops(v, traversal_info, s.ops, partial_of);
        traversal_info.clear();
        plk_end a{0, v.slot(partial_of(dad_branch))};
        plk_end b = dad->isLeaf() ? plk_end{1, dad->id} : plk_end{0, v.slot(partial_of(node_branch))};
        plk_check(plk_likelihood_buffer(g_ctx, s.instance, s.ops.data(), (int)s.ops.size(), a, b),
                  "plk_likelihood_buffer");
        theta_computed = true;
    }
    const plk_status st = plk_likelihood_derv(g_ctx, s.instance, dad_branch->length, df, ddf, nullptr);
    if (st == PLK_ERR_NONFINITE) {
        // df = ddf = 0 (2629-2632); the normal kernel stops (2577-2579)
        if (!safe_numeric)
            outError("Numerical underflow (lh-derivative). Run again with the safe likelihood kernel via `-safe` option");
        outWarning("Numerical underflow for lh-derivative");
    } else {
        plk_check(st, "plk_likelihood_derv");
    }
}

/* computeLikelihoodFromBufferSIMD (tree/phylokernelnew.h:3232-3428) */
template <>
double PhyloTree::computeLikelihoodFromBufferGenericSIMD<PlkDevice, false, false>() {
    PLK_TREE_VIEW(v);
    PlkTreeState& s = sync(v);
    double lnl = 0.0;
    const plk_status st = plk_likelihood_from_buffer(g_ctx, s.instance, current_it->length, &lnl, _pattern_lh);
    if (st == PLK_ERR_NONFINITE)
        outError("Numerical underflow (lh-from-buffer). Run again with the safe likelihood kernel via `-safe` option");
    plk_check(st, "plk_likelihood_from_buffer");
    return lnl;
}

/* setLikelihoodKernel for LK_GPU: the four members above, no mixlen
 * derivative (as tree/phylokernelfma.cpp:44). Reversible models with 4, 20
 * or 61 states and 1, 2, 4 or 8 rate categories. */
void plkUseDeviceKernel(PhyloTree* tree) {
    tree->computePartialLikelihoodPointer =
        &PhyloTree::computePartialLikelihoodGenericSIMD<PlkDevice, false, false, false>;
    tree->computeLikelihoodBranchPointer =
        &PhyloTree::computeLikelihoodBranchGenericSIMD<PlkDevice, false, false, false>;
    tree->computeLikelihoodDervPointer =
        &PhyloTree::computeLikelihoodDervGenericSIMD<PlkDevice, false, false, false>;
    tree->computeLikelihoodFromBufferPointer =
        &PhyloTree::computeLikelihoodFromBufferGenericSIMD<PlkDevice, false, false>;
    tree->computeLikelihoodDervMixlenPointer = nullptr;
}

/* releases a tree's engine instance (call from PhyloTree::deleteAllPartialLh
 * or the destructor) */
void plkReleaseTree(const PhyloTree* tree) {
    std::lock_guard<std::mutex> lock(g_mu);
    auto it = g_trees.find(tree);
    if (it == g_trees.end()) return;
    if (g_ctx && it->second.instance >= 0) plk_instance_destroy(g_ctx, it->second.instance);
    g_trees.erase(it);
}
