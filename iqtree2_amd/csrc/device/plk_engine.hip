// C-ABI device engine: contexts, instances (one PhyloTree's device state),
// launches of the kernels in plk_kernels.h. See include/plk.h for the
// contract of every entry point.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <unordered_map>
#include <mutex>
#include <string>
#include <vector>

#include "plk.h"
#include "plk_kernels.h"
#include "plk_sweep.h"
#include "plk_batch.h"
#include "plk_codon.h"
#include "plk_mfma.h"
#include "plk_pars.h"
#include "plk_nonrev.h"
#include "newton.h"

using namespace plkdev;

namespace {

thread_local std::string g_last_error = "";

plk_status fail(plk_status st, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return st;
}

#define HIP_TRY(expr)                                                                        \
    do {                                                                                     \
        hipError_t _e = (expr);                                                              \
        if (_e != hipSuccess)                                                                \
            return fail(PLK_ERR_DEVICE, "%s failed: %s", #expr, hipGetErrorString(_e));      \
    } while (0)

constexpr int kPadPatterns = 256;  // Ppad is a multiple of every tile size
constexpr int kSweepMaxBranches = PLK_SWEEP_MAX_BRANCHES;

size_t round_up(size_t x, size_t m) { return (x + m - 1) / m * m; }

// Scratch of plk_likelihood_derv_batch_noalloc (items, tile list, tile sums
// and results of one call): a module-scope device array, so that entry point
// allocates no device memory. Each GPU the library runs on holds its own copy
// until the process ends; in host code the array's name is not a device
// address, so host code only ever uses the per-device address that
// hipGetSymbolAddress returns (scratch_address).
constexpr size_t kNoAllocScratchBytes = (size_t)4 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

// this GPU's scratch address, looked up once per device with the device current
hipError_t scratch_address(int dev, unsigned char** out) {
    *out = nullptr;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[dev] = static_cast<unsigned char*>(p);
    }
    *out = g_scratch_ptr[dev];
    return hipSuccess;
}

struct Instance {
    plk_instance_desc d{};
    int Ppad = 0;
    size_t block = 0;
    size_t elem = 8;            // bytes of one stored value
    size_t scale_elem = 2;      // bytes of one scale counter
    int table_size = 0;         // elements per (op, child) table
    int sweep_table_size = 0;   // elements per child table in the sweep's LDS layout
    size_t bytes = 0;           // HBM held
    // device buffers
    uint8_t* states = nullptr;
    double* freq = nullptr;
    double* invar = nullptr;
    double* model = nullptr;    // eval | evec | inv_evec | tip | rates | props
    void* partial = nullptr;
    void* scale = nullptr;
    void* theta = nullptr;
    double* buffer_scale = nullptr;
    double* ptn_lh = nullptr;
    double* block_sums = nullptr;
    double* red_out = nullptr;
    double* root_tab = nullptr;
    plk_op* ops = nullptr;
    int ops_cap = 0;
    // device-resident Newton sweep
    plk_sweep_op* sw_ops = nullptr;
    size_t sw_ops_cap = 0;
    plk_sweep_branch* sw_br = nullptr;
    size_t sw_br_cap = 0;
    double* sw_len = nullptr;
    size_t sw_len_cap = 0;
    int* sw_nderv = nullptr;
    double* sw_part = nullptr;
    size_t sw_part_cap = 0;
    unsigned* sw_flags = nullptr;     // arrival flags, release flag, abort flag
    size_t sw_flags_cap = 0;
    double* sw_out = nullptr;         // [0..1] lnL, status; [2..3] reduced sums
    double* sw_res = nullptr;         // per-branch lengths, then requested lnL values
    size_t sw_res_cap = 0;
    long long* sw_timing = nullptr;   // phase ticks of the last sweeps (PLK_SWEEP_TIMING)
    unsigned char* sw_tabs = nullptr; // codon sweep: two steps' child tables
    size_t sw_tabs_cap = 0;
    void* tables = nullptr;
    bool patterns_set = false, model_set = false, theta_ready = false, ptn_lh_ready = false;
    double theta_length = 0.0;
    // Fitch parsimony state (allocated by the first parsimony call)
    uint32_t* pars_tip_planes = nullptr;   // [ntaxa][n][Ppad/32] (DNA, protein)
    uint32_t* pars_planes = nullptr;       // [nslots][n][Ppad/32] (DNA, protein)
    double* pars_counts = nullptr;         // [nslots][Ppad/32] (DNA, protein)
    uint64_t* pars_sets = nullptr;         // [nslots][Ppad] (codon)
    uint16_t* pars_changes = nullptr;      // [nslots][Ppad] (codon)
    bool pars_tips_ready = false;          // tip planes of the current patterns
    plk_op* pars_ops = nullptr;            // ops sorted by level, then level starts (int)
    int pars_ops_cap = 0;
    // +ASC (Lewis): the last nasc patterns are the unobserved ones
    int nasc = 0;
    double* asc_vals = nullptr;  // more than kAscMax of them: [nasc][4] values, then the 4 sums
    size_t asc_cap = 0;
    // state-basis models (non-reversible, mixtures): plk_set_model_nonrev
    bool nonrev = false;
    int nr_nmix = 0, nr_rooted = 0;
    double* nr_model = nullptr;   // ceval | cevec | cinv | Q | pi | rate | wt, then int mix[C]
    size_t nr_model_cap = 0;
    double* nr_root = nullptr;    // root tables T[3][C][n][n], lookups [3][SU+1][C][n]
    size_t nr_root_cap = 0;
    // Sankoff parsimony: cost matrix [n][n], tip costs [SU+1][n], costs [nslots][Ppad][n]
    uint32_t* sk_cost = nullptr;
    uint32_t* sk_part = nullptr;
    bool sk_ready = false;
    uint8_t* nr_flags = nullptr;  // child flags of the staged ops [2*nops]
    size_t nr_flags_cap = 0;
    plk_end nr_node{0, 0}, nr_dad{0, 0};   // ends of the last buffer call
    std::vector<double> hfreq;   // pattern weights (host copy)
    double nsite = 0.0;          // sites of the observed patterns
    double staged_bytes = 0.0, staged_flops = 0.0;   // per-pattern cost of the staged ops
};

size_t alloc_class(size_t bytes);

// red_out: 8 reduced scalars, then four values per unobserved pattern of
// +ASC when there are at most kAscMax of them (the constant patterns of
// ASC_VARIANT, one per state); the larger sets of ASC_INFORMATIVE live in
// the instance's asc_vals
constexpr int kAscMax = 64;
constexpr int kRedOutDoubles = 8 + 4 * kAscMax;

// device bytes of an instance, as allocated (size classes included), with
// the op tables of a full traversal
size_t footprint_of(const plk_instance_desc* d, int Ppad) {
    const size_t elem = d->precision == PLK_PREC_F32 ? 4 : 8;
    const size_t scale_elem = d->precision == PLK_PREC_F32 ? 4 : 2;
    const size_t n = d->nstates, SU = d->state_unknown;
    const size_t block = n * d->ncat;
    const size_t ncnt = d->safe_numeric ? d->ncat : 1;
    const size_t table = std::max(block * n, (SU + 1) * block);
    const size_t bufs[] = {
        (size_t)d->ntaxa * Ppad,                                  // states
        (size_t)Ppad * 8, (size_t)Ppad * 8,                       // freq, invar
        (n + 2 * n * n + (SU + 1) * n + 2 * (size_t)d->ncat) * 8, // model
        (size_t)d->nslots * Ppad * block * elem,                  // partials
        (size_t)d->nslots * Ppad * ncnt * scale_elem,             // scale counters
        (size_t)Ppad * block * elem,                              // theta
        (size_t)Ppad * 8, (size_t)Ppad * 8,                       // buffer_scale, ptn_lh
        (size_t)(Ppad / (kThreads / 8) + 1) * 3 * 8,              // block sums
        kRedOutDoubles * 8,                                       // reduced scalars, +ASC values
        (SU + 2) * block * 8,                                     // root tables
        sizeof(plk_op) * (size_t)d->nslots,                       // ops of a full traversal
        (size_t)d->nslots * 2 * table * elem,                     // and their tables
    };
    size_t b = 0;
    for (size_t x : bufs) b += alloc_class(x);
    return b;
}

}  // namespace

enum ProfKind { K_TRAVERSE = 0, K_TRAVERSE_LNL, K_TRAVERSE_THETA, K_DERV, K_FROM_BUFFER,
                K_PREP_OPS, K_PREP_ROOT, K_REDUCE, K_RELL, K_SWEEP, K_BATCH, K_PREP_BATCH,
                K_TRAVERSE_MFMA, K_ROOT, K_BATCH_DNA, K_BATCH_PROT, K_BATCH_CODON, K_PARSIMONY,
                K_NR_PREP, K_NR_ROOT, K_SANKOFF, K_NKIND };
static const char* kProfNames[K_NKIND] = {"traverse", "traverse+lnl", "traverse+theta",
                                          "derv", "lnl_from_buffer", "prep_ops", "prep_root",
                                          "reduce", "rell", "newton_sweep", "batch", "prep_batch",
                                          "traverse_mfma", "root", "batch_dna", "batch_protein",
                                          "batch_codon", "parsimony", "nonrev_prep", "nonrev_root", "sankoff"};

struct ProfRec {
    int kind;
    double bytes, flops;
    int ev;
};

struct plk_context {
    int device = 0;
    int num_cu = 0;
    int sweep_blocks_per_cu = 2;
    bool sweep_lds = true;              // theta resident in LDS during sweeps when it fits
    int lds_per_cu = 0;                 // bytes of LDS per CU
    int mfma_sub = 1;                   // 16-pattern sub-tiles per protein traversal workgroup
    bool prof_on = false;
    std::vector<hipEvent_t> prof_ev;
    size_t prof_next = 0;
    std::vector<ProfRec> prof_recs;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    size_t budget = 0, used = 0;
    std::mutex mu;
    std::vector<Instance*> inst;
    double* h_out = nullptr;  // pinned host staging for reduced scalars
    unsigned char* h_stage = nullptr;   // pinned host staging for uploads
    size_t h_stage_cap = 0;
    bool h_stage_inflight = false;      // an upload from h_stage may still run
    // batched evaluation buffers (grown on demand)
    unsigned char* bt_upload = nullptr;   // items | ops | tile list | root lengths
    size_t bt_upload_cap = 0;
    unsigned char* bt_tables = nullptr;
    size_t bt_tables_cap = 0;
    double* bt_roots = nullptr;
    size_t bt_roots_cap = 0;
    double* bt_tile_sums = nullptr;
    size_t bt_tile_sums_cap = 0;
    double* bt_vals = nullptr;      // lnl out
    size_t bt_vals_cap = 0;
    unsigned* bt_counter = nullptr;
    size_t bt_counter_cap = 0;
    // batched derivative buffers
    DervItem* db_items = nullptr;
    size_t db_items_cap = 0;
    int* db_tile_item = nullptr;
    size_t db_tile_cap = 0;
    double* db_tile_sums = nullptr;
    size_t db_tile_sums_cap = 0;
    double* db_out = nullptr;
    size_t db_out_cap = 0;
    // device address of this GPU's copy of the no-allocation scratch array
    unsigned char* scratch = nullptr;
    int batch_blocks_per_cu = 0;    // 0: autotuned
    // caching device allocator: live blocks by address, free blocks by size
    // class; `used` counts live bytes, `cached` the free blocks kept
    std::unordered_map<void*, size_t> live;
    std::multimap<size_t, void*> cache;
    size_t cached = 0;
    // autotuner: key (class mix, log2 tiles) -> candidate shapes tried so far
    // (one per call: the launch a call makes anyway is the trial) and the best
    struct BatchTune {
        int next = 0;            // next candidate to time
        float best_ms = 1e30f;
        int split = 1, bpc = 2;  // best so far (default until every candidate ran)
        bool done = false;
    };
    std::map<long, BatchTune> batch_tune;
    hipEvent_t tune_ev[2] = {nullptr, nullptr};
    bool autotune = true;
};

namespace {

// Allocation size classes: powers of two up to 64 KiB, then four steps per
// octave, so a freed block serves any later request of its class (at most
// 25% rounding above 64 KiB).
size_t alloc_class(size_t bytes) {
    bytes = round_up(bytes ? bytes : 8, 256);
    size_t p = 256;
    while (p < bytes) p <<= 1;
    if (p <= 65536) return p;
    const size_t step = p / 8;
    size_t c = p / 2;
    while (c < bytes) c += step;
    return c;
}

void release_cached(plk_context* ctx, size_t need) {
    // largest blocks first, until live + cached + need fits the budget
    while (!ctx->cache.empty() && ctx->used + ctx->cached + need > ctx->budget) {
        auto last = std::prev(ctx->cache.end());
        (void)hipFree(last->second);
        ctx->cached -= last->first;
        ctx->cache.erase(last);
    }
}

// Device allocation through the context's block cache. Every use of a
// block is ordered on the context stream, so a block freed by one instance
// can back the next one without a device synchronisation.
template <typename P>
plk_status dmalloc(plk_context* ctx, Instance* in, P** ptr, size_t bytes) {
    const size_t cls = alloc_class(bytes);
    if (ctx->used + cls > ctx->budget)
        return fail(PLK_ERR_NOMEM, "HBM budget exceeded: need %zu more bytes, %zu of %zu in use",
                    cls, ctx->used, ctx->budget);
    void* p = nullptr;
    auto it = ctx->cache.find(cls);
    if (it != ctx->cache.end()) {
        p = it->second;
        ctx->cache.erase(it);
        ctx->cached -= cls;
    } else {
        release_cached(ctx, cls);
        hipError_t e = hipMalloc(&p, cls);
        if (e != hipSuccess && !ctx->cache.empty()) {
            (void)hipGetLastError();
            release_cached(ctx, ctx->budget);   // everything cached
            e = hipMalloc(&p, cls);
        }
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return fail(PLK_ERR_NOMEM, "hipMalloc(%zu): %s", cls, hipGetErrorString(e));
        }
    }
    ctx->live[p] = cls;
    ctx->used += cls;
    if (in) in->bytes += cls;
    *ptr = reinterpret_cast<P*>(p);
    return PLK_OK;
}

void dfree(plk_context* ctx, Instance* in, void* p, size_t) {
    if (!p) return;
    auto it = ctx->live.find(p);
    if (it == ctx->live.end()) {
        (void)hipFree(p);
        return;
    }
    const size_t cls = it->second;
    ctx->live.erase(it);
    ctx->used = ctx->used >= cls ? ctx->used - cls : 0;
    if (in) in->bytes = in->bytes >= cls ? in->bytes - cls : 0;
    ctx->cache.insert({cls, p});
    ctx->cached += cls;
}

Instance* get_instance(plk_context* ctx, int id);

// returns the instance's blocks to the cache (caller holds ctx->mu)
plk_status destroy_instance(plk_context* ctx, int id) {
    Instance* in = get_instance(ctx, id);
    if (!in) return fail(PLK_ERR_ARG, "unknown instance %d", id);
    void* bufs[] = {in->states, in->freq, in->invar, in->model, in->partial, in->scale,
                    in->theta, in->buffer_scale, in->ptn_lh, in->block_sums, in->red_out,
                    in->root_tab, in->ops, in->tables, in->sw_ops, in->sw_br, in->sw_len,
                    in->sw_nderv, in->sw_part, in->sw_flags, in->sw_out, in->sw_timing, in->sw_res, in->sw_tabs,
                    in->pars_tip_planes, in->pars_planes, in->pars_counts, in->pars_sets,
                    in->pars_changes, in->pars_ops, in->asc_vals, in->nr_model, in->nr_root,
                    in->nr_flags, in->sk_cost, in->sk_part};
    for (void* b : bufs) dfree(ctx, nullptr, b, 0);
    delete in;
    ctx->inst[id] = nullptr;
    return PLK_OK;
}

// event pair around a launch when profiling is on; returns the pair index or -1
int prof_begin(plk_context* ctx) {
    if (!ctx->prof_on) return -1;
    if (ctx->prof_next + 2 > ctx->prof_ev.size()) {
        const size_t old = ctx->prof_ev.size();
        ctx->prof_ev.resize(old ? 2 * old : 8192);
        for (size_t i = old; i < ctx->prof_ev.size(); i++)
            if (hipEventCreate(&ctx->prof_ev[i]) != hipSuccess) return -1;
    }
    const int idx = (int)ctx->prof_next;
    ctx->prof_next += 2;
    (void)hipEventRecord(ctx->prof_ev[idx], ctx->stream);
    return idx;
}

void prof_end(plk_context* ctx, int idx, int kind, double bytes, double flops) {
    if (idx < 0) return;
    (void)hipEventRecord(ctx->prof_ev[idx + 1], ctx->stream);
    ctx->prof_recs.push_back(ProfRec{kind, bytes, flops, idx});
}

// algorithmic bytes / flops per pattern of one traversal step: child partials
// read once, the parent written once, one state byte per tip child, scale
// counters read and written
void op_cost(const Instance* in, const plk_op& op, double& bytes, double& flops) {
    const double s = (double)in->elem, n = in->d.nstates, C = in->d.ncat;
    int k_int = 0, k_tip = 0;
    for (int k = 0; k < op.nchildren; k++) (op.child[k].is_tip ? k_tip : k_int)++;
    const double ncnt = in->d.safe_numeric ? C : 1;
    bytes = C * n * s * (k_int + 1) + k_tip + in->scale_elem * ncnt * (k_int + 1);
    flops = C * (2.0 * n * n * k_int + n * (k_int + k_tip - 1) + 2.0 * n * n);
}

Instance* get_instance(plk_context* ctx, int id) {
    if (!ctx || id < 0 || id >= (int)ctx->inst.size()) return nullptr;
    return ctx->inst[id];
}

InstDev inst_dev(const Instance* in) {
    InstDev I;
    I.P = in->d.npatterns;
    I.Ppad = in->Ppad;
    I.n = in->d.nstates;
    I.C = in->d.ncat;
    I.SU = in->d.state_unknown;
    I.safe = in->d.safe_numeric;
    I.states = in->states;
    I.freq = in->freq;
    I.invar = in->invar;
    I.partial = in->partial;
    I.scale = in->scale;
    I.theta = in->theta;
    I.buffer_scale = in->buffer_scale;
    I.ptn_lh = in->ptn_lh;
    I.slot_stride = (size_t)in->Ppad * in->block;
    I.scale_stride = (size_t)in->Ppad * (in->d.safe_numeric ? in->d.ncat : 1);
    I.nasc = in->nasc;
    I.asc = in->nasc > kAscMax ? in->asc_vals : in->red_out + 8;
    return I;
}

ModelDev model_dev(const Instance* in) {
    const int n = in->d.nstates;
    ModelDev m;
    m.eval = in->model;
    m.evec = m.eval + n;
    m.inv_evec = m.evec + n * n;
    m.tip = m.inv_evec + n * n;
    m.rates = m.tip + (size_t)(in->d.state_unknown + 1) * n;
    m.props = m.rates + in->d.ncat;
    return m;
}

NrDev nr_dev(const Instance* in) {
    const int n = in->d.nstates, C = in->d.ncat, nm = in->nr_nmix;
    NrDev D;
    D.ceval = in->nr_model;
    D.cevec = D.ceval + (size_t)nm * 2 * n;
    D.cinv = D.cevec + (size_t)nm * 2 * n * n;
    D.Q = D.cinv + (size_t)nm * 2 * n * n;
    D.pi = D.Q + (size_t)nm * n * n;
    D.rate = D.pi + (size_t)nm * n;
    D.wt = D.rate + C;
    D.mix = reinterpret_cast<const int*>(D.wt + C);
    D.n = n;
    D.C = C;
    D.rooted = in->nr_rooted;
    return D;
}

plk_status check_ops(const Instance* in, const plk_op* ops, int nops) {
    if (nops < 0 || (nops > 0 && !ops)) return fail(PLK_ERR_ARG, "invalid op list");
    for (int o = 0; o < nops; o++) {
        const plk_op& op = ops[o];
        if (op.dad_slot < 0 || op.dad_slot >= in->d.nslots)
            return fail(PLK_ERR_ARG, "op %d: dad slot %d out of range", o, op.dad_slot);
        if (op.nchildren != 2)
            return fail(PLK_ERR_UNSUPPORTED, "op %d: %d children (binary nodes only)", o, op.nchildren);
        for (int k = 0; k < 2; k++) {
            const plk_child& ch = op.child[k];
            // state-basis instances: bit 1 toward the root, bit 2 the root leaf
            const int allowed = in->nonrev ? (kNrTip | kNrTowardRoot | kNrRootLeaf) : 1;
            if ((ch.is_tip & ~allowed) || ((ch.is_tip & kNrRootLeaf) && !(ch.is_tip & kNrTip)))
                return fail(PLK_ERR_ARG, "op %d child %d: invalid flags %d", o, k, ch.is_tip);
            const int lim = (ch.is_tip & kNrTip) ? in->d.ntaxa : in->d.nslots;
            if (ch.id < 0 || ch.id >= lim)
                return fail(PLK_ERR_ARG, "op %d child %d: id %d out of range", o, k, ch.id);
            if (!(ch.is_tip & kNrTip) && ch.id == op.dad_slot)
                return fail(PLK_ERR_ARG, "op %d: child slot equals dad slot", o);
            if (!(ch.length >= 0.0) || !std::isfinite(ch.length))
                return fail(PLK_ERR_ARG, "op %d child %d: invalid length", o, k);
        }
    }
    return PLK_OK;
}

plk_status check_end(const Instance* in, plk_end e, bool allow_tip) {
    const int allowed = in->nonrev ? (kNrTip | kNrRootLeaf) : 1;
    if ((e.is_tip & ~allowed) || ((e.is_tip & kNrRootLeaf) && !(e.is_tip & kNrTip)))
        return fail(PLK_ERR_ARG, "invalid end flags %d", e.is_tip);
    if (e.is_tip) {
        if (!allow_tip) return fail(PLK_ERR_ARG, "branch end must be a partial slot");
        if (e.id < 0 || e.id >= in->d.ntaxa) return fail(PLK_ERR_ARG, "tip id %d out of range", e.id);
    } else if (e.id < 0 || e.id >= in->d.nslots) {
        return fail(PLK_ERR_ARG, "slot %d out of range", e.id);
    }
    return PLK_OK;
}

// ---- kernel dispatch over (precision, states, categories) -------------------
template <typename T, int N, int C>
plk_status launch_traverse_t(plk_context* ctx, Instance* in, int nops, int mode, plk_end node,
                             plk_end dad) {
    const int PB = kThreads / C;
    const int grid = in->Ppad / PB;
    const size_t tab_bytes = 2 * (size_t)in->table_size * sizeof(T);
    const size_t root_bytes = (size_t)(in->d.state_unknown + 2) * in->block * 8;
    const size_t smem = round_up((size_t)N * N * sizeof(T), 16) + std::max(tab_bytes, root_bytes);
    if (smem > 160 * 1024) return fail(PLK_ERR_UNSUPPORTED, "LDS need %zu bytes", smem);
    const int pe = prof_begin(ctx);
    hipLaunchKernelGGL((k_traverse<T, N, C>), dim3(grid), dim3(kThreads), smem, ctx->stream,
                       inst_dev(in), model_dev(in), in->ops, nops,
                       reinterpret_cast<const T*>(in->tables), in->table_size, mode, node, dad,
                       in->root_tab, in->block_sums);
    HIP_TRY(hipGetLastError());
    if (pe >= 0) {
        const double P = in->d.npatterns, s = (double)in->elem, blk = (double)in->block;
        const double ncnt = in->d.safe_numeric ? in->d.ncat : 1;
        double b = in->staged_bytes, f = in->staged_flops;
        int kind = K_TRAVERSE;
        if (mode != MODE_PARTIALS) {
            const int nint = 1 + (dad.is_tip ? 0 : 1);
            b += blk * s * nint + (dad.is_tip ? 1 : 0) + in->scale_elem * ncnt * nint;
            if (mode == MODE_BRANCH) {
                b += 24.0;   // invar, freq, pattern lnL
                f += 3.0 * blk;
                kind = K_TRAVERSE_LNL;
            } else {
                b += blk * s + 8.0;   // theta, buffer_scale
                f += blk;
                kind = K_TRAVERSE_THETA;
            }
        }
        prof_end(ctx, pe, kind, b * P, f * P);
    }
    return PLK_OK;
}

template <typename T, int N>
plk_status launch_traverse_n(plk_context* ctx, Instance* in, int nops, int mode, plk_end node,
                             plk_end dad) {
    switch (in->d.ncat) {
    case 1: return launch_traverse_t<T, N, 1>(ctx, in, nops, mode, node, dad);
    case 2: return launch_traverse_t<T, N, 2>(ctx, in, nops, mode, node, dad);
    case 4: return launch_traverse_t<T, N, 4>(ctx, in, nops, mode, node, dad);
    case 8: return launch_traverse_t<T, N, 8>(ctx, in, nops, mode, node, dad);
    }
    return fail(PLK_ERR_UNSUPPORTED, "ncat %d not built", in->d.ncat);
}

template <typename T>
plk_status launch_traverse_p(plk_context* ctx, Instance* in, int nops, int mode, plk_end node,
                             plk_end dad) {
    // 20 and 61 states run on the matrix cores (launch_traverse)
    switch (in->d.nstates) {
    case 4: return launch_traverse_n<T, 4>(ctx, in, nops, mode, node, dad);
    }
    return fail(PLK_ERR_UNSUPPORTED, "nstates %d not built", in->d.nstates);
}

// branch-end bytes per pattern of the root stage (lnL or theta)
double root_bytes(const Instance* in, int mode, plk_end dad) {
    const double s = (double)in->elem, blk = (double)in->block;
    const double ncnt = in->d.safe_numeric ? in->d.ncat : 1;
    const int nint = 1 + (dad.is_tip ? 0 : 1);
    double b = blk * s * nint + (dad.is_tip ? 1 : 0) + in->scale_elem * ncnt * nint;
    b += mode == MODE_BRANCH ? 24.0 : blk * s + 8.0;
    return b;
}

template <typename T, int N, int C>
plk_status launch_root_stream(plk_context* ctx, Instance* in, int mode, plk_end node, plk_end dad) {
    const int grid = in->Ppad / (kThreads / C);
    const int pe = prof_begin(ctx);
    hipLaunchKernelGGL((k_root_stream<T, N, C>), dim3(grid), dim3(kThreads), 0, ctx->stream,
                       inst_dev(in), model_dev(in), mode, node, dad, in->root_tab, in->block_sums);
    HIP_TRY(hipGetLastError());
    prof_end(ctx, pe, K_ROOT, root_bytes(in, mode, dad) * in->d.npatterns,
             (mode == MODE_BRANCH ? 3.0 : 1.0) * in->block * in->d.npatterns);
    return PLK_OK;
}

// 61 states: the MFMA traversal kernel, then the streaming root stage
template <typename T>
plk_status launch_traverse_codon(plk_context* ctx, Instance* in, int nops, int mode, plk_end node,
                                 plk_end dad) {
    if (nops > 0) {
        const dim3 grid(in->Ppad / kCodonTile, in->d.ncat);
        const int pe = prof_begin(ctx);
        hipLaunchKernelGGL(k_traverse_codon<T>, grid, dim3(kThreads), 0, ctx->stream,
                           inst_dev(in), model_dev(in), in->ops, nops,
                           reinterpret_cast<const T*>(in->tables), in->table_size);
        HIP_TRY(hipGetLastError());
        prof_end(ctx, pe, K_TRAVERSE_MFMA, in->staged_bytes * in->d.npatterns,
                 in->staged_flops * in->d.npatterns);
    }
    if (mode == MODE_PARTIALS) return PLK_OK;
    switch (in->d.ncat) {
    case 1: return launch_root_stream<T, kCodonN, 1>(ctx, in, mode, node, dad);
    case 2: return launch_root_stream<T, kCodonN, 2>(ctx, in, mode, node, dad);
    case 4: return launch_root_stream<T, kCodonN, 4>(ctx, in, mode, node, dad);
    case 8: return launch_root_stream<T, kCodonN, 8>(ctx, in, mode, node, dad);
    }
    return fail(PLK_ERR_UNSUPPORTED, "ncat %d not built", in->d.ncat);
}

// 20 states: the MFMA traversal kernel, then the streaming root stage
template <typename T, int C>
plk_status launch_traverse_mfma_c(plk_context* ctx, Instance* in, int nops, int mode, plk_end node,
                                  plk_end dad) {
    constexpr int N = 20;
    if (nops > 0) {
        const int sub = ctx->mfma_sub;
        const int grid = in->Ppad / (kMfmaTile * sub);
        const int pe = prof_begin(ctx);
        auto go = [&](auto kern) {
            hipLaunchKernelGGL(kern, dim3(grid), dim3(kThreads), 0, ctx->stream, inst_dev(in), model_dev(in),
                               in->ops, nops, reinterpret_cast<const T*>(in->tables), in->table_size);
        };
        if (sub == 4) go(k_traverse_mfma<T, N, C, 4>);
        else if (sub == 2) go(k_traverse_mfma<T, N, C, 2>);
        else go(k_traverse_mfma<T, N, C, 1>);
        HIP_TRY(hipGetLastError());
        prof_end(ctx, pe, K_TRAVERSE_MFMA, in->staged_bytes * in->d.npatterns,
                 in->staged_flops * in->d.npatterns);
    }
    if (mode == MODE_PARTIALS) return PLK_OK;
    return launch_root_stream<T, N, C>(ctx, in, mode, node, dad);
}

template <typename T>
plk_status launch_traverse_protein(plk_context* ctx, Instance* in, int nops, int mode, plk_end node,
                                   plk_end dad) {
    switch (in->d.ncat) {
    case 1: return launch_traverse_mfma_c<T, 1>(ctx, in, nops, mode, node, dad);
    case 2: return launch_traverse_mfma_c<T, 2>(ctx, in, nops, mode, node, dad);
    case 4: return launch_traverse_mfma_c<T, 4>(ctx, in, nops, mode, node, dad);
    case 8: return launch_traverse_mfma_c<T, 8>(ctx, in, nops, mode, node, dad);
    }
    return fail(PLK_ERR_UNSUPPORTED, "ncat %d not built", in->d.ncat);
}

plk_status launch_traverse(plk_context* ctx, Instance* in, int nops, int mode, plk_end node,
                           plk_end dad) {
    const bool f32 = in->d.precision == PLK_PREC_F32;
    if (in->d.nstates == kCodonN)
        return f32 ? launch_traverse_codon<float>(ctx, in, nops, mode, node, dad)
                   : launch_traverse_codon<double>(ctx, in, nops, mode, node, dad);
    if (in->d.nstates == 20)
        return f32 ? launch_traverse_protein<float>(ctx, in, nops, mode, node, dad)
                   : launch_traverse_protein<double>(ctx, in, nops, mode, node, dad);
    if (f32) return launch_traverse_p<float>(ctx, in, nops, mode, node, dad);
    return launch_traverse_p<double>(ctx, in, nops, mode, node, dad);
}

template <typename T, int N, int C>
plk_status launch_buffer_t(plk_context* ctx, Instance* in, double length, int mode) {
    const int PB = kThreads / C;
    const int grid = in->Ppad / PB;
    int pe = prof_begin(ctx);
    hipLaunchKernelGGL((k_buffer<T, N, C>), dim3(grid), dim3(kThreads), 0, ctx->stream,
                       inst_dev(in), model_dev(in), length, mode, in->block_sums);
    HIP_TRY(hipGetLastError());
    {
        const double P = in->d.npatterns, blk = (double)in->block, s = (double)in->elem;
        // theta + invar + freq + buffer_scale (+ pattern lnL written by lnL-from-buffer)
        const double b = blk * s + 24.0 + (mode == BUF_LNL ? 8.0 : 0.0);
        const double f = (mode == BUF_DERV ? 6.0 : 2.0) * blk + 10.0;
        prof_end(ctx, pe, mode == BUF_DERV ? K_DERV : K_FROM_BUFFER, b * P, f * P);
    }
    pe = prof_begin(ctx);
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(kThreads), 0, ctx->stream, in->block_sums, grid, 3,
                       in->red_out);
    HIP_TRY(hipGetLastError());
    prof_end(ctx, pe, K_REDUCE, 3.0 * 8 * grid, 3.0 * grid);
    return PLK_OK;
}

template <typename T, int N>
plk_status launch_buffer_n(plk_context* ctx, Instance* in, double length, int mode) {
    switch (in->d.ncat) {
    case 1: return launch_buffer_t<T, N, 1>(ctx, in, length, mode);
    case 2: return launch_buffer_t<T, N, 2>(ctx, in, length, mode);
    case 4: return launch_buffer_t<T, N, 4>(ctx, in, length, mode);
    case 8: return launch_buffer_t<T, N, 8>(ctx, in, length, mode);
    }
    return fail(PLK_ERR_UNSUPPORTED, "ncat %d not built", in->d.ncat);
}

plk_status launch_buffer(plk_context* ctx, Instance* in, double length, int mode) {
    const bool f32 = in->d.precision == PLK_PREC_F32;
    switch (in->d.nstates) {
    case 4: return f32 ? launch_buffer_n<float, 4>(ctx, in, length, mode)
                       : launch_buffer_n<double, 4>(ctx, in, length, mode);
    case 20: return f32 ? launch_buffer_n<float, 20>(ctx, in, length, mode)
                        : launch_buffer_n<double, 20>(ctx, in, length, mode);
    case kCodonN: return f32 ? launch_buffer_n<float, kCodonN>(ctx, in, length, mode)
                             : launch_buffer_n<double, kCodonN>(ctx, in, length, mode);
    }
    return fail(PLK_ERR_UNSUPPORTED, "nstates %d not built", in->d.nstates);
}

template <typename T>
plk_status grow(plk_context* ctx, Instance* in, T** p, size_t& cap, size_t need);

// upload ops and build their transition tables
plk_status stage_ops(plk_context* ctx, Instance* in, const plk_op* ops, int nops) {
    plk_status st = check_ops(in, ops, nops);
    if (st != PLK_OK) return st;
    in->staged_bytes = in->staged_flops = 0.0;
    if (nops == 0) return PLK_OK;
    if (nops > in->ops_cap) {
        const size_t elem = in->elem;
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        dfree(ctx, in, in->ops, sizeof(plk_op) * in->ops_cap);
        dfree(ctx, in, in->tables, (size_t)in->ops_cap * 2 * in->table_size * elem);
        in->ops = nullptr;
        in->tables = nullptr;
        const int cap = std::max(nops, 2 * in->ops_cap);
        st = dmalloc(ctx, in, &in->ops, sizeof(plk_op) * cap);
        if (st != PLK_OK) return st;
        st = dmalloc(ctx, in, &in->tables, (size_t)cap * 2 * in->table_size * elem);
        if (st != PLK_OK) return st;
        in->ops_cap = cap;
    }
    std::vector<plk_op> stripped;
    if (in->nonrev) {
        // the traversal kernels see plain tip bits; the flags go beside them
        stripped.assign(ops, ops + nops);
        std::vector<uint8_t> fl(2 * (size_t)nops);
        for (int o = 0; o < nops; o++)
            for (int k = 0; k < 2; k++) {
                fl[2 * o + k] = (uint8_t)(stripped[o].child[k].is_tip & (kNrTowardRoot | kNrRootLeaf));
                stripped[o].child[k].is_tip &= kNrTip;
            }
        st = grow(ctx, in, &in->nr_flags, in->nr_flags_cap, fl.size());
        if (st != PLK_OK) return st;
        HIP_TRY(hipMemcpyAsync(in->nr_flags, fl.data(), fl.size(), hipMemcpyHostToDevice, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));   // fl leaves scope
        ops = stripped.data();
    }
    HIP_TRY(hipMemcpyAsync(in->ops, ops, sizeof(plk_op) * nops, hipMemcpyHostToDevice, ctx->stream));
    if (in->nonrev) HIP_TRY(hipStreamSynchronize(ctx->stream));   // stripped leaves scope
    in->staged_bytes = in->staged_flops = 0.0;
    for (int o = 0; o < nops; o++) {
        double b, f;
        op_cost(in, ops[o], b, f);
        in->staged_bytes += b;
        in->staged_flops += f;
    }
    const size_t n = in->d.nstates, SU = in->d.state_unknown;
    const size_t smem = (n + n * n + SU * n) * sizeof(double);
    const dim3 grid(2 * nops, in->d.ncat);
    const int pe = prof_begin(ctx);
    if (in->nonrev) {
        const size_t nsmem = (2 * n + n * n) * sizeof(double);
        const double* tip = model_dev(in).tip;
        if (in->d.precision == PLK_PREC_F32) {
            hipLaunchKernelGGL(k_nr_prep_ops<float>, grid, dim3(kThreads), nsmem, ctx->stream, in->ops,
                               in->nr_flags, nops, nr_dev(in), tip, in->d.state_unknown, in->table_size,
                               reinterpret_cast<float*>(in->tables));
        } else {
            hipLaunchKernelGGL(k_nr_prep_ops<double>, grid, dim3(kThreads), nsmem, ctx->stream, in->ops,
                               in->nr_flags, nops, nr_dev(in), tip, in->d.state_unknown, in->table_size,
                               reinterpret_cast<double*>(in->tables));
        }
    } else if (in->d.precision == PLK_PREC_F32) {
        hipLaunchKernelGGL(k_prep_ops<float>, grid, dim3(kThreads), smem, ctx->stream,
                           in->ops, nops, model_dev(in), in->d.nstates, in->d.ncat,
                           in->d.state_unknown, in->table_size, reinterpret_cast<float*>(in->tables));
    } else {
        hipLaunchKernelGGL(k_prep_ops<double>, grid, dim3(kThreads), smem, ctx->stream,
                           in->ops, nops, model_dev(in), in->d.nstates, in->d.ncat,
                           in->d.state_unknown, in->table_size, reinterpret_cast<double*>(in->tables));
    }
    HIP_TRY(hipGetLastError());
    prof_end(ctx, pe, K_PREP_OPS, 2.0 * nops * in->table_size * in->elem, 0.0);
    return PLK_OK;
}

plk_status ready(Instance* in) {
    if (!in) return fail(PLK_ERR_ARG, "unknown instance");
    if (!in->patterns_set) return fail(PLK_ERR_STATE, "patterns not set");
    if (!in->model_set) return fail(PLK_ERR_STATE, "model not set");
    return PLK_OK;
}

plk_status fetch_scalars(plk_context* ctx, Instance* in, int count) {
    HIP_TRY(hipMemcpyAsync(ctx->h_out, in->red_out, sizeof(double) * count, hipMemcpyDeviceToHost,
                           ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return PLK_OK;
}

// per-pattern fix-up of a non-finite tree lnL (tree/phylokernelnew.h:3115-3132)
plk_status fixup_nonfinite(plk_context* ctx, Instance* in, double* lnl, double* pattern_lh) {
    const int P = in->d.npatterns - in->nasc;   // observed patterns
    std::vector<double> plh(P), freq(P);
    HIP_TRY(hipMemcpyAsync(plh.data(), in->ptn_lh, sizeof(double) * P, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipMemcpyAsync(freq.data(), in->freq, sizeof(double) * P, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    double s = 0.0;
    for (int p = 0; p < P; p++) {
        if (!std::isfinite(plh[p])) plh[p] = kLogScaleThreshold * 4;
        s += plh[p] * freq[p];
    }
    HIP_TRY(hipMemcpyAsync(in->ptn_lh, plh.data(), sizeof(double) * P, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (pattern_lh) std::memcpy(pattern_lh, plh.data(), sizeof(double) * P);
    *lnl = s;
    return PLK_OK;
}

// +ASC: the four per-pattern values of the unobserved patterns, each summed
// in pattern order on the host, or by the fixed-order device reduction for
// sets larger than kAscMax
plk_status asc_sums(plk_context* ctx, Instance* in, double sums[4]) {
    if (in->nasc > kAscMax) {
        double* out = in->asc_vals + 4 * (size_t)in->nasc;
        hipLaunchKernelGGL(k_reduce, dim3(1), dim3(kThreads), 0, ctx->stream, in->asc_vals, in->nasc, 4, out);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(sums, out, sizeof(double) * 4, hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        return PLK_OK;
    }
    double h[4 * kAscMax];
    HIP_TRY(hipMemcpyAsync(h, in->red_out + 8, sizeof(double) * 4 * in->nasc, hipMemcpyDeviceToHost,
                           ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    for (int j = 0; j < 4; j++) sums[j] = 0.0;
    for (int k = 0; k < in->nasc; k++)
        for (int j = 0; j < 4; j++) sums[j] += h[4 * k + j];
    return PLK_OK;
}

// Lewis's correction of a tree lnL and its pattern lnL (ASC_VARIANT,
// tree/phylokernelnew.h:3198-3221, 3405-3424): lnL -= nsite * log(1 - P(const)),
// the device pattern lnL shifted the same way (the unobserved patterns' own
// entries zeroed), so RELL and later reads see the conditioned values
plk_status asc_correct_lnl(plk_context* ctx, Instance* in, double prob_const, double* lnl, double* pattern_lh) {
    if (!(prob_const >= 0.0 && prob_const < 1.0))
        return fail(PLK_ERR_NONFINITE, "probability of the unobserved constant patterns %g outside [0, 1)",
                    prob_const);
    const double corr = std::log(1.0 - prob_const);
    *lnl -= in->nsite * corr;
    const int P = in->d.npatterns, pobs = P - in->nasc;
    hipLaunchKernelGGL(k_asc_shift, dim3((P + kThreads - 1) / kThreads), dim3(kThreads), 0, ctx->stream,
                       in->ptn_lh, pobs, P, corr);
    HIP_TRY(hipGetLastError());
    if (pattern_lh) {
        HIP_TRY(hipMemcpyAsync(pattern_lh, in->ptn_lh, sizeof(double) * P, hipMemcpyDeviceToHost, ctx->stream));
    }
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return PLK_OK;
}

// Grows a device buffer to hold at least `need` elements (capacities are
// element counts; byte sizes never pass through int).
template <typename T>
plk_status grow(plk_context* ctx, Instance* in, T** p, size_t& cap, size_t need) {
    if (need <= cap && *p) return PLK_OK;
    if (need > (std::numeric_limits<size_t>::max() / 2) / sizeof(T))
        return fail(PLK_ERR_NOMEM, "buffer of %zu elements exceeds the address space", need);
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    dfree(ctx, in, *p, sizeof(T) * cap);
    *p = nullptr;
    cap = std::max(need, 2 * cap);
    plk_status st = dmalloc(ctx, in, p, sizeof(T) * cap);
    if (st != PLK_OK) cap = 0;
    return st;
}

template <typename T, int N>
plk_status launch_nr_root_n(plk_context* ctx, Instance* in, const NrRoot& R, int grid) {
    switch (in->d.ncat) {
    case 1: hipLaunchKernelGGL((k_nr_root<T, N, 1>), dim3(grid), dim3(kThreads), 0, ctx->stream, inst_dev(in), R, in->block_sums); break;
    case 2: hipLaunchKernelGGL((k_nr_root<T, N, 2>), dim3(grid), dim3(kThreads), 0, ctx->stream, inst_dev(in), R, in->block_sums); break;
    case 4: hipLaunchKernelGGL((k_nr_root<T, N, 4>), dim3(grid), dim3(kThreads), 0, ctx->stream, inst_dev(in), R, in->block_sums); break;
    case 8: hipLaunchKernelGGL((k_nr_root<T, N, 8>), dim3(grid), dim3(kThreads), 0, ctx->stream, inst_dev(in), R, in->block_sums); break;
    default: return fail(PLK_ERR_UNSUPPORTED, "ncat %d not built", in->d.ncat);
    }
    HIP_TRY(hipGetLastError());
    return PLK_OK;
}

// State-basis branch evaluation between the buffered ends at `length`:
// h_out[0] = lnL (derv = 0); h_out[0..2] = df, ddf, lnL (derv = 1)
plk_status nr_eval(plk_context* ctx, Instance* in, double length, int derv) {
    const int n = in->d.nstates, C = in->d.ncat, SU = in->d.state_unknown;
    const plk_end node = in->nr_node, dad = in->nr_dad;
    const int dad_tip = dad.is_tip & kNrTip, dad_root = (dad.is_tip & kNrRootLeaf) ? 1 : 0;
    double* T = in->nr_root;
    double* lut = T + 3 * (size_t)C * n * n;
    const size_t smem = (2 * (size_t)n + 3 * (size_t)n * n) * sizeof(double);
    int pe = prof_begin(ctx);
    hipLaunchKernelGGL(k_nr_prep_root, dim3(C), dim3(kThreads), smem, ctx->stream, nr_dev(in),
                       model_dev(in).tip, SU, length, derv, dad_tip, dad_root, T, lut);
    HIP_TRY(hipGetLastError());
    prof_end(ctx, pe, K_NR_PREP, 3.0 * C * n * (n + SU + 1.0) * 8, 0.0);
    NrRoot R{T, lut, node.id, dad_tip ? 0 : dad.id, dad_tip, dad_tip ? dad.id : 0, derv};
    // one thread per (pattern, category)
    const int grid = (int)(((size_t)in->Ppad * C + kThreads - 1) / kThreads);
    pe = prof_begin(ctx);
    const bool f32 = in->d.precision == PLK_PREC_F32;
    plk_status st;
    if (n == 4) st = f32 ? launch_nr_root_n<float, 4>(ctx, in, R, grid) : launch_nr_root_n<double, 4>(ctx, in, R, grid);
    else st = f32 ? launch_nr_root_n<float, kProtN>(ctx, in, R, grid) : launch_nr_root_n<double, kProtN>(ctx, in, R, grid);
    if (st != PLK_OK) return st;
    {
        const double nint = dad_tip ? 1.0 : 2.0, ncnt = in->d.safe_numeric ? C : 1;
        prof_end(ctx, pe, K_NR_ROOT,
                 (double)in->d.npatterns * (in->block * in->elem * nint + in->scale_elem * ncnt * nint + 24.0),
                 (double)in->d.npatterns * C * (dad_tip ? 2.0 * n : 2.0 * n * (n + 1)) * (derv ? 3 : 1));
    }
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(kThreads), 0, ctx->stream, in->block_sums, grid, 3,
                       in->red_out);
    HIP_TRY(hipGetLastError());
    return fetch_scalars(ctx, in, 3);
}

// plk_likelihood_branch / plk_likelihood_from_buffer of a state-basis
// instance once its ends are set: lnL, the per-pattern fix-up
// (tree/phylokernelnonrev.h:1486-1497), pattern lnL
plk_status nr_lnl(plk_context* ctx, Instance* in, double length, double* lnl, double* pattern_lh) {
    plk_status st = nr_eval(ctx, in, length, 0);
    if (st != PLK_OK) return st;
    *lnl = ctx->h_out[0];
    in->ptn_lh_ready = true;
    const bool nonfinite = !std::isfinite(*lnl);
    if (nonfinite && (st = fixup_nonfinite(ctx, in, lnl, pattern_lh)) != PLK_OK) return st;
    if (in->nasc) {
        double a[4];
        if ((st = asc_sums(ctx, in, a)) != PLK_OK) return st;
        if ((st = asc_correct_lnl(ctx, in, a[0], lnl, pattern_lh)) != PLK_OK) return st;
    }
    if (nonfinite) return fail(PLK_ERR_NONFINITE, "numerical underflow for lh-branch; per-pattern fix-up applied");
    if (pattern_lh && !in->nasc) {
        HIP_TRY(hipMemcpyAsync(pattern_lh, in->ptn_lh, sizeof(double) * in->d.npatterns,
                               hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
    }
    return PLK_OK;
}

}  // namespace

namespace {

// Pinned host staging of at least need bytes for an upload on ctx->stream.
// An upload that may still read the buffer (a call that returned early) is
// waited for before the buffer is handed out again.
plk_status host_stage(plk_context* ctx, size_t need, unsigned char** out) {
    if (ctx->h_stage_inflight) {
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        ctx->h_stage_inflight = false;
    }
    if (need > ctx->h_stage_cap) {
        const size_t cap = std::max(need, 2 * ctx->h_stage_cap);
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        if (ctx->h_stage) (void)hipHostFree(ctx->h_stage);
        ctx->h_stage = nullptr;
        ctx->h_stage_cap = 0;
        HIP_TRY(hipHostMalloc((void**)&ctx->h_stage, cap, hipHostMallocDefault));
        ctx->h_stage_cap = cap;
    }
    *out = ctx->h_stage;
    return PLK_OK;
}

}  // namespace

extern "C" {

const char* plk_last_error(void) { return g_last_error.c_str(); }

plk_status plk_create(const plk_config* cfg, plk_context** out) {
    if (!out) return fail(PLK_ERR_ARG, "null output");
    *out = nullptr;
    int ndev = 0;
    HIP_TRY(hipGetDeviceCount(&ndev));
    const int dev = cfg ? cfg->device : 0;
    if (dev < 0 || dev >= ndev) return fail(PLK_ERR_ARG, "device %d of %d", dev, ndev);
    HIP_TRY(hipSetDevice(dev));
    plk_context* ctx = new plk_context();
    ctx->device = dev;
    if (cfg && cfg->stream) {
        ctx->stream = reinterpret_cast<hipStream_t>(cfg->stream);
    } else {
        hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            delete ctx;
            return fail(PLK_ERR_DEVICE, "hipStreamCreate: %s", hipGetErrorString(e));
        }
        ctx->own_stream = true;
    }
    size_t free_b = 0, total_b = 0;
    {
        // a failure from here on releases what was created so far
        hipError_t q = hipMemGetInfo(&free_b, &total_b);
        if (q == hipSuccess) q = hipDeviceGetAttribute(&ctx->num_cu, hipDeviceAttributeMultiprocessorCount, dev);
        if (q == hipSuccess)
            q = hipDeviceGetAttribute(&ctx->lds_per_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev);
        if (q == hipSuccess) q = scratch_address(dev, &ctx->scratch);
        if (q != hipSuccess) {
            if (ctx->own_stream) (void)hipStreamDestroy(ctx->stream);
            delete ctx;
            return fail(PLK_ERR_DEVICE, "device query: %s", hipGetErrorString(q));
        }
    }
    if (const char* e = std::getenv("PLK_SWEEP_BLOCKS_PER_CU")) ctx->sweep_blocks_per_cu = std::max(1, std::atoi(e));
    if (const char* e = std::getenv("PLK_SWEEP_LDS")) ctx->sweep_lds = std::atoi(e) != 0;
    if (const char* e = std::getenv("PLK_MFMA_SUB")) {
        const int v = std::atoi(e);
        ctx->mfma_sub = v == 4 ? 4 : v == 2 ? 2 : 1;
    }
    if (const char* e = std::getenv("PLK_BATCH_BLOCKS_PER_CU")) ctx->batch_blocks_per_cu = std::max(1, std::atoi(e));
    if (std::getenv("PLK_NO_AUTOTUNE")) ctx->autotune = false;
    ctx->budget = (cfg && cfg->hbm_budget) ? cfg->hbm_budget : (size_t)(0.9 * (double)free_b);
    hipError_t e = hipHostMalloc((void**)&ctx->h_out, 64 * sizeof(double), hipHostMallocDefault);
    if (e != hipSuccess) {
        if (ctx->own_stream) (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return fail(PLK_ERR_DEVICE, "hipHostMalloc: %s", hipGetErrorString(e));
    }
    *out = ctx;
    return PLK_OK;
}

plk_status plk_destroy(plk_context* ctx) {
    if (!ctx) return PLK_OK;
    (void)hipStreamSynchronize(ctx->stream);
    for (int i = 0; i < (int)ctx->inst.size(); i++)
        if (ctx->inst[i]) destroy_instance(ctx, i);
    for (hipEvent_t e : ctx->prof_ev) (void)hipEventDestroy(e);
    for (hipEvent_t e : ctx->tune_ev)
        if (e) (void)hipEventDestroy(e);
    // every block still live (batch buffers) and every cached one
    for (auto& kv : ctx->live) (void)hipFree(kv.first);
    for (auto& kv : ctx->cache) (void)hipFree(kv.second);
    if (ctx->own_stream) (void)hipStreamDestroy(ctx->stream);
    (void)hipHostFree(ctx->h_out);
    if (ctx->h_stage) (void)hipHostFree(ctx->h_stage);
    delete ctx;
    return PLK_OK;
}

plk_status plk_memory(plk_context* ctx, size_t* used, size_t* budget) {
    if (!ctx) return fail(PLK_ERR_ARG, "null context");
    if (used) *used = ctx->used;
    if (budget) *budget = ctx->budget;
    return PLK_OK;
}

void* plk_stream(plk_context* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

plk_status plk_synchronize(plk_context* ctx) {
    if (!ctx) return fail(PLK_ERR_ARG, "null context");
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return PLK_OK;
}

size_t plk_instance_footprint(const plk_instance_desc* d) {
    if (!d || d->npatterns <= 0) return 0;
    return footprint_of(d, (int)round_up(d->npatterns, kPadPatterns));
}

plk_status plk_instance_create(plk_context* ctx, const plk_instance_desc* d, int* id) {
    if (!ctx || !d || !id) return fail(PLK_ERR_ARG, "null argument");
    if (d->ntaxa < 3 || d->npatterns <= 0 || d->nslots <= 0 || d->nstates < 2 || d->ncat < 1 ||
        d->state_unknown < d->nstates || d->state_unknown > 255)
        return fail(PLK_ERR_ARG, "invalid instance shape");
    if (d->nstates != 4 && d->nstates != 20 && d->nstates != kCodonN)
        return fail(PLK_ERR_UNSUPPORTED, "nstates %d not built into this library", d->nstates);
    // IQ-TREE forces per-category scaling for every other state count
    // (tree/phylotreesse.cpp:95-96); the 61-state kernel implements only that
    if (d->nstates == kCodonN && !d->safe_numeric)
        return fail(PLK_ERR_UNSUPPORTED, "61-state models need safe_numeric = 1");
    if (d->ncat != 1 && d->ncat != 2 && d->ncat != 4 && d->ncat != 8)
        return fail(PLK_ERR_UNSUPPORTED, "ncat %d not built into this library", d->ncat);
    std::lock_guard<std::mutex> lock(ctx->mu);
    const int Ppad = (int)round_up(d->npatterns, kPadPatterns);
    const size_t need = footprint_of(d, Ppad);
    if (ctx->used + need > ctx->budget)
        return fail(PLK_ERR_NOMEM, "instance needs %zu bytes, %zu of %zu in use", need, ctx->used,
                    ctx->budget);
    HIP_TRY(hipSetDevice(ctx->device));
    Instance* in = new Instance();
    in->d = *d;
    in->Ppad = Ppad;
    in->block = (size_t)d->nstates * d->ncat;
    in->elem = d->precision == PLK_PREC_F32 ? 4 : 8;
    in->scale_elem = d->precision == PLK_PREC_F32 ? 4 : 2;
    in->sweep_table_size = (int)std::max(in->block * d->nstates, (size_t)(d->state_unknown + 1) * in->block);
    in->table_size = d->nstates == kProtN ? prot_table_size(d->ncat, d->state_unknown) : in->sweep_table_size;
    const int n = d->nstates;
    const size_t ncnt = d->safe_numeric ? d->ncat : 1;
    const size_t model_doubles = n + 2 * (size_t)n * n + (size_t)(d->state_unknown + 1) * n + 2 * d->ncat;
    const int max_grid = Ppad / (kThreads / 8) + 1;
    plk_status st = PLK_OK;
    auto A = [&](auto** p, size_t bytes) {
        if (st == PLK_OK) st = dmalloc(ctx, in, p, bytes);
    };
    A(&in->states, (size_t)d->ntaxa * Ppad);
    A(&in->freq, (size_t)Ppad * 8);
    A(&in->invar, (size_t)Ppad * 8);
    A(&in->model, model_doubles * 8);
    A(&in->partial, (size_t)d->nslots * Ppad * in->block * in->elem);
    A(&in->scale, (size_t)d->nslots * Ppad * ncnt * in->scale_elem);
    A(&in->theta, (size_t)Ppad * in->block * in->elem);
    A(&in->buffer_scale, (size_t)Ppad * 8);
    A(&in->ptn_lh, (size_t)Ppad * 8);
    A(&in->block_sums, (size_t)max_grid * 3 * 8);
    A(&in->red_out, kRedOutDoubles * 8);
    A(&in->root_tab, (size_t)(d->state_unknown + 2) * in->block * 8);
    if (st == PLK_OK) {
        // padding patterns: unknown tips, zero weight
        if (hipMemsetAsync(in->states, d->state_unknown, (size_t)d->ntaxa * Ppad, ctx->stream) != hipSuccess ||
            hipMemsetAsync(in->freq, 0, (size_t)Ppad * 8, ctx->stream) != hipSuccess ||
            hipMemsetAsync(in->invar, 0, (size_t)Ppad * 8, ctx->stream) != hipSuccess ||
            hipMemsetAsync(in->scale, 0, (size_t)d->nslots * Ppad * ncnt * in->scale_elem, ctx->stream) != hipSuccess ||
            hipStreamSynchronize(ctx->stream) != hipSuccess)
            st = fail(PLK_ERR_DEVICE, "initialising instance buffers failed");
    }
    // ids of destroyed instances are reused
    int slot = 0;
    while (slot < (int)ctx->inst.size() && ctx->inst[slot]) slot++;
    if (slot == (int)ctx->inst.size()) ctx->inst.push_back(nullptr);
    ctx->inst[slot] = in;
    *id = slot;
    if (st != PLK_OK) {
        std::string msg = g_last_error;
        destroy_instance(ctx, *id);
        *id = -1;
        g_last_error = msg;
        return st;
    }
    return PLK_OK;
}

plk_status plk_instance_destroy(plk_context* ctx, int id) {
    if (!ctx) return fail(PLK_ERR_ARG, "null context");
    std::lock_guard<std::mutex> lock(ctx->mu);
    return destroy_instance(ctx, id);
}

plk_status plk_set_patterns(plk_context* ctx, int id, const uint8_t* states, const double* freq,
                            const double* invar) {
    Instance* in = get_instance(ctx, id);
    if (!in || !states || !freq || !invar) return fail(PLK_ERR_ARG, "null argument");
    const int P = in->d.npatterns;
    for (size_t i = 0; i < (size_t)in->d.ntaxa * P; i++)
        if (states[i] > in->d.state_unknown) return fail(PLK_ERR_ARG, "tip state %d > STATE_UNKNOWN", states[i]);
    std::lock_guard<std::mutex> lock(ctx->mu);
    // padded rows, weights and +I mass packed into pinned memory: three
    // contiguous copies instead of one staged copy per taxon
    const size_t Ppad = in->Ppad, nst = (size_t)in->d.ntaxa * Ppad;
    const size_t need = nst + 2 * Ppad * sizeof(double) + 16;
    unsigned char* hs = nullptr;
    plk_status st = host_stage(ctx, need, &hs);
    if (st != PLK_OK) return st;
    for (int t = 0; t < in->d.ntaxa; t++) {
        std::memcpy(hs + (size_t)t * Ppad, states + (size_t)t * P, P);
        std::memset(hs + (size_t)t * Ppad + P, in->d.state_unknown, Ppad - P);
    }
    double* hf = reinterpret_cast<double*>(hs + round_up(nst, 16));
    double* hv = hf + Ppad;
    std::memcpy(hf, freq, sizeof(double) * P);
    std::memset(hf + P, 0, sizeof(double) * (Ppad - P));
    std::memcpy(hv, invar, sizeof(double) * P);
    std::memset(hv + P, 0, sizeof(double) * (Ppad - P));
    ctx->h_stage_inflight = true;
    HIP_TRY(hipMemcpyAsync(in->states, hs, nst, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(in->freq, hf, sizeof(double) * Ppad, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(in->invar, hv, sizeof(double) * Ppad, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    ctx->h_stage_inflight = false;
    in->hfreq.assign(freq, freq + P);
    in->nsite = 0.0;
    for (int p = 0; p < P - in->nasc; p++) in->nsite += freq[p];
    in->patterns_set = true;
    in->pars_tips_ready = false;
    in->theta_ready = false;
    return PLK_OK;
}

plk_status plk_set_asc(plk_context* ctx, int id, int nasc) {
    Instance* in = get_instance(ctx, id);
    if (!in) return fail(PLK_ERR_ARG, "invalid instance");
    if (nasc < 0 || nasc >= in->d.npatterns)
        return fail(PLK_ERR_ARG, "%d unobserved patterns of %d", nasc, in->d.npatterns);
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (nasc > kAscMax) {
        plk_status st = grow(ctx, in, &in->asc_vals, in->asc_cap, 4 * ((size_t)nasc + 1));
        if (st != PLK_OK) return st;
    }
    in->nasc = nasc;
    in->nsite = 0.0;
    for (int p = 0; p < (int)in->hfreq.size() - nasc; p++) in->nsite += in->hfreq[p];
    in->ptn_lh_ready = false;
    return PLK_OK;
}

plk_status plk_set_model(plk_context* ctx, int id, const double* eval, const double* evec,
                         const double* inv_evec, const double* tip_eigen, const double* cat_rates,
                         const double* cat_props) {
    Instance* in = get_instance(ctx, id);
    if (!in || !eval || !evec || !inv_evec || !tip_eigen || !cat_rates || !cat_props)
        return fail(PLK_ERR_ARG, "null argument");
    const int n = in->d.nstates, C = in->d.ncat, SU = in->d.state_unknown;
    std::vector<double> h;
    h.insert(h.end(), eval, eval + n);
    h.insert(h.end(), evec, evec + n * n);
    h.insert(h.end(), inv_evec, inv_evec + n * n);
    h.insert(h.end(), tip_eigen, tip_eigen + (size_t)(SU + 1) * n);
    h.insert(h.end(), cat_rates, cat_rates + C);
    h.insert(h.end(), cat_props, cat_props + C);
    for (double v : h)
        if (!std::isfinite(v)) return fail(PLK_ERR_ARG, "non-finite model parameter");
    std::lock_guard<std::mutex> lock(ctx->mu);
    HIP_TRY(hipMemcpyAsync(in->model, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice,
                           ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    in->model_set = true;
    in->nonrev = false;
    in->theta_ready = false;
    return PLK_OK;
}

plk_status plk_set_model_nonrev(plk_context* ctx, int id, int nmix, int fused, const double* ceval,
                                const double* cevec, const double* cinv_evec, const double* rate_matrix,
                                const double* state_freq, const double* mix_weights, const double* site_rates,
                                const double* site_props, int nrates, const double* tip, int rooted) {
    Instance* in = get_instance(ctx, id);
    if (!in || !ceval || !cevec || !cinv_evec || !rate_matrix || !state_freq || !mix_weights || !site_rates ||
        !site_props || !tip)
        return fail(PLK_ERR_ARG, "null argument");
    const int n = in->d.nstates, C = in->d.ncat, SU = in->d.state_unknown;
    if (n != 4 && n != kProtN) return fail(PLK_ERR_UNSUPPORTED, "state-basis models for %d states", n);
    const int Cm = fused ? nrates : nrates * nmix;   // ncat_mix
    if (nmix < 1 || nrates < 1 || Cm > C || (fused && nmix != nrates))
        return fail(PLK_ERR_ARG, "%d categories for %d mixture components and %d rates%s", C, nmix, nrates,
                    fused ? " (fused)" : "");
    const size_t nn = (size_t)n * n;
    std::vector<double> h;
    h.insert(h.end(), ceval, ceval + (size_t)nmix * 2 * n);
    h.insert(h.end(), cevec, cevec + (size_t)nmix * 2 * nn);
    h.insert(h.end(), cinv_evec, cinv_evec + (size_t)nmix * 2 * nn);
    h.insert(h.end(), rate_matrix, rate_matrix + (size_t)nmix * nn);
    h.insert(h.end(), state_freq, state_freq + (size_t)nmix * n);
    // ncat_mix categories: rate category c % ncat of component c / denom
    // (tree/phylokernelnonrev.h:1107-1114)
    const int denom = fused ? 1 : nrates;
    std::vector<double> rate(C), wt(C);
    std::vector<int> mix(C);
    for (int c = 0; c < C; c++) {
        // categories beyond ncat_mix (the instance's count is 1, 2, 4 or 8):
        // copies of category 0 with weight 0, which leave every likelihood,
        // derivative and rescale decision as they are
        const int cc = c < Cm ? c : 0;
        mix[c] = cc / denom;
        rate[c] = site_rates[cc % nrates];
        wt[c] = c < Cm ? site_props[cc % nrates] * mix_weights[mix[c]] : 0.0;
    }
    h.insert(h.end(), rate.begin(), rate.end());
    h.insert(h.end(), wt.begin(), wt.end());
    for (double v : h)
        if (!std::isfinite(v)) return fail(PLK_ERR_ARG, "non-finite model parameter");
    for (int i = 0; i < (SU + 1) * n; i++)
        if (!std::isfinite(tip[i])) return fail(PLK_ERR_ARG, "non-finite tip vector");
    const size_t ints = (C * sizeof(int) + 7) / 8;
    const size_t words = h.size() + ints;
    h.resize(words, 0.0);
    std::memcpy(h.data() + words - ints, mix.data(), C * sizeof(int));
    // the traversal kernels' model: identity output transform, plain tip vectors
    std::vector<double> m;
    m.resize(n, 0.0);
    for (int pass = 0; pass < 2; pass++)
        for (int x = 0; x < n; x++)
            for (int i = 0; i < n; i++) m.push_back(x == i ? 1.0 : 0.0);
    m.insert(m.end(), tip, tip + (size_t)(SU + 1) * n);
    m.insert(m.end(), rate.begin(), rate.end());
    m.insert(m.end(), wt.begin(), wt.end());
    std::lock_guard<std::mutex> lock(ctx->mu);
    plk_status st = grow(ctx, in, &in->nr_model, in->nr_model_cap, words);
    if (st != PLK_OK) return st;
    st = grow(ctx, in, &in->nr_root, in->nr_root_cap, 3 * (size_t)C * nn + 3 * (size_t)(SU + 1) * C * n);
    if (st != PLK_OK) return st;
    HIP_TRY(hipMemcpyAsync(in->nr_model, h.data(), sizeof(double) * words, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(in->model, m.data(), sizeof(double) * m.size(), hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    in->nr_nmix = nmix;
    in->nr_rooted = rooted ? 1 : 0;
    in->nonrev = true;
    in->model_set = true;
    in->theta_ready = false;
    return PLK_OK;
}

plk_status plk_update_partials(plk_context* ctx, int id, const plk_op* ops, int nops) {
    Instance* in = get_instance(ctx, id);
    plk_status st = ready(in);
    if (st != PLK_OK) return st;
    std::lock_guard<std::mutex> lock(ctx->mu);
    st = stage_ops(ctx, in, ops, nops);
    if (st != PLK_OK || nops == 0) return st;
    plk_end none{0, 0};
    st = launch_traverse(ctx, in, nops, MODE_PARTIALS, none, none);
    if (st != PLK_OK) return st;
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return PLK_OK;
}

plk_status plk_likelihood_branch(plk_context* ctx, int id, const plk_op* ops, int nops,
                                 plk_end node, plk_end dad, double length, double* lnl,
                                 double* pattern_lh) {
    Instance* in = get_instance(ctx, id);
    plk_status st = ready(in);
    if (st != PLK_OK) return st;
    if (!lnl) return fail(PLK_ERR_ARG, "null lnl");
    if ((st = check_end(in, node, false)) != PLK_OK) return st;
    if ((st = check_end(in, dad, true)) != PLK_OK) return st;
    if (!(length >= 0.0) || !std::isfinite(length)) return fail(PLK_ERR_ARG, "invalid length");
    std::lock_guard<std::mutex> lock(ctx->mu);
    st = stage_ops(ctx, in, ops, nops);
    if (st != PLK_OK) return st;
    if (in->nonrev) {
        const plk_end none{0, 0};
        if (nops && (st = launch_traverse(ctx, in, nops, MODE_PARTIALS, none, none)) != PLK_OK) return st;
        in->nr_node = node;
        in->nr_dad = dad;
        in->theta_ready = false;
        return nr_lnl(ctx, in, length, lnl, pattern_lh);
    }
    int pe = prof_begin(ctx);
    hipLaunchKernelGGL(k_prep_root, dim3(1), dim3(kThreads), 0, ctx->stream, model_dev(in),
                       in->d.nstates, in->d.ncat, in->d.state_unknown, length, dad.is_tip,
                       in->root_tab);
    HIP_TRY(hipGetLastError());
    prof_end(ctx, pe, K_PREP_ROOT, (in->d.state_unknown + 2.0) * in->block * 8, 0.0);
    st = launch_traverse(ctx, in, nops, MODE_BRANCH, node, dad);
    if (st != PLK_OK) return st;
    const int grid = in->Ppad / (kThreads / in->d.ncat);
    pe = prof_begin(ctx);
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(kThreads), 0, ctx->stream, in->block_sums, grid, 1,
                       in->red_out);
    HIP_TRY(hipGetLastError());
    prof_end(ctx, pe, K_REDUCE, 8.0 * grid, (double)grid);
    if ((st = fetch_scalars(ctx, in, 1)) != PLK_OK) return st;
    *lnl = ctx->h_out[0];
    in->ptn_lh_ready = true;
    const bool nonfinite = !std::isfinite(*lnl);
    if (nonfinite && (st = fixup_nonfinite(ctx, in, lnl, pattern_lh)) != PLK_OK) return st;
    if (in->nasc) {
        double a[4];
        if ((st = asc_sums(ctx, in, a)) != PLK_OK) return st;
        if ((st = asc_correct_lnl(ctx, in, a[0], lnl, pattern_lh)) != PLK_OK) return st;
    }
    if (nonfinite) return fail(PLK_ERR_NONFINITE, "numerical underflow for lh-branch; per-pattern fix-up applied");
    if (pattern_lh && !in->nasc) {
        HIP_TRY(hipMemcpyAsync(pattern_lh, in->ptn_lh, sizeof(double) * in->d.npatterns,
                               hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
    }
    return PLK_OK;
}

plk_status plk_likelihood_buffer(plk_context* ctx, int id, const plk_op* ops, int nops,
                                 plk_end node, plk_end dad) {
    Instance* in = get_instance(ctx, id);
    plk_status st = ready(in);
    if (st != PLK_OK) return st;
    if ((st = check_end(in, node, false)) != PLK_OK) return st;
    if ((st = check_end(in, dad, true)) != PLK_OK) return st;
    std::lock_guard<std::mutex> lock(ctx->mu);
    st = stage_ops(ctx, in, ops, nops);
    if (st != PLK_OK) return st;
    if (in->nonrev) {
        // no theta buffer in the nonrev kernel set: the ends, evaluated on demand
        const plk_end none{0, 0};
        if (nops && (st = launch_traverse(ctx, in, nops, MODE_PARTIALS, none, none)) != PLK_OK) return st;
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        in->nr_node = node;
        in->nr_dad = dad;
        in->theta_ready = true;
        return PLK_OK;
    }
    st = launch_traverse(ctx, in, nops, MODE_THETA, node, dad);
    if (st != PLK_OK) return st;
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    in->theta_ready = true;
    return PLK_OK;
}

plk_status plk_likelihood_derv(plk_context* ctx, int id, double length, double* df, double* ddf,
                               double* lnl) {
    Instance* in = get_instance(ctx, id);
    plk_status st = ready(in);
    if (st != PLK_OK) return st;
    if (!df || !ddf) return fail(PLK_ERR_ARG, "null output");
    if (!in->theta_ready) return fail(PLK_ERR_STATE, "theta buffer not computed");
    if (!std::isfinite(length)) return fail(PLK_ERR_ARG, "invalid length");
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (in->nonrev) {
        if ((st = nr_eval(ctx, in, length, 1)) != PLK_OK) return st;
        *df = ctx->h_out[0];
        *ddf = ctx->h_out[1];
        if (lnl) *lnl = ctx->h_out[2];
        if (in->nasc) {
            // ASC_Lewis (tree/phylokernelnonrev.h:1040-1047)
            double a[4];
            if ((st = asc_sums(ctx, in, a)) != PLK_OK) return st;
            const double pv = 1.0 - a[1];
            const double df_frac = a[2] / pv, ddf_frac = a[3] / pv;
            *df += in->nsite * df_frac;
            *ddf += in->nsite * (ddf_frac + df_frac * df_frac);
            if (lnl) *lnl -= in->nsite * std::log(1.0 - a[0]);
        }
        if (!std::isfinite(*df)) {
            *df = *ddf = 0.0;
            return fail(PLK_ERR_NONFINITE, "numerical underflow for lh-derivative");
        }
        return PLK_OK;
    }
    st = launch_buffer(ctx, in, length, BUF_DERV);
    if (st != PLK_OK) return st;
    if ((st = fetch_scalars(ctx, in, 3)) != PLK_OK) return st;
    *df = ctx->h_out[0];
    *ddf = ctx->h_out[1];
    if (lnl) *lnl = ctx->h_out[2];
    if (in->nasc) {
        // ASC_Lewis (tree/phylokernelnew.h:2620-2627): the derivative sums count
        // the unobserved patterns' lh as the derivative kernel does, the lnL
        // beside them the plain lh
        double a[4];
        if ((st = asc_sums(ctx, in, a)) != PLK_OK) return st;
        const double pv = 1.0 - a[1];
        const double df_frac = a[2] / pv, ddf_frac = a[3] / pv;
        *df += in->nsite * df_frac;
        *ddf += in->nsite * (ddf_frac + df_frac * df_frac);
        if (lnl) *lnl -= in->nsite * std::log(1.0 - a[0]);
    }
    if (!std::isfinite(*df)) {
        *df = *ddf = 0.0;
        return fail(PLK_ERR_NONFINITE, "numerical underflow for lh-derivative");
    }
    return PLK_OK;
}

plk_status plk_likelihood_from_buffer(plk_context* ctx, int id, double length, double* lnl,
                                      double* pattern_lh) {
    Instance* in = get_instance(ctx, id);
    plk_status st = ready(in);
    if (st != PLK_OK) return st;
    if (!lnl) return fail(PLK_ERR_ARG, "null output");
    if (!in->theta_ready) return fail(PLK_ERR_STATE, "theta buffer not computed");
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (in->nonrev) return nr_lnl(ctx, in, length, lnl, pattern_lh);
    st = launch_buffer(ctx, in, length, BUF_LNL);
    if (st != PLK_OK) return st;
    if ((st = fetch_scalars(ctx, in, 3)) != PLK_OK) return st;
    *lnl = ctx->h_out[2];
    in->ptn_lh_ready = true;
    const bool nonfinite = !std::isfinite(*lnl);
    if (nonfinite && (st = fixup_nonfinite(ctx, in, lnl, pattern_lh)) != PLK_OK) return st;
    if (in->nasc) {
        double a[4];
        if ((st = asc_sums(ctx, in, a)) != PLK_OK) return st;
        if ((st = asc_correct_lnl(ctx, in, a[0], lnl, pattern_lh)) != PLK_OK) return st;
    }
    if (nonfinite) {
        if (!in->d.safe_numeric)
            return fail(PLK_ERR_NONFINITE, "numerical underflow (lh-from-buffer); use the safe kernel");
        return PLK_OK;
    }
    if (pattern_lh && !in->nasc) {
        HIP_TRY(hipMemcpyAsync(pattern_lh, in->ptn_lh, sizeof(double) * in->d.npatterns,
                               hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
    }
    return PLK_OK;
}

plk_status plk_optimize_branch(plk_context* ctx, int id, const plk_op* ops, int nops,
                               plk_end node, plk_end dad, double length, double min_len,
                               double max_len, int max_nr, double* out_length, double* out_lnl,
                               int* n_derv) {
    if (!out_length) return fail(PLK_ERR_ARG, "null output");
    if (!(min_len > 0.0) || !(max_len > min_len)) return fail(PLK_ERR_ARG, "invalid bracket");
    plk_status st = plk_likelihood_buffer(ctx, id, ops, nops, node, dad);
    if (st != PLK_OK) return st;
    const bool safe = get_instance(ctx, id)->d.safe_numeric != 0;
    // Optimization::minimizeNewton (utils/optimization.cpp:422-499) on
    // f = -dlnL/dt (PhyloTree::computeFuncDerv, tree/phylotree.cpp:2529-2536)
    plk::NewtonRun nr;
    nr.start(length, min_len, max_len, min_len, max_nr);
    while (!nr.done()) {
        double d1, d2;
        st = plk_likelihood_derv(ctx, id, nr.x(), &d1, &d2, nullptr);
        // a non-finite derivative is an error of the normal kernel; the safe
        // kernel warns and continues with df = ddf = 0
        // (tree/phylokernelnew.h:2577-2579, 2629-2632)
        if (st == PLK_ERR_NONFINITE && safe) st = PLK_OK;
        if (st != PLK_OK) return st;
        nr.feed(-d1, -d2);
    }
    if (nr.failed()) return fail(PLK_ERR_NEWTON, "maximum number of iterations exceeded in minimizeNewton");
    const int evals = nr.evals();
    double optx = nr.result();
    // divergence guard of optimizeOneBranch (tree/phylotree.cpp:2616-2625):
    // lnL at the length the last derivative evaluation left in the tree
    // against lnL at the starting length
    if (optx > max_len * 0.95) {
        double opt_lh, orig_lh;
        if ((st = plk_likelihood_from_buffer(ctx, id, nr.last_x(), &opt_lh, nullptr)) != PLK_OK) return st;
        if ((st = plk_likelihood_from_buffer(ctx, id, length, &orig_lh, nullptr)) != PLK_OK) return st;
        if (orig_lh > opt_lh) optx = length;
    }
    *out_length = optx;
    if (n_derv) *n_derv = evals;
    if (out_lnl) {
        if ((st = plk_likelihood_from_buffer(ctx, id, optx, out_lnl, nullptr)) != PLK_OK) return st;
    }
    return PLK_OK;
}

}  // extern "C"

namespace {

template <typename T, int N, int C>
plk_status launch_sweep_t(plk_context* ctx, Instance* in, SweepArgs& a, int& grid_out) {
    // codon: the codon tile's E and U^-1 need more than 64 KB in FP64, and
    // one workgroup per CU is enough for its 64-pattern tiles
    const int kLdsMax = N == kCodonN && ctx->lds_per_cu > 64 * 1024 ? ctx->lds_per_cu : 64 * 1024;
    const SweepSmem L0(N, C, in->d.state_unknown, in->sweep_table_size, a.nedges, (int)sizeof(T), 0);
    if (L0.total > kLdsMax) return fail(PLK_ERR_UNSUPPORTED, "sweep LDS need %d bytes", L0.total);
    int nb = 0;
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_sweep<T, N, C>, kThreads, L0.total));
    if (nb < 1) return fail(PLK_ERR_UNSUPPORTED, "sweep kernel does not fit a CU");
    const int PB = kThreads / C;
    const int ntiles = in->Ppad / PB;
    const int per_cu = std::min(nb, ctx->sweep_blocks_per_cu);
    int grid = ctx->num_cu * per_cu;
    grid = std::max(1, std::min(grid, std::min(ntiles, 2 * kThreads)));   // the exchange polls 2 per thread
    // theta (and the per-pattern scalars) of each workgroup's tiles resident
    // in LDS when they fit without lowering the workgroups per CU
    int lds_tiles = 0, smem = L0.total;
    if (ctx->sweep_lds) {
        const int tiles = (ntiles + grid - 1) / grid;
        const SweepSmem L1(N, C, in->d.state_unknown, in->sweep_table_size, a.nedges, (int)sizeof(T), tiles);
        int nb1 = 0;
        if (L1.total <= kLdsMax)
            HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb1, k_sweep<T, N, C>, kThreads, L1.total));
        if (nb1 >= per_cu) {
            lds_tiles = tiles;
            smem = L1.total;
        }
    }
    a.lds_tiles = lds_tiles;
    // exactly per_cu workgroups on every CU: with room for one more, the
    // dispatcher may stack three on some CUs and leave others with one, and
    // every grid-wide exchange waits for the slowest CU
    if (ctx->lds_per_cu > 0) {
        const int pad = ctx->lds_per_cu / (per_cu + 1) + 256;
        int nbp = 0;
        if (pad > smem && pad <= kLdsMax &&
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&nbp, k_sweep<T, N, C>, kThreads, pad) == hipSuccess &&
            nbp == per_cu)
            smem = pad;
    }
    plk_status st = grow(ctx, in, &in->sw_part, in->sw_part_cap, 6 * (size_t)grid);   // three slot buffers
    if (st != PLK_OK) return st;
    HIP_TRY(hipMemsetD32Async(reinterpret_cast<hipDeviceptr_t>(in->sw_part), kSlotEmptyWord,
                              12 * (size_t)grid, ctx->stream));
    const int nflags = 2 * kFlagStride;
    if ((st = grow(ctx, in, &in->sw_flags, in->sw_flags_cap, (size_t)nflags)) != PLK_OK) return st;
    HIP_TRY(hipMemsetAsync(in->sw_flags, 0, sizeof(unsigned) * nflags, ctx->stream));
    a.part = in->sw_part;
    a.release = in->sw_flags;
    void* args[] = {&a};
    HIP_TRY(hipLaunchCooperativeKernel((const void*)k_sweep<T, N, C>, dim3(grid), dim3(kThreads),
                                       args, (unsigned)smem, ctx->stream));
    grid_out = grid;
    return PLK_OK;
}

template <typename T, int N>
plk_status launch_sweep_n(plk_context* ctx, Instance* in, SweepArgs& a, int& grid) {
    switch (in->d.ncat) {
    case 1: return launch_sweep_t<T, N, 1>(ctx, in, a, grid);
    case 2: return launch_sweep_t<T, N, 2>(ctx, in, a, grid);
    case 4: return launch_sweep_t<T, N, 4>(ctx, in, a, grid);
    case 8:
        // codon tiles are 64 patterns, the sweep's tiles 256/C: C <= 4
        if constexpr (N != kCodonN) return launch_sweep_t<T, N, 8>(ctx, in, a, grid);
        break;
    }
    return fail(PLK_ERR_UNSUPPORTED, "ncat %d not built for %d states in the sweep", in->d.ncat, N);
}

plk_status launch_sweep(plk_context* ctx, Instance* in, SweepArgs& a, int& grid) {
    const bool f32 = in->d.precision == PLK_PREC_F32;
    switch (in->d.nstates) {
    case 4: return f32 ? launch_sweep_n<float, 4>(ctx, in, a, grid) : launch_sweep_n<double, 4>(ctx, in, a, grid);
    case 20: return f32 ? launch_sweep_n<float, 20>(ctx, in, a, grid) : launch_sweep_n<double, 20>(ctx, in, a, grid);
    case kCodonN:
        return f32 ? launch_sweep_n<float, kCodonN>(ctx, in, a, grid) : launch_sweep_n<double, kCodonN>(ctx, in, a, grid);
    }
    return fail(PLK_ERR_UNSUPPORTED, "nstates %d not built", in->d.nstates);
}

}  // namespace

namespace {

// k_derv_batch over tiles [begin, end) whose items all have this shape
// the device buffers of one derivative batch
struct DervBufs {
    DervItem* items;
    int* tile_item;
    double* tile_sums;
    double* out;
};

template <typename T, int N>
plk_status launch_derv_batch_n(plk_context* ctx, const DervBufs& b, int C, int begin, int end) {
    const dim3 grid(end - begin), blk(kThreads);
    switch (C) {
    case 1: hipLaunchKernelGGL((k_derv_batch<T, N, 1>), grid, blk, 0, ctx->stream, b.items, b.tile_item, begin, b.tile_sums); break;
    case 2: hipLaunchKernelGGL((k_derv_batch<T, N, 2>), grid, blk, 0, ctx->stream, b.items, b.tile_item, begin, b.tile_sums); break;
    case 4: hipLaunchKernelGGL((k_derv_batch<T, N, 4>), grid, blk, 0, ctx->stream, b.items, b.tile_item, begin, b.tile_sums); break;
    case 8: hipLaunchKernelGGL((k_derv_batch<T, N, 8>), grid, blk, 0, ctx->stream, b.items, b.tile_item, begin, b.tile_sums); break;
    default: return fail(PLK_ERR_UNSUPPORTED, "ncat %d not built", C);
    }
    HIP_TRY(hipGetLastError());
    return PLK_OK;
}

plk_status launch_derv_batch(plk_context* ctx, const DervBufs& b, int f32, int n, int C, int begin, int end) {
    switch (n) {
    case 4: return f32 ? launch_derv_batch_n<float, 4>(ctx, b, C, begin, end) : launch_derv_batch_n<double, 4>(ctx, b, C, begin, end);
    case 20: return f32 ? launch_derv_batch_n<float, 20>(ctx, b, C, begin, end) : launch_derv_batch_n<double, 20>(ctx, b, C, begin, end);
    case kCodonN: return f32 ? launch_derv_batch_n<float, kCodonN>(ctx, b, C, begin, end) : launch_derv_batch_n<double, kCodonN>(ctx, b, C, begin, end);
    }
    return fail(PLK_ERR_UNSUPPORTED, "nstates %d not built", n);
}

// One persistent work-list launch over items whose shape k_batch is built
// for (4 or 20 states, 4 categories); arguments already validated.
plk_status batch_fused(plk_context* ctx, const plk_batch_item* items, const int* modes, int nitems,
                       Instance* const* ins, double* lnl) {
    plk_status st;
    std::lock_guard<std::mutex> lock(ctx->mu);
    // largest partitions first (PhyloSuperTree::computePartitionOrder)
    std::vector<double> cost(nitems);
    std::vector<int> cls(nitems), ntiles_of(nitems);
    int total_tiles = 0, mix = 0, nclasses = 0;
    for (int i = 0; i < nitems; i++) {
        const Instance* in = ins[i];
        const double n = in->d.nstates;
        cost[i] = (double)in->d.npatterns * (items[i].nops + 1) * n * n;
        cls[i] = batch_class(in->d.nstates, in->d.precision == PLK_PREC_F32);
        const int PB = batch_tile_patterns(in->d.nstates, in->d.ncat);
        ntiles_of[i] = (in->d.npatterns + PB - 1) / PB;
        total_tiles += ntiles_of[i];
        if (!(mix & cls[i])) nclasses++;
        mix |= cls[i];
    }
    // launch shape: the tuned one for a recurring batch shape (class mix,
    // log2 tiles), else the default. Online tuning: a recurring shape's
    // first calls each run one candidate (fused or per-class launches x 1,
    // 2, 4, 8 workgroups per CU) and time it; results do not depend on the
    // shape, so no call repeats work. Batches of fewer tiles than a few per
    // CU keep the default. Likely-good shapes first, so exploring costs little.
    int lg = 0;
    while ((1 << lg) < total_tiles) lg++;
    const long key = (long)mix * 64 + lg;
    int split = 1, bpc = ctx->batch_blocks_per_cu > 0 ? ctx->batch_blocks_per_cu : 2;
    static const std::pair<int, int> kOrder[8] = {{1, 2}, {1, 4}, {0, 2}, {0, 4}, {1, 8}, {0, 8}, {1, 1}, {0, 1}};
    std::pair<int, int> cands[8];
    int ncand = 0;
    for (const auto& cd : kOrder)
        if (cd.first == 0 || nclasses > 1) cands[ncand++] = cd;
    plk_context::BatchTune* tune = nullptr;
    if (ctx->batch_blocks_per_cu == 0 && ctx->autotune && total_tiles >= 4 * ctx->num_cu) {
        plk_context::BatchTune& bt = ctx->batch_tune[key];
        if (bt.done) {
            split = bt.split;
            bpc = bt.bpc;
        } else {
            split = cands[bt.next].first;
            bpc = cands[bt.next].second;
            tune = &bt;
        }
    }
    // launch groups: one per class (split), or the DNA and protein classes
    // fused and each codon class alone; inside a group every item's tiles go
    // to one XCD segment, items bin-packed over the segments by cost (largest
    // first onto the least loaded), so an item's tables are fetched into one
    // XCD's L2
    auto group_of = [&](int c) { return (split || !(c & BC_ALL)) ? c : BC_ALL; };
    auto rank = [](int g) { return (g & BC_ALL) ? g + kBatchClasses * 64 : g; };   // fused / DNA / protein first
    std::vector<int> sorted(nitems);
    for (int i = 0; i < nitems; i++) sorted[i] = i;
    std::stable_sort(sorted.begin(), sorted.end(), [&](int a, int b) {
        const int ga = group_of(cls[a]), gb = group_of(cls[b]);
        if (ga != gb) return rank(ga) > rank(gb);
        return cost[a] > cost[b];
    });
    std::vector<int> order;
    order.reserve(nitems);
    struct Group {
        int mask;
        XcdSplit xs;
    };
    std::vector<Group> groups;
    int t = 0;
    for (int r0 = 0; r0 < nitems;) {
        const int g = group_of(cls[sorted[r0]]);
        int r1 = r0;
        while (r1 < nitems && group_of(cls[sorted[r1]]) == g) r1++;
        double load[kXcds] = {};
        std::vector<int> seg[kXcds];
        for (int r = r0; r < r1; r++) {
            int q = 0;
            for (int k = 1; k < kXcds; k++)
                if (load[k] < load[q]) q = k;
            load[q] += cost[sorted[r]];
            seg[q].push_back(sorted[r]);
        }
        Group G;
        G.mask = 0;
        for (int q = 0; q < kXcds; q++) {
            G.xs.b[q] = t;
            for (int i : seg[q]) {
                order.push_back(i);
                t += ntiles_of[i];
                G.mask |= cls[i];
            }
        }
        G.xs.b[kXcds] = t;
        groups.push_back(G);
        r0 = r1;
    }
    // one pinned upload: items | ops | tile list | root lengths
    std::vector<BatchItem> hb(nitems);
    std::vector<size_t> op_off(nitems);
    size_t tab_bytes = 0, root_doubles = 0, smem_prep = 0, nops_total = 0;
    int tiles_so_far = 0, max_nops = 0;
    // algorithmic bytes and flops of the batch, per instance class (index:
    // the BatchClass bit)
    double cls_bytes[kBatchClasses] = {}, cls_flops[kBatchClasses] = {};
    size_t cls_smem[kBatchClasses] = {};   // dynamic LDS of each class's tiles
    for (int r = 0; r < nitems; r++) {
        const int i = order[r];
        const Instance* in = ins[i];
        op_off[r] = nops_total;
        nops_total += items[i].nops;
        max_nops = std::max(max_nops, items[i].nops);
        const int PB = batch_tile_patterns(in->d.nstates, in->d.ncat);
        BatchItem& b = hb[r];
        b.I = inst_dev(in);
        b.m = model_dev(in);
        b.nops = items[i].nops;
        b.table_size = in->table_size;
        b.tile_begin = tiles_so_far;
        b.ntiles = (in->d.npatterns + PB - 1) / PB;
        b.node = items[i].node;
        b.dad = items[i].dad;
        b.n = in->d.nstates;
        b.C = in->d.ncat;
        b.f32 = in->d.precision == PLK_PREC_F32;
        b.mode = modes ? modes[i] : BATCH_LNL;
        tiles_so_far += b.ntiles;
        b.tables = reinterpret_cast<const void*>(tab_bytes);   // offsets until the buffers exist
        tab_bytes += round_up((size_t)items[i].nops * 2 * in->table_size * in->elem, 256);
        b.root = reinterpret_cast<const double*>(root_doubles);
        root_doubles += (size_t)(in->d.state_unknown + 2) * in->block;
        const size_t item_smem = batch_smem_bytes(b.n, b.C, in->d.state_unknown, in->table_size, in->elem);
        smem_prep = std::max(smem_prep, (size_t)(b.n * b.n + in->d.state_unknown * b.n + 2 * in->block) * 8);
        double ob = 0.0, of = 0.0;
        for (int o = 0; o < items[i].nops; o++) {
            double bb, ff;
            op_cost(in, items[i].ops[o], bb, ff);
            ob += bb;
            of += ff;
        }
        const int nint = 1 + (items[i].dad.is_tip ? 0 : 1);
        ob += (double)in->block * in->elem * nint + (items[i].dad.is_tip ? 1 : 0) + in->scale_elem * nint + 24.0;
        of += 3.0 * in->block;
        int cb = 0;
        while (!((batch_class(b.n, b.f32) >> cb) & 1)) cb++;
        cls_bytes[cb] += ob * in->d.npatterns;
        cls_flops[cb] += of * in->d.npatterns;
        cls_smem[cb] = std::max(cls_smem[cb], item_smem);
    }
    const size_t off_ops = round_up(sizeof(BatchItem) * nitems, 256);
    const size_t off_tiles = off_ops + round_up(sizeof(plk_op) * std::max<size_t>(1, nops_total), 256);
    const size_t off_lens = off_tiles + round_up(sizeof(int) * total_tiles, 256);
    const size_t up_bytes = off_lens + sizeof(double) * nitems;
    if ((st = grow(ctx, nullptr, &ctx->bt_upload, ctx->bt_upload_cap, up_bytes)) != PLK_OK) return st;
    if ((st = grow(ctx, nullptr, &ctx->bt_tables, ctx->bt_tables_cap, std::max<size_t>(256, tab_bytes))) != PLK_OK) return st;
    if ((st = grow(ctx, nullptr, &ctx->bt_roots, ctx->bt_roots_cap, root_doubles)) != PLK_OK) return st;
    if ((st = grow(ctx, nullptr, &ctx->bt_tile_sums, ctx->bt_tile_sums_cap, (size_t)total_tiles)) != PLK_OK) return st;
    if ((st = grow(ctx, nullptr, &ctx->bt_vals, ctx->bt_vals_cap, (size_t)nitems)) != PLK_OK) return st;
    if ((st = grow(ctx, nullptr, &ctx->bt_counter, ctx->bt_counter_cap, 64)) != PLK_OK) return st;
    unsigned char* hs = nullptr;
    if ((st = host_stage(ctx, up_bytes, &hs)) != PLK_OK) return st;
    const BatchItem* d_items = reinterpret_cast<const BatchItem*>(ctx->bt_upload);
    const plk_op* d_ops = reinterpret_cast<const plk_op*>(ctx->bt_upload + off_ops);
    const int* d_tiles = reinterpret_cast<const int*>(ctx->bt_upload + off_tiles);
    const double* d_lens = reinterpret_cast<const double*>(ctx->bt_upload + off_lens);
    plk_op* h_ops = reinterpret_cast<plk_op*>(hs + off_ops);
    int* h_tiles = reinterpret_cast<int*>(hs + off_tiles);
    double* h_lens = reinterpret_cast<double*>(hs + off_lens);
    for (int r = 0; r < nitems; r++) {
        const int i = order[r];
        BatchItem& b = hb[r];
        b.ops = d_ops + op_off[r];
        b.tables = ctx->bt_tables + reinterpret_cast<size_t>(b.tables);
        b.root = ctx->bt_roots + reinterpret_cast<size_t>(b.root);
        if (items[i].nops) std::memcpy(h_ops + op_off[r], items[i].ops, sizeof(plk_op) * items[i].nops);
        for (int t = 0; t < b.ntiles; t++) h_tiles[b.tile_begin + t] = r;
        h_lens[r] = items[i].length;
    }
    std::memcpy(hs, hb.data(), sizeof(BatchItem) * nitems);
    HIP_TRY(hipMemcpyAsync(ctx->bt_upload, hs, up_bytes, hipMemcpyHostToDevice, ctx->stream));
    ctx->h_stage_inflight = true;
    // tables: rows of ops per item so that the launch has a few thousand
    // workgroups (a single item still gets one per op)
    const int rows = std::max(1, std::min(max_nops, std::max(4, (4096 + nitems - 1) / nitems)));
    int pe = prof_begin(ctx);
    hipLaunchKernelGGL(k_prep_batch, dim3(nitems, rows), dim3(kThreads), smem_prep, ctx->stream,
                       d_items, d_lens, ctx->bt_roots, ctx->bt_counter, kXcds * kBatchClasses);
    HIP_TRY(hipGetLastError());
    prof_end(ctx, pe, K_PREP_BATCH, (double)tab_bytes, 0.0);
    // profiling record of one launch over the classes in mask
    auto prof_mask = [&](int pe2, int mask) {
        if (pe2 < 0) return;
        double b2 = 0.0, f2 = 0.0;
        for (int cb = 0; cb < kBatchClasses; cb++)
            if ((mask >> cb) & 1) {
                b2 += cls_bytes[cb];
                f2 += cls_flops[cb];
            }
        const int kind = mask == BC_DNA64 || mask == BC_DNA32 ? K_BATCH_DNA
                         : mask == BC_PROT64 || mask == BC_PROT32 ? K_BATCH_PROT
                         : mask == BC_COD64 || mask == BC_COD32 ? K_BATCH_CODON : K_BATCH;
        prof_end(ctx, pe2, kind, b2, f2);
    };
    // one launch per group, each with its own kXcds work counters
    auto launch = [&]() -> plk_status {
        for (size_t k = 0; k < groups.size(); k++) {
            const Group& G = groups[k];
            const int pe2 = prof_begin(ctx);
            const int fmask = __builtin_popcount(G.mask) > 1 ? BC_ALL : G.mask;   // fused DNA + protein
            const void* fn;
            switch (fmask) {
            case BC_DNA64: fn = (const void*)k_batch<BC_DNA64>; break;
            case BC_PROT64: fn = (const void*)k_batch<BC_PROT64>; break;
            case BC_DNA32: fn = (const void*)k_batch<BC_DNA32>; break;
            case BC_PROT32: fn = (const void*)k_batch<BC_PROT32>; break;
            case BC_COD64: fn = (const void*)k_batch<BC_COD64>; break;
            case BC_COD32: fn = (const void*)k_batch<BC_COD32>; break;
            default: fn = (const void*)k_batch<BC_ALL>; break;
            }
            size_t smem = 0;
            for (int cb = 0; cb < kBatchClasses; cb++)
                if ((G.mask >> cb) & 1) smem = std::max(smem, cls_smem[cb]);
            int nb = 0;
            HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, kThreads, smem));
            const int ntl = G.xs.b[kXcds] - G.xs.b[0];
            const int grid = std::max(1, std::min(ntl, ctx->num_cu * std::max(1, std::min(nb, bpc))));
            const BatchItem* a0 = d_items;
            const int* a1 = d_tiles;
            XcdSplit xs = G.xs;
            unsigned* a3 = ctx->bt_counter + kXcds * k;
            double* a4 = ctx->bt_tile_sums;
            void* args[] = {(void*)&a0, (void*)&a1, (void*)&xs, (void*)&a3, (void*)&a4};
            HIP_TRY(hipLaunchKernel(fn, dim3(grid), dim3(kThreads), args, smem, ctx->stream));
            prof_mask(pe2, G.mask);
        }
        return PLK_OK;
    };
    if (tune) {
        if (!ctx->tune_ev[0]) {
            HIP_TRY(hipEventCreate(&ctx->tune_ev[0]));
            HIP_TRY(hipEventCreate(&ctx->tune_ev[1]));
        }
        HIP_TRY(hipEventRecord(ctx->tune_ev[0], ctx->stream));
    }
    if ((st = launch()) != PLK_OK) return st;
    if (tune) HIP_TRY(hipEventRecord(ctx->tune_ev[1], ctx->stream));
    hipLaunchKernelGGL(k_batch_reduce, dim3((nitems + kThreads - 1) / kThreads), dim3(kThreads), 0,
                       ctx->stream, d_items, nitems, ctx->bt_tile_sums, ctx->bt_vals);
    HIP_TRY(hipGetLastError());
    std::vector<double> out(nitems);
    HIP_TRY(hipMemcpyAsync(out.data(), ctx->bt_vals, sizeof(double) * nitems, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    ctx->h_stage_inflight = false;
    if (tune) {
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, ctx->tune_ev[0], ctx->tune_ev[1]));
        if (ms < tune->best_ms) {
            tune->best_ms = ms;
            tune->split = split;
            tune->bpc = bpc;
        }
        if (++tune->next >= ncand) tune->done = true;
    }
    bool nonfinite = false;
    for (int r = 0; r < nitems; r++) {
        const int i = order[r];
        if (hb[r].mode == BATCH_THETA) {
            ins[i]->theta_ready = true;
            lnl[i] = 0.0;
            continue;
        }
        ins[i]->ptn_lh_ready = true;
        double v = out[r];
        if (!std::isfinite(v)) {
            if ((st = fixup_nonfinite(ctx, ins[i], &v, nullptr)) != PLK_OK) return st;
            nonfinite = true;
        }
        lnl[i] = v;
    }
    if (nonfinite) return fail(PLK_ERR_NONFINITE, "numerical underflow for lh-branch in a batch item; fix-up applied");
    return PLK_OK;
}

}  // namespace

extern "C" {

plk_status plk_likelihood_branch_batch(plk_context* ctx, const plk_batch_item* items, int nitems,
                                       double* lnl) {
    return plk_batch_evaluate(ctx, items, nullptr, nitems, lnl);
}

plk_status plk_batch_evaluate(plk_context* ctx, const plk_batch_item* items, const int* modes,
                              int nitems, double* lnl) {
    if (!ctx || nitems < 0 || (nitems > 0 && (!items || !lnl))) return fail(PLK_ERR_ARG, "invalid batch");
    if (modes)
        for (int i = 0; i < nitems; i++)
            if (modes[i] != PLK_BATCH_LNL && modes[i] != PLK_BATCH_THETA)
                return fail(PLK_ERR_ARG, "item %d: invalid mode %d", i, modes[i]);
    if (nitems == 0) return PLK_OK;
    plk_status st;
    std::vector<Instance*> ins(nitems);
    for (int i = 0; i < nitems; i++) {
        Instance* in = get_instance(ctx, items[i].instance);
        if ((st = ready(in)) != PLK_OK) return st;
        if ((st = check_ops(in, items[i].ops, items[i].nops)) != PLK_OK) return st;
        if ((st = check_end(in, items[i].node, false)) != PLK_OK) return st;
        if ((st = check_end(in, items[i].dad, true)) != PLK_OK) return st;
        if (!(items[i].length >= 0.0) || !std::isfinite(items[i].length)) return fail(PLK_ERR_ARG, "item %d: invalid length", i);
        ins[i] = in;
    }
    {
        std::vector<int> ids(nitems);
        for (int i = 0; i < nitems; i++) ids[i] = items[i].instance;
        std::sort(ids.begin(), ids.end());
        for (int i = 1; i < nitems; i++)
            if (ids[i] == ids[i - 1]) return fail(PLK_ERR_ARG, "instance %d appears twice in a batch", ids[i]);
    }
    // shapes the fused kernel is built for go into one launch; the rest
    // (61 states, category counts other than 4) run one traversal each
    std::vector<plk_batch_item> fused;
    std::vector<Instance*> fused_ins;
    std::vector<int> fused_at, fused_mode;
    bool nonfinite = false;
    for (int i = 0; i < nitems; i++) {
        const Instance* in = ins[i];
        if (in->d.ncat == 4 && in->nasc == 0 && !in->nonrev &&
            (in->d.nstates == 4 || in->d.nstates == 20 || in->d.nstates == kCodonN)) {
            fused.push_back(items[i]);
            fused_ins.push_back(ins[i]);
            fused_at.push_back(i);
            fused_mode.push_back(modes ? modes[i] : PLK_BATCH_LNL);
            continue;
        }
        if (modes && modes[i] == PLK_BATCH_THETA) {
            lnl[i] = 0.0;
            st = plk_likelihood_buffer(ctx, items[i].instance, items[i].ops, items[i].nops,
                                       items[i].node, items[i].dad);
        } else {
            st = plk_likelihood_branch(ctx, items[i].instance, items[i].ops, items[i].nops,
                                       items[i].node, items[i].dad, items[i].length, &lnl[i], nullptr);
        }
        if (st == PLK_ERR_NONFINITE) nonfinite = true;
        else if (st != PLK_OK) return st;
    }
    if (!fused.empty()) {
        std::vector<double> out(fused.size());
        st = batch_fused(ctx, fused.data(), fused_mode.data(), (int)fused.size(), fused_ins.data(),
                         out.data());
        if (st != PLK_OK && st != PLK_ERR_NONFINITE) return st;
        if (st == PLK_ERR_NONFINITE) nonfinite = true;
        for (size_t k = 0; k < fused.size(); k++) lnl[fused_at[k]] = out[k];
    }
    if (nonfinite) return fail(PLK_ERR_NONFINITE, "numerical underflow for lh-branch in a batch item; fix-up applied");
    return PLK_OK;
}

}  // extern "C"

namespace {

// plk_likelihood_derv_batch and its no-allocation form: the same items,
// kernels and reduction; the buffers come from the context's caching
// allocator, or from this GPU's scratch array (noalloc).
plk_status derv_batch_run(plk_context* ctx, const int* instances, const double* lengths, int n,
                          double* df, double* ddf, double* lnl, bool noalloc) {
    if (!ctx || n < 0 || (n > 0 && (!instances || !lengths || !df || !ddf)))
        return fail(PLK_ERR_ARG, "invalid derivative batch");
    if (n == 0) return PLK_OK;
    plk_status st;
    std::vector<Instance*> ins(n);
    for (int i = 0; i < n; i++) {
        ins[i] = get_instance(ctx, instances[i]);
        if ((st = ready(ins[i])) != PLK_OK) return st;
        if (!ins[i]->theta_ready) return fail(PLK_ERR_STATE, "item %d: theta buffer not computed", i);
        if (ins[i]->nasc) return fail(PLK_ERR_UNSUPPORTED, "item %d: +ASC instances take plk_likelihood_derv", i);
        if (ins[i]->nonrev)
            return fail(PLK_ERR_UNSUPPORTED, "item %d: state-basis instances take plk_likelihood_derv", i);
        if (!std::isfinite(lengths[i])) return fail(PLK_ERR_ARG, "item %d: invalid length", i);
    }
    std::lock_guard<std::mutex> lock(ctx->mu);
    // items grouped by shape, so each shape's tiles are one contiguous launch
    auto shape = [&](int i) {
        const Instance* in = ins[i];
        return (in->d.precision == PLK_PREC_F32 ? 1 : 0) * 100000 + in->d.nstates * 100 + in->d.ncat;
    };
    std::vector<int> order(n);
    for (int i = 0; i < n; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return shape(a) < shape(b); });
    std::vector<DervItem> hd(n);
    std::vector<int> tile_item;
    double bytes = 0.0;
    for (int r = 0; r < n; r++) {
        const Instance* in = ins[order[r]];
        DervItem& d = hd[r];
        d.I = inst_dev(in);
        d.m = model_dev(in);
        d.length = lengths[order[r]];
        d.tile_begin = (int)tile_item.size();
        d.ntiles = in->Ppad / (kThreads / in->d.ncat);
        for (int t = 0; t < d.ntiles; t++) tile_item.push_back(r);
        bytes += ((double)in->block * in->elem + 24.0) * in->d.npatterns;
    }
    const int total = (int)tile_item.size();
    DervBufs b{};
    if (noalloc) {
        // items | tile list | tile sums | results, 256-byte aligned, in the
        // scratch array; every tile writes its sums before the reduction
        // reads them, so nothing needs clearing
        size_t off = 0;
        auto carve = [&](size_t bytes) {
            const size_t at = off;
            off = round_up(off + bytes, 256);
            return at;
        };
        const size_t o_items = carve(sizeof(DervItem) * n), o_tiles = carve(sizeof(int) * (size_t)total);
        const size_t o_sums = carve(sizeof(double) * 3 * (size_t)total), o_out = carve(sizeof(double) * 3 * (size_t)n);
        if (off > kNoAllocScratchBytes)
            return fail(PLK_ERR_NOMEM, "derivative batch needs %zu bytes of scratch; the no-allocation form holds %zu",
                        off, kNoAllocScratchBytes);
        if (!ctx->scratch) return fail(PLK_ERR_DEVICE, "no scratch address for device %d", ctx->device);
        b.items = reinterpret_cast<DervItem*>(ctx->scratch + o_items);
        b.tile_item = reinterpret_cast<int*>(ctx->scratch + o_tiles);
        b.tile_sums = reinterpret_cast<double*>(ctx->scratch + o_sums);
        b.out = reinterpret_cast<double*>(ctx->scratch + o_out);
    } else {
        if ((st = grow(ctx, nullptr, &ctx->db_items, ctx->db_items_cap, (size_t)n)) != PLK_OK) return st;
        if ((st = grow(ctx, nullptr, &ctx->db_tile_item, ctx->db_tile_cap, (size_t)total)) != PLK_OK) return st;
        if ((st = grow(ctx, nullptr, &ctx->db_tile_sums, ctx->db_tile_sums_cap, 3 * (size_t)total)) != PLK_OK) return st;
        if ((st = grow(ctx, nullptr, &ctx->db_out, ctx->db_out_cap, 3 * (size_t)n)) != PLK_OK) return st;
        b = DervBufs{ctx->db_items, ctx->db_tile_item, ctx->db_tile_sums, ctx->db_out};
    }
    HIP_TRY(hipMemcpyAsync(b.items, hd.data(), sizeof(DervItem) * n, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(b.tile_item, tile_item.data(), sizeof(int) * total, hipMemcpyHostToDevice, ctx->stream));
    const int pe = prof_begin(ctx);
    for (int r = 0; r < n;) {
        int e = r;
        while (e < n && shape(order[e]) == shape(order[r])) e++;
        const Instance* in = ins[order[r]];
        st = launch_derv_batch(ctx, b, in->d.precision == PLK_PREC_F32, in->d.nstates, in->d.ncat,
                               hd[r].tile_begin, hd[e - 1].tile_begin + hd[e - 1].ntiles);
        if (st != PLK_OK) return st;
        r = e;
    }
    prof_end(ctx, pe, K_DERV, bytes, 0.0);
    hipLaunchKernelGGL(k_derv_reduce, dim3(n), dim3(kThreads), 0, ctx->stream, b.items, b.tile_sums, b.out);
    HIP_TRY(hipGetLastError());
    std::vector<double> out(3 * (size_t)n);
    HIP_TRY(hipMemcpyAsync(out.data(), b.out, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    bool nonfinite = false;
    for (int r = 0; r < n; r++) {
        const int i = order[r];
        df[i] = out[3 * r];
        ddf[i] = out[3 * r + 1];
        if (lnl) lnl[i] = out[3 * r + 2];
        if (!std::isfinite(df[i])) {
            // as plk_likelihood_derv; lnl = NaN marks the item
            df[i] = ddf[i] = 0.0;
            if (lnl) lnl[i] = std::nan("");
            nonfinite = true;
        }
    }
    if (nonfinite) return fail(PLK_ERR_NONFINITE, "numerical underflow for lh-derivative in a batch item");
    return PLK_OK;
}

}  // namespace

extern "C" {

plk_status plk_likelihood_derv_batch(plk_context* ctx, const int* instances, const double* lengths,
                                     int n, double* df, double* ddf, double* lnl) {
    return derv_batch_run(ctx, instances, lengths, n, df, ddf, lnl, false);
}

plk_status plk_likelihood_derv_batch_noalloc(plk_context* ctx, const int* instances, const double* lengths,
                                             int n, double* df, double* ddf, double* lnl) {
    return derv_batch_run(ctx, instances, lengths, n, df, ddf, lnl, true);
}


plk_status plk_optimize_sweep(plk_context* ctx, int id, const plk_sweep_op* ops, int nops,
                              const plk_sweep_branch* branches, int nbranches, double* lengths,
                              int nedges, double min_len, double max_len, int max_nr,
                              int* n_derv, double* lnl) {
    return plk_optimize_sweep_ex(ctx, id, ops, nops, branches, nbranches, lengths, nedges, min_len,
                                 max_len, max_nr, n_derv, nullptr, nullptr, 0, lnl);
}

plk_status plk_optimize_sweep_ex(plk_context* ctx, int id, const plk_sweep_op* ops, int nops,
                                 const plk_sweep_branch* branches, int nbranches, double* lengths,
                                 int nedges, double min_len, double max_len, int max_nr,
                                 int* n_derv, double* branch_len, double* lnl_out, int nlnl,
                                 double* lnl) {
    Instance* in = get_instance(ctx, id);
    plk_status st = ready(in);
    if (st != PLK_OK) return st;
    if (in->nasc) return fail(PLK_ERR_UNSUPPORTED, "the device sweep does not apply +ASC");
    if (in->nonrev) return fail(PLK_ERR_UNSUPPORTED, "the device sweep runs reversible models");
    if (nops < 0 || (nops > 0 && !ops) || nbranches < 1 || !branches || !lengths || nedges < 1 ||
        max_nr < 1 || !(min_len > 0.0) || !(max_len > min_len))
        return fail(PLK_ERR_ARG, "invalid sweep arguments");
    for (int e = 0; e < nedges; e++)
        if (!(lengths[e] >= 0.0) || !std::isfinite(lengths[e])) return fail(PLK_ERR_ARG, "invalid length %d", e);
    for (int o = 0; o < nops; o++) {
        const plk_sweep_op& op = ops[o];
        if (op.dad_slot < 0 || op.dad_slot >= in->d.nslots || op.nchildren != 2)
            return fail(PLK_ERR_ARG, "sweep op %d: bad dad slot / arity", o);
        for (int k = 0; k < 2; k++) {
            const plk_sweep_child& ch = op.child[k];
            const int lim = ch.is_tip ? in->d.ntaxa : in->d.nslots;
            if (ch.id < 0 || ch.id >= lim || ch.edge < 0 || ch.edge >= nedges ||
                (!ch.is_tip && ch.id == op.dad_slot))
                return fail(PLK_ERR_ARG, "sweep op %d child %d out of range", o, k);
        }
    }
    for (int b = 0; b < nbranches; b++) {
        const plk_sweep_branch& B = branches[b];
        if (B.op_begin < 0 || B.op_count < 0 || B.op_begin + B.op_count > nops || B.edge < 0 ||
            B.edge >= nedges)
            return fail(PLK_ERR_ARG, "sweep branch %d: bad op range / edge", b);
        if ((st = check_end(in, B.node, false)) != PLK_OK) return st;
        if ((st = check_end(in, B.dad, true)) != PLK_OK) return st;
        if (B.lnl_index < -1 || B.lnl_index >= nlnl)
            return fail(PLK_ERR_ARG, "sweep branch %d: lnl_index %d outside [-1, %d)", b, B.lnl_index, nlnl);
    }
    if (nlnl < 0 || (nlnl > 0 && !lnl_out)) return fail(PLK_ERR_ARG, "invalid lnl_out");
    if (nbranches > kSweepMaxBranches)
        return fail(PLK_ERR_UNSUPPORTED, "more than %d branches in one sweep", kSweepMaxBranches);
    std::lock_guard<std::mutex> lock(ctx->mu);
    if ((st = grow(ctx, in, &in->sw_ops, in->sw_ops_cap, (size_t)std::max(nops, 1))) != PLK_OK) return st;
    if ((st = grow(ctx, in, &in->sw_br, in->sw_br_cap, (size_t)nbranches)) != PLK_OK) return st;
    if ((st = grow(ctx, in, &in->sw_len, in->sw_len_cap, (size_t)nedges)) != PLK_OK) return st;
    if (!in->sw_nderv) {
        if ((st = dmalloc(ctx, in, &in->sw_nderv, sizeof(int) * (size_t)kSweepMaxBranches)) != PLK_OK) return st;
        if ((st = dmalloc(ctx, in, &in->sw_out, sizeof(double) * 8)) != PLK_OK) return st;
    }
    // per-branch lengths, then the requested lnL values
    if ((st = grow(ctx, in, &in->sw_res, in->sw_res_cap, (size_t)nbranches + (size_t)std::max(nlnl, 1))) != PLK_OK)
        return st;
    if (nops > 0)
        HIP_TRY(hipMemcpyAsync(in->sw_ops, ops, sizeof(plk_sweep_op) * nops, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(in->sw_br, branches, sizeof(plk_sweep_branch) * nbranches, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(in->sw_len, lengths, sizeof(double) * nedges, hipMemcpyHostToDevice, ctx->stream));
    SweepArgs a;
    a.I = inst_dev(in);
    a.m = model_dev(in);
    a.ops = in->sw_ops;
    a.br = in->sw_br;
    a.nbranches = nbranches;
    a.nedges = nedges;
    a.len = in->sw_len;
    a.min_len = min_len;
    a.max_len = max_len;
    a.max_nr = max_nr;
    a.table_size = in->sweep_table_size;
    a.part = nullptr;
    a.release = nullptr;
    a.nderv = in->sw_nderv;
    a.br_len = in->sw_res;
    a.lnl_out = in->sw_res + nbranches;
    a.nlnl = nlnl;
    a.out = in->sw_out;
    a.timing = nullptr;
    a.lds_tiles = 0;
    a.tabs = nullptr;
    if (in->d.nstates == kCodonN) {
        // codon: a step's tables live in HBM, two steps' worth
        const size_t elem = in->d.precision == PLK_PREC_F32 ? 4 : 8;
        if ((st = grow(ctx, in, &in->sw_tabs, in->sw_tabs_cap, 4 * (size_t)in->sweep_table_size * elem)) != PLK_OK)
            return st;
        a.tabs = in->sw_tabs;
    }
    static const bool want_timing = std::getenv("PLK_SWEEP_TIMING") != nullptr;
    if (want_timing) {
        if (!in->sw_timing) {
            if ((st = dmalloc(ctx, in, &in->sw_timing, sizeof(long long) * PH_COUNT)) != PLK_OK) return st;
            HIP_TRY(hipMemsetAsync(in->sw_timing, 0, sizeof(long long) * PH_COUNT, ctx->stream));
        }
        a.timing = in->sw_timing;
    }
    int grid = 0;
    const int pe = prof_begin(ctx);
    st = launch_sweep(ctx, in, a, grid);
    if (st != PLK_OK) return st;
    prof_end(ctx, pe, K_SWEEP, 0.0, 0.0);
    std::vector<int> nd(nbranches);
    double out[2];
    HIP_TRY(hipMemcpyAsync(lengths, in->sw_len, sizeof(double) * nedges, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipMemcpyAsync(nd.data(), in->sw_nderv, sizeof(int) * nbranches, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipMemcpyAsync(out, in->sw_out, sizeof(double) * 2, hipMemcpyDeviceToHost, ctx->stream));
    std::vector<double> res((size_t)nbranches + (size_t)nlnl);
    if (branch_len || nlnl > 0)
        HIP_TRY(hipMemcpyAsync(res.data(), in->sw_res, sizeof(double) * res.size(), hipMemcpyDeviceToHost,
                               ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (branch_len) std::memcpy(branch_len, res.data(), sizeof(double) * nbranches);
    if (nlnl > 0) std::memcpy(lnl_out, res.data() + nbranches, sizeof(double) * nlnl);
    // theta and pattern lnL hold the last branch only after a complete sweep
    in->theta_ready = out[1] == 0.0;
    in->ptn_lh_ready = out[1] == 0.0;
    if (a.timing) {
        long long t[PH_COUNT];
        int khz = 0;
        HIP_TRY(hipMemcpy(t, a.timing, sizeof(t), hipMemcpyDeviceToHost));
        HIP_TRY(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, ctx->device));
        const double us = 1e3 / (double)khz;
        std::fprintf(stderr, "[plk sweep timing] grid %d, LDS tiles %d: ops %.1f us, theta %.1f us, "
                     "eval work %.1f us, exchange %.1f us, total %.1f us (cumulative)\n", grid,
                     a.lds_tiles, t[PH_OPS] * us,
                     t[PH_THETA] * us, t[PH_EVAL_WORK] * us, t[PH_EXCHANGE] * us, t[PH_TOTAL] * us);
    }
    if (n_derv) std::memcpy(n_derv, nd.data(), sizeof(int) * nbranches);
    double final_lnl = out[0];
    bool final_nonfinite = false;
    if (out[1] == 0.0 && !std::isfinite(final_lnl)) {
        // computeLikelihoodFromBuffer's per-pattern fix-up (tree/phylokernelnew.h:3371-3380)
        if ((st = fixup_nonfinite(ctx, in, &final_lnl, nullptr)) != PLK_OK) return st;
        final_nonfinite = !in->d.safe_numeric;
    }
    if (lnl) *lnl = final_lnl;
    if (pe >= 0) {
        // algorithmic bytes: the ops, theta per branch, and every evaluation
        // pass over theta (+invar, freq; lnL passes also buffer_scale)
        const double P = in->d.npatterns, s = (double)in->elem, blk = (double)in->block;
        double bytes = 0.0, flops = 0.0;
        for (int o = 0; o < nops; o++) {
            plk_op tmp{};
            tmp.nchildren = 2;
            for (int k = 0; k < 2; k++) tmp.child[k].is_tip = ops[o].child[k].is_tip;
            double b, f;
            op_cost(in, tmp, b, f);
            bytes += b;
            flops += f;
        }
        long evals = 0;
        for (int b = 0; b < nbranches; b++) {
            evals += nd[b];
            const int nint = 1 + (branches[b].dad.is_tip ? 0 : 1);
            bytes += blk * s * nint + (branches[b].dad.is_tip ? 1 : 0) + in->scale_elem * nint +
                     blk * s + 8.0;
        }
        bytes += evals * (blk * s + 16.0) + (blk * s + 32.0);
        flops += evals * 6.0 * blk;
        ctx->prof_recs.back().bytes = bytes * P;
        ctx->prof_recs.back().flops = flops * P;
    }
    const int status = (int)out[1];
    if (final_nonfinite)
        return fail(PLK_ERR_NONFINITE, "numerical underflow (lh-from-buffer) after the Newton sweep; use the safe kernel");
    if (status == PLK_ERR_NONFINITE) return fail(PLK_ERR_NONFINITE, "non-finite derivative in the Newton sweep");
    if (status == PLK_ERR_DEVICE) return fail(PLK_ERR_DEVICE, "grid reduction in the Newton sweep timed out");
    if (status != 0) return fail((plk_status)status, "Newton sweep failed (status %d)", status);
    return PLK_OK;
}

plk_status plk_set_cost_matrix(plk_context* ctx, int id, const uint32_t* cost) {
    Instance* in = get_instance(ctx, id);
    if (!in || !cost) return fail(PLK_ERR_ARG, "null argument");
    const int n = in->d.nstates, SU = in->d.state_unknown;
    if ((n != 4 && n != kProtN && n != kCodonN) || SU > kSankoffMaxSU)
        return fail(PLK_ERR_UNSUPPORTED, "Sankoff parsimony for %d states", n);
    // the tip costs of computeTipPartialParsimony (tree/phylotreepars.cpp:561-633):
    // a state's row of the cost matrix, an ambiguity set 0 inside and the least
    // cost into it outside (DNA bit mask + 3, protein B/Z/J), STATE_UNKNOWN 0
    static const int ambi_aa[3] = {4 + 8, 32 + 64, 512 + 1024};
    std::vector<uint32_t> h((size_t)n * n + (size_t)(SU + 1) * n, 0u);
    std::memcpy(h.data(), cost, sizeof(uint32_t) * n * n);
    uint32_t* tip = h.data() + (size_t)n * n;
    std::memcpy(tip, cost, sizeof(uint32_t) * n * n);
    for (int s = n; s < SU; s++) {   // ambiguity codes (none for codons)
        const int mask = n == 4 ? s - n + 1 : ambi_aa[s - kProtN];
        for (int i = 0; i < n; i++) {
            uint32_t v = 0;
            if (!(mask & (1 << i))) {
                v = 0xFFFFFFFFu;
                for (int j = 0; j < n; j++)
                    if ((mask & (1 << j)) && cost[i * n + j] < v) v = cost[i * n + j];
            }
            tip[(size_t)s * n + i] = v;
        }
    }
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (!in->sk_cost) {
        plk_status st = dmalloc(ctx, in, &in->sk_cost, h.size() * sizeof(uint32_t));
        if (st != PLK_OK) return st;
    }
    HIP_TRY(hipMemcpyAsync(in->sk_cost, h.data(), h.size() * sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    in->sk_ready = true;
    return PLK_OK;
}

plk_status plk_parsimony_sankoff(plk_context* ctx, int id, const plk_op* ops, int nops, plk_end node,
                                 plk_end dad, double* score, double* branch_subst) {
    Instance* in = get_instance(ctx, id);
    if (!in) return fail(PLK_ERR_ARG, "invalid instance");
    if (!in->patterns_set) return fail(PLK_ERR_STATE, "patterns not set");
    if (!in->sk_ready) return fail(PLK_ERR_STATE, "cost matrix not set");
    if (!score) return fail(PLK_ERR_ARG, "null score");
    plk_status st = check_ops(in, ops, nops);
    if (st != PLK_OK) return st;
    if ((st = check_end(in, node, false)) != PLK_OK) return st;
    if ((st = check_end(in, dad, true)) != PLK_OK) return st;
    std::vector<plk_op> plain(ops, ops + nops);   // parsimony has no direction
    for (plk_op& o : plain)
        for (int k = 0; k < 2; k++) o.child[k].is_tip &= kNrTip;
    dad.is_tip &= kNrTip;
    const int n = in->d.nstates, SU = in->d.state_unknown;
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (!in->sk_part) {
        if ((st = dmalloc(ctx, in, &in->sk_part, (size_t)in->d.nslots * in->Ppad * n * sizeof(uint32_t))) != PLK_OK)
            return st;
    }
    if (nops > in->ops_cap) {
        // the op buffer is shared with the likelihood calls: grow it as stage_ops does
        if ((st = stage_ops(ctx, in, plain.data(), nops)) != PLK_OK) return st;
    }
    if (nops) HIP_TRY(hipMemcpyAsync(in->ops, plain.data(), sizeof(plk_op) * nops, hipMemcpyHostToDevice, ctx->stream));
    const int grid = (in->d.npatterns + kThreads - 1) / kThreads;
    const int pe = prof_begin(ctx);
    if (n == 4)
        hipLaunchKernelGGL(k_sankoff<4>, dim3(grid), dim3(kThreads), 0, ctx->stream, in->ops, nops, in->sk_cost,
                           in->sk_cost + 16, SU, in->states, in->d.npatterns, in->Ppad, in->sk_part, node, dad,
                           in->freq, in->block_sums);
    else if (n == kProtN)
        hipLaunchKernelGGL(k_sankoff<kProtN>, dim3(grid), dim3(kThreads), 0, ctx->stream, in->ops, nops, in->sk_cost,
                           in->sk_cost + kProtN * kProtN, SU, in->states, in->d.npatterns, in->Ppad, in->sk_part,
                           node, dad, in->freq, in->block_sums);
    else
        hipLaunchKernelGGL(k_sankoff<kCodonN>, dim3(grid), dim3(kThreads), 0, ctx->stream, in->ops, nops,
                           in->sk_cost, in->sk_cost + kCodonN * kCodonN, SU, in->states, in->d.npatterns, in->Ppad,
                           in->sk_part, node, dad, in->freq, in->block_sums);
    HIP_TRY(hipGetLastError());
    prof_end(ctx, pe, K_SANKOFF, (double)in->d.npatterns * n * 4 * (2.0 * nops + 2), 0.0);
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(kThreads), 0, ctx->stream, in->block_sums, grid, 3, in->red_out);
    HIP_TRY(hipGetLastError());
    if ((st = fetch_scalars(ctx, in, 2)) != PLK_OK) return st;
    *score = ctx->h_out[0];
    if (branch_subst) *branch_subst = ctx->h_out[1];
    return PLK_OK;
}

plk_status plk_parsimony_branch(plk_context* ctx, int id, const plk_op* ops, int nops, plk_end node,
                                plk_end dad, double* score, double* branch_subst) {
    Instance* in = get_instance(ctx, id);
    if (!in) return fail(PLK_ERR_ARG, "invalid instance");
    if (!in->patterns_set) return fail(PLK_ERR_STATE, "patterns not set");
    if (!score) return fail(PLK_ERR_ARG, "null score");
    plk_status st = check_ops(in, ops, nops);
    if (st != PLK_OK) return st;
    if ((st = check_end(in, node, true)) != PLK_OK) return st;
    if ((st = check_end(in, dad, true)) != PLK_OK) return st;
    // parsimony has no direction: a state-basis instance's flags are dropped
    std::vector<plk_op> plain;
    if (in->nonrev && nops > 0) {
        plain.assign(ops, ops + nops);
        for (plk_op& o : plain)
            for (int k = 0; k < 2; k++) o.child[k].is_tip &= kNrTip;
        ops = plain.data();
        node.is_tip &= kNrTip;
        dad.is_tip &= kNrTip;
    }
    if (node.is_tip && dad.is_tip && in->d.ntaxa > 2) return fail(PLK_ERR_ARG, "branch between two tips");
    if (in->d.nstates == kCodonN && in->d.ntaxa > 65535)
        return fail(PLK_ERR_UNSUPPORTED, "%d taxa: codon change counts are 16-bit", in->d.ntaxa);
    std::lock_guard<std::mutex> lock(ctx->mu);
    const int n = in->d.nstates;
    const bool planes = n != kCodonN;
    const size_t Ppad = in->Ppad, nw = Ppad / 32, nslots = in->d.nslots;
    if (planes && !in->pars_planes) {
        if ((st = dmalloc(ctx, in, &in->pars_tip_planes, (size_t)in->d.ntaxa * n * nw * 4)) != PLK_OK) return st;
        if ((st = dmalloc(ctx, in, &in->pars_planes, nslots * n * nw * 4)) != PLK_OK) return st;
        if ((st = dmalloc(ctx, in, &in->pars_counts, nslots * nw * 8)) != PLK_OK) return st;
    }
    if (!planes && !in->pars_sets) {
        if ((st = dmalloc(ctx, in, &in->pars_sets, nslots * Ppad * 8)) != PLK_OK) return st;
        if ((st = dmalloc(ctx, in, &in->pars_changes, nslots * Ppad * 2)) != PLK_OK) return st;
    }
    // levels: an op's level is one more than the level of the ops in this
    // list that write its children's slots; ops of one level are independent.
    // A list that writes a slot twice, or reads a slot a later op writes,
    // keeps its order, one op per level.
    std::vector<int> level(nops, 0), writer(nslots, -1);
    bool ordered = true;
    for (int o = 0; o < nops && ordered; o++) {
        for (int k = 0; k < ops[o].nchildren; k++)
            if (!ops[o].child[k].is_tip && writer[ops[o].child[k].id] >= 0)
                level[o] = std::max(level[o], level[writer[ops[o].child[k].id]] + 1);
        if (writer[ops[o].dad_slot] >= 0) ordered = false;
        writer[ops[o].dad_slot] = o;
    }
    for (int o = 0; o < nops && ordered; o++)
        for (int k = 0; k < ops[o].nchildren; k++)
            if (!ops[o].child[k].is_tip && writer[ops[o].child[k].id] > o) ordered = false;
    if (!ordered)
        for (int o = 0; o < nops; o++) level[o] = o;
    const int nlevels = nops ? *std::max_element(level.begin(), level.end()) + 1 : 0;
    std::vector<plk_op> sorted(nops);
    std::vector<int> level_start(nlevels + 1, 0);
    for (int o = 0; o < nops; o++) level_start[level[o] + 1]++;
    for (int l = 0; l < nlevels; l++) level_start[l + 1] += level_start[l];
    {
        std::vector<int> fill(level_start.begin(), level_start.end() - 1);
        for (int o = 0; o < nops; o++) sorted[fill[level[o]]++] = ops[o];
    }
    // LDS form (DNA, protein, an ordered list): number the branch ends the
    // call touches, inputs (tips, slots of earlier calls) first, then the
    // slots the ops write; the widest word count whose working set fits
    // 60 KB of LDS (two workgroups per CU), else the HBM form
    std::vector<plk_end> inputs;
    std::vector<int4> lops;
    int pars_w = 0, bnode = -1, bdad = -1, nends = 0;
    if (planes && ordered) {
        std::vector<int> of_tip(in->d.ntaxa, -1), of_slot(nslots, -1);
        auto input = [&](plk_end e) {
            int& m = e.is_tip ? of_tip[e.id] : of_slot[e.id];
            if (m < 0 && (e.is_tip || writer[e.id] < 0)) {
                m = (int)inputs.size();
                inputs.push_back(e);
            }
        };
        for (const plk_op& op : sorted)
            for (int k = 0; k < 2; k++) input(plk_end{op.child[k].is_tip, op.child[k].id});
        input(node);
        input(dad);
        nends = (int)inputs.size();
        for (const plk_op& op : sorted) of_slot[op.dad_slot] = nends++;
        auto idx = [&](int is_tip, int id) { return is_tip ? of_tip[id] : of_slot[id]; };
        lops.resize(nops);
        for (int o = 0; o < nops; o++)
            lops[o] = make_int4(idx(sorted[o].child[0].is_tip, sorted[o].child[0].id),
                                idx(sorted[o].child[1].is_tip, sorted[o].child[1].id), of_slot[sorted[o].dad_slot],
                                sorted[o].dad_slot);
        bnode = idx(node.is_tip, node.id);
        bdad = idx(dad.is_tip, dad.id);
        for (int W : {8, 4, 2, 1})
            if (pars_lds_bytes(n, W, nends) <= 60 * 1024 && nw % W == 0) {
                pars_w = W;
                break;
            }
    }
    // one upload: sorted ops | level starts | LDS ops | LDS inputs
    const size_t off_lvl = round_up(sizeof(plk_op) * nops, 16);
    const size_t off_lops = off_lvl + round_up(sizeof(int) * (nlevels + 1), 16);
    const size_t off_in = off_lops + round_up(sizeof(int4) * lops.size(), 16);
    const size_t need = off_in + sizeof(plk_end) * inputs.size();
    if ((size_t)in->pars_ops_cap < need) {
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        dfree(ctx, in, in->pars_ops, in->pars_ops_cap);
        in->pars_ops = nullptr;
        in->pars_ops_cap = 0;
        const size_t cap = std::max(need, (size_t)64 * (in->d.ntaxa + nslots + 16));
        if ((st = dmalloc(ctx, in, &in->pars_ops, cap)) != PLK_OK) return st;
        in->pars_ops_cap = (int)cap;
    }
    std::vector<unsigned char> up(need);
    if (nops) std::memcpy(up.data(), sorted.data(), sizeof(plk_op) * nops);
    std::memcpy(up.data() + off_lvl, level_start.data(), sizeof(int) * (nlevels + 1));
    if (!lops.empty()) std::memcpy(up.data() + off_lops, lops.data(), sizeof(int4) * lops.size());
    if (!inputs.empty()) std::memcpy(up.data() + off_in, inputs.data(), sizeof(plk_end) * inputs.size());
    HIP_TRY(hipMemcpyAsync(in->pars_ops, up.data(), need, hipMemcpyHostToDevice, ctx->stream));
    unsigned char* dbase = reinterpret_cast<unsigned char*>(in->pars_ops);
    const int* d_level_start = reinterpret_cast<const int*>(dbase + off_lvl);
    const int4* d_lops = reinterpret_cast<const int4*>(dbase + off_lops);
    const plk_end* d_inputs = reinterpret_cast<const plk_end*>(dbase + off_in);
    ParsDev D;
    D.P = in->d.npatterns;
    D.Ppad = (int)Ppad;
    D.nwords = (int)nw;
    D.n = n;
    D.SU = in->d.state_unknown;
    D.seq = n == 4 ? 0 : (n == 20 ? 1 : 2);
    D.states = in->states;
    D.freq = in->freq;
    D.tip_planes = in->pars_tip_planes;
    D.planes = in->pars_planes;
    D.sets = in->pars_sets;
    D.counts = in->pars_counts;
    D.changes = in->pars_changes;
    if (planes && !in->pars_tips_ready) {
        const int g = (int)(((size_t)in->d.ntaxa * nw + kThreads - 1) / kThreads);
        if (n == 4) hipLaunchKernelGGL(k_pars_tip_planes<4>, dim3(g), dim3(kThreads), 0, ctx->stream, D, in->d.ntaxa);
        else hipLaunchKernelGGL(k_pars_tip_planes<20>, dim3(g), dim3(kThreads), 0, ctx->stream, D, in->d.ntaxa);
        HIP_TRY(hipGetLastError());
        in->pars_tips_ready = true;
    }
    const int grid = pars_w ? (int)(nw / pars_w)
                     : planes ? (int)((nw + kParsWords - 1) / kParsWords) : (int)((Ppad + kThreads - 1) / kThreads);
    const int pe = prof_begin(ctx);
    if (pars_w) {
        const size_t smem = pars_lds_bytes(n, pars_w, nends);
        const int nin = (int)inputs.size();
        auto go = [&](auto kern) {
            hipLaunchKernelGGL(kern, dim3(grid), dim3(kThreads), smem, ctx->stream, D, d_lops, d_level_start,
                               nlevels, d_inputs, nin, nends, bnode, bdad, in->block_sums);
        };
        if (n == 4) {
            if (pars_w == 8) go(k_pars_lds<4, 8>);
            else if (pars_w == 4) go(k_pars_lds<4, 4>);
            else if (pars_w == 2) go(k_pars_lds<4, 2>);
            else go(k_pars_lds<4, 1>);
        } else {
            if (pars_w == 8) go(k_pars_lds<20, 8>);
            else if (pars_w == 4) go(k_pars_lds<20, 4>);
            else if (pars_w == 2) go(k_pars_lds<20, 2>);
            else go(k_pars_lds<20, 1>);
        }
    } else if (n == 4)
        hipLaunchKernelGGL(k_pars_levels<4>, dim3(grid), dim3(kThreads), 0, ctx->stream, D, in->pars_ops,
                           d_level_start, nlevels, node, dad, in->block_sums);
    else if (n == 20)
        hipLaunchKernelGGL(k_pars_levels<20>, dim3(grid), dim3(kThreads), 0, ctx->stream, D, in->pars_ops,
                           d_level_start, nlevels, node, dad, in->block_sums);
    else
        hipLaunchKernelGGL(k_pars_sets, dim3(grid), dim3(kThreads), 0, ctx->stream, D, in->pars_ops, nops, node,
                           dad, in->block_sums);
    HIP_TRY(hipGetLastError());
    {
        // algorithmic bytes per pattern: an internal end is read as its sets
        // and count, a tip as its planes (DNA, protein) or state byte (codon);
        // every op writes one set and count
        const double set_b = planes ? n / 8.0 : 8.0, cnt_b = planes ? 8.0 / 32 : 2.0;
        const double tip_b = planes ? n / 8.0 : 1.0;
        double b = 0.0;
        for (int o = 0; o < nops; o++)
            for (int k = 0; k < 2; k++) b += ops[o].child[k].is_tip ? tip_b : set_b + cnt_b;
        b += nops * (set_b + cnt_b);
        for (const plk_end& e : {node, dad}) b += e.is_tip ? tip_b : set_b + cnt_b;
        prof_end(ctx, pe, K_PARSIMONY, b * in->d.npatterns, 0.0);
    }
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(kThreads), 0, ctx->stream, in->block_sums, grid, 2, in->red_out);
    HIP_TRY(hipGetLastError());
    if ((st = fetch_scalars(ctx, in, 2)) != PLK_OK) return st;
    *score = ctx->h_out[0];
    if (branch_subst) *branch_subst = ctx->h_out[1];
    return PLK_OK;
}

plk_status plk_copy_bandwidth(plk_context* ctx, size_t bytes, int reps, double* gbps, double* avg_ms) {
    if (!ctx || bytes < 4096 || reps < 1 || !gbps) return fail(PLK_ERR_ARG, "invalid copy measurement");
    std::lock_guard<std::mutex> lock(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    const size_t n = bytes / 32;   // 32-byte vectors
    unsigned char *a = nullptr, *b = nullptr;
    plk_status st = dmalloc(ctx, nullptr, &a, n * 32);
    if (st != PLK_OK) return st;
    if ((st = dmalloc(ctx, nullptr, &b, n * 32)) != PLK_OK) {
        dfree(ctx, nullptr, a, n * 32);
        return st;
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMemsetAsync(a, 0, n * 32, ctx->stream);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    const dim3 grid(ctx->num_cu * 16), blk(kThreads);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_copy, grid, blk, 0, ctx->stream, reinterpret_cast<const double4*>(a),
                           reinterpret_cast<double4*>(b), n);   // warm-up
        e = hipEventRecord(e0, ctx->stream);
    }
    for (int r = 0; r < reps && e == hipSuccess; r++) {
        hipLaunchKernelGGL(k_copy, grid, blk, 0, ctx->stream, reinterpret_cast<const double4*>(a),
                           reinterpret_cast<double4*>(b), n);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(e1, ctx->stream);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    dfree(ctx, nullptr, a, n * 32);
    dfree(ctx, nullptr, b, n * 32);
    if (e != hipSuccess) return fail(PLK_ERR_DEVICE, "copy measurement: %s", hipGetErrorString(e));
    *gbps = 2.0 * (double)n * 32 * reps / (ms * 1e-3) * 1e-9;
    if (avg_ms) *avg_ms = ms / reps;
    return PLK_OK;
}

plk_status plk_rell(plk_context* ctx, int id, const float* weights, int nboot, double* out) {
    Instance* in = get_instance(ctx, id);
    if (!in || !weights || !out || nboot <= 0) return fail(PLK_ERR_ARG, "invalid argument");
    if (!in->ptn_lh_ready) return fail(PLK_ERR_STATE, "no pattern likelihoods evaluated yet");
    std::lock_guard<std::mutex> lock(ctx->mu);
    const int P = in->d.npatterns;
    float* dw = nullptr;
    double* dout = nullptr;
    plk_status st = dmalloc(ctx, nullptr, &dw, (size_t)nboot * P * sizeof(float));
    if (st != PLK_OK) return st;
    st = dmalloc(ctx, nullptr, &dout, (size_t)nboot * sizeof(double));
    if (st != PLK_OK) {
        dfree(ctx, nullptr, dw, (size_t)nboot * P * sizeof(float));
        return st;
    }
    hipError_t e = hipMemcpyAsync(dw, weights, (size_t)nboot * P * sizeof(float), hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_rell, dim3(nboot), dim3(kThreads), 0, ctx->stream, dw, in->ptn_lh, P, dout);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(out, dout, nboot * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    dfree(ctx, nullptr, dw, (size_t)nboot * P * sizeof(float));
    dfree(ctx, nullptr, dout, (size_t)nboot * sizeof(double));
    if (e != hipSuccess) return fail(PLK_ERR_DEVICE, "rell: %s", hipGetErrorString(e));
    return PLK_OK;
}

plk_status plk_profile_start(plk_context* ctx) {
    if (!ctx) return fail(PLK_ERR_ARG, "null context");
    std::lock_guard<std::mutex> lock(ctx->mu);
    ctx->prof_on = true;
    ctx->prof_next = 0;
    ctx->prof_recs.clear();
    return PLK_OK;
}

plk_status plk_profile_stop(plk_context* ctx, plk_kernel_stat* stats, int cap, int* nstats) {
    if (!ctx) return fail(PLK_ERR_ARG, "null context");
    std::lock_guard<std::mutex> lock(ctx->mu);
    ctx->prof_on = false;
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    plk_kernel_stat agg[K_NKIND];
    std::memset(agg, 0, sizeof(agg));
    for (int k = 0; k < K_NKIND; k++) std::snprintf(agg[k].name, sizeof(agg[k].name), "%s", kProfNames[k]);
    for (const ProfRec& r : ctx->prof_recs) {
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, ctx->prof_ev[r.ev], ctx->prof_ev[r.ev + 1]));
        agg[r.kind].launches++;
        agg[r.kind].total_ms += ms;
        agg[r.kind].bytes += r.bytes;
        agg[r.kind].flops += r.flops;
    }
    int n = 0;
    for (int k = 0; k < K_NKIND; k++)
        if (agg[k].launches > 0) {
            if (stats && n < cap) stats[n] = agg[k];
            n++;
        }
    if (nstats) *nstats = n;
    ctx->prof_recs.clear();
    ctx->prof_next = 0;
    return PLK_OK;
}

plk_status plk_read_partial(plk_context* ctx, int id, int slot, double* partial, int32_t* scale) {
    Instance* in = get_instance(ctx, id);
    if (!in) return fail(PLK_ERR_ARG, "unknown instance");
    if (slot < 0 || slot >= in->d.nslots) return fail(PLK_ERR_ARG, "slot out of range");
    std::lock_guard<std::mutex> lock(ctx->mu);
    const int P = in->d.npatterns;
    const size_t ncnt = in->d.safe_numeric ? in->d.ncat : 1;
    if (partial) {
        const size_t cnt = (size_t)P * in->block;
        std::vector<unsigned char> h(cnt * in->elem);
        const unsigned char* src = reinterpret_cast<const unsigned char*>(in->partial) +
                                   (size_t)slot * in->Ppad * in->block * in->elem;
        HIP_TRY(hipMemcpyAsync(h.data(), src, h.size(), hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        for (size_t i = 0; i < cnt; i++)
            partial[i] = in->elem == 8 ? reinterpret_cast<const double*>(h.data())[i]
                                       : (double)reinterpret_cast<const float*>(h.data())[i];
    }
    if (scale) {
        const size_t cnt = (size_t)P * ncnt;
        std::vector<unsigned char> h(cnt * in->scale_elem);
        const unsigned char* src = reinterpret_cast<const unsigned char*>(in->scale) +
                                   (size_t)slot * in->Ppad * ncnt * in->scale_elem;
        HIP_TRY(hipMemcpyAsync(h.data(), src, h.size(), hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        for (size_t i = 0; i < cnt; i++)
            scale[i] = in->scale_elem == 2 ? (int32_t)reinterpret_cast<const uint16_t*>(h.data())[i]
                                           : reinterpret_cast<const int32_t*>(h.data())[i];
    }
    return PLK_OK;
}

}  // extern "C"
