// Fitch parsimony (tree/phylotreepars.cpp:32-300, computePartialParsimonyFast
// and computeParsimonyBranchFast; the SIMD form tree/phylokernel.h:1465-2094).
//
// Fitch's step at a node whose two children hold state sets X and Y:
//   Z = X & Y;  where Z is empty, Z = X | Y and the site costs one change.
// The score of a branch is the changes below both ends plus the sites whose
// two end sets are disjoint.
//
// DNA and protein keep the reference's bit-plane layout: per node, one 32-bit
// word per state for every 32 patterns (bit b of plane s of word w: pattern
// 32w+b admits state s), the n planes of a word adjacent ([slot][word][n], as
// the reference's partial_pars). The reference expands each pattern into
// `frequency` sites so that a popcount gives the score; here patterns stay
// compressed and the empty-intersection mask is weighted by the pattern
// frequencies, so the per-node cache is a weighted change count per word
// ([slot][word], double) in place of the reference's one subtree score per
// node. Both give the same integer score for integer frequencies. Tip planes
// are built once per pattern set (the reference's computeTipPartialParsimony,
// tree/phylotreepars.cpp:561-633).
//
// One thread per word walking the whole op list would chain one memory round
// trip per op; instead the host groups the ops into levels (an op's level is
// one more than its children's), and a workgroup owns kParsWords words and
// runs each level's ops side by side (kThreads / kParsWords ops at a time),
// with a barrier between levels. A wave then reads 8 ops x 8 adjacent words.
//
// Codons (61 states) keep one 64-bit set per pattern: 61 planes per 32
// patterns hold the same 61 bits, so the layout costs nothing there, and a
// thread per pattern keeps a wave's loads contiguous.
#pragma once
#include "plk_kernels.h"

namespace plkdev {

struct ParsDev {
    int P, Ppad, nwords, n, SU, seq;   // seq: 0 DNA, 1 protein, 2 codon
    const uint8_t* states;             // [ntaxa][Ppad]
    const double* freq;                // [Ppad], 0 on padding
    uint32_t* tip_planes;              // [ntaxa][nwords][n] (DNA, protein)
    uint32_t* planes;                  // [nslots][nwords][n] (DNA, protein)
    uint64_t* sets;                    // [nslots][Ppad] (codon)
    double* counts;                    // [nslots][nwords] weighted changes in the subtree (DNA, protein)
    uint16_t* changes;                 // [nslots][Ppad] changes in the subtree (codon)
};

// the Fitch set of an observed tip state: the state, the ambiguity set, or
// every state for STATE_UNKNOWN (tree/phylotreepars.cpp:60-160)
__device__ inline uint64_t tip_set(int n, int SU, int seq, int s) {
    const uint64_t all = n >= 64 ? ~0ull : ((1ull << n) - 1);
    if (s < n) return 1ull << s;
    if (s == SU) return all;
    if (seq == 0) return (uint64_t)(s - 3) & all;   // DNA ambiguity: bit mask + 3
    if (seq == 1) {                                  // protein B = N|D, Z = Q|E, J = I|L
        const int lo = s == 20 ? 2 : (s == 21 ? 5 : 9);
        return (1ull << lo) | (1ull << (lo + 1));
    }
    return all;
}

// weighted count of the patterns of word w whose bit is set in e
__device__ inline double weigh(uint32_t e, const double* __restrict__ freq, int w) {
    double s = 0.0;
    while (e) {
        s += freq[32 * w + __ffs(e) - 1];
        e &= e - 1;
    }
    return s;
}

constexpr int kParsWords = 8;   // words per workgroup in k_pars_levels

// N planes of one word as 16-byte vectors (N = 4: one, N = 20: five)
template <int N>
struct Planes {
    uint4 v[N / 4];
};

template <int N>
__device__ inline Planes<N> load_planes(const uint32_t* p) {
    Planes<N> x;
#pragma unroll
    for (int k = 0; k < N / 4; k++) x.v[k] = reinterpret_cast<const uint4*>(p)[k];
    return x;
}

template <int N>
__device__ inline void store_planes(uint32_t* p, const Planes<N>& x) {
#pragma unroll
    for (int k = 0; k < N / 4; k++) reinterpret_cast<uint4*>(p)[k] = x.v[k];
}

// bits of the patterns whose two sets are disjoint
template <int N>
__device__ inline uint32_t disjoint(const Planes<N>& x, const Planes<N>& y) {
    uint32_t any = 0u;
#pragma unroll
    for (int k = 0; k < N / 4; k++)
        any |= (x.v[k].x & y.v[k].x) | (x.v[k].y & y.v[k].y) | (x.v[k].z & y.v[k].z) | (x.v[k].w & y.v[k].w);
    return ~any;
}

// Tip planes: one thread per (tip, word) gathers its 32 patterns' sets.
template <int N>
__global__ void __launch_bounds__(kThreads) k_pars_tip_planes(ParsDev D, int ntaxa) {
    const size_t t = (size_t)blockIdx.x * kThreads + threadIdx.x;
    if (t >= (size_t)ntaxa * D.nwords) return;
    const int tip = (int)(t / D.nwords), w = (int)(t % D.nwords);
    uint32_t pl[N];
#pragma unroll
    for (int s = 0; s < N; s++) pl[s] = 0u;
    const uint8_t* st = D.states + (size_t)tip * D.Ppad + 32 * w;
    for (int b = 0; b < 32; b++) {
        const uint32_t set = (uint32_t)tip_set(N, D.SU, D.seq, st[b]);
#pragma unroll
        for (int s = 0; s < N; s++) pl[s] |= ((set >> s) & 1u) << b;
    }
    uint32_t* out = D.tip_planes + t * N;
#pragma unroll
    for (int s = 0; s < N; s++) out[s] = pl[s];
}

// The ops level by level, then the branch (node, dad). level_start[l] is the
// first op of level l (ops sorted by level on the host). block_sums[2b]: the
// tree score, [2b+1]: the changes on the branch itself.
template <int N>
__global__ void __launch_bounds__(kThreads)
k_pars_levels(ParsDev D, const plk_op* __restrict__ ops, const int* __restrict__ level_start, int nlevels,
              plk_end node, plk_end dad, double* __restrict__ block_sums) {
    __shared__ double red[2 * kThreads / 64];
    constexpr int kLanes = kThreads / kParsWords;   // ops side by side
    const int w = blockIdx.x * kParsWords + (int)threadIdx.x % kParsWords;
    const int lane = (int)threadIdx.x / kParsWords;
    const bool live = w < D.nwords;
    const size_t nw = D.nwords;
    auto end_planes = [&](bool is_tip, int id) {
        return load_planes<N>((is_tip ? D.tip_planes : D.planes) + ((size_t)id * nw + w) * N);
    };
    auto end_count = [&](bool is_tip, int id) { return is_tip ? 0.0 : D.counts[(size_t)id * nw + w]; };
    for (int l = 0; l < nlevels; l++) {
        if (live)
            for (int o = level_start[l] + lane; o < level_start[l + 1]; o += kLanes) {
                const plk_op op = ops[o];
                const Planes<N> x = end_planes(op.child[0].is_tip, op.child[0].id);
                const Planes<N> y = end_planes(op.child[1].is_tip, op.child[1].id);
                const double c = end_count(op.child[0].is_tip, op.child[0].id) +
                                 end_count(op.child[1].is_tip, op.child[1].id);
                const uint32_t e = disjoint<N>(x, y);
                Planes<N> z;
#pragma unroll
                for (int k = 0; k < N / 4; k++) {
                    const uint4 a = x.v[k], b = y.v[k];
                    z.v[k].x = (a.x & b.x) | (e & (a.x | b.x));
                    z.v[k].y = (a.y & b.y) | (e & (a.y | b.y));
                    z.v[k].z = (a.z & b.z) | (e & (a.z | b.z));
                    z.v[k].w = (a.w & b.w) | (e & (a.w | b.w));
                }
                store_planes<N>(D.planes + ((size_t)op.dad_slot * nw + w) * N, z);
                D.counts[(size_t)op.dad_slot * nw + w] = c + weigh(e, D.freq, w);
            }
        __syncthreads();   // this level's sets visible to the next
    }
    double score = 0.0, bsub = 0.0;
    if (live && lane == 0) {
        bsub = weigh(disjoint<N>(end_planes(node.is_tip, node.id), end_planes(dad.is_tip, dad.id)), D.freq, w);
        score = end_count(node.is_tip, node.id) + end_count(dad.is_tip, dad.id) + bsub;
    }
    double v[2] = {score, bsub};
    block_sum<2, false>(v, red);
    if (threadIdx.x == 0) {
        block_sums[2 * blockIdx.x] = v[0];
        block_sums[2 * blockIdx.x + 1] = v[1];
    }
}

// The same ops with the workgroup's working set in LDS. The host numbers
// every branch end the call touches: inputs first (tips, and internal slots
// computed by an earlier call), loaded once, then the slots the ops write.
// LDS holds their planes ([end][W][N]), counts ([end][W]) and the W words'
// pattern frequencies; a level reads and writes LDS only (results are also
// written back to HBM), so a level costs an LDS round trip and a barrier
// instead of an L2 round trip. lops[o] = (LDS end of child 0, of child 1,
// LDS end of the dad, the dad's slot).
template <int N, int W>
__global__ void __launch_bounds__(kThreads)
k_pars_lds(ParsDev D, const int4* __restrict__ lops, const int* __restrict__ level_start, int nlevels,
           const plk_end* __restrict__ inputs, int nin, int nends, int bnode, int bdad,
           double* __restrict__ block_sums) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ double red[2 * kThreads / 64];
    uint32_t* pl = reinterpret_cast<uint32_t*>(smem);                      // [nends][W][N]
    double* cnt = reinterpret_cast<double*>(smem + (size_t)nends * W * N * 4);   // [nends][W]
    double* fq = cnt + (size_t)nends * W;                                  // [W * 32]
    const size_t nw = D.nwords;
    const int w0 = blockIdx.x * W;
    for (int t = threadIdx.x; t < nin * W * N; t += kThreads) {
        const int e = t / (W * N), r = t - e * (W * N), word = r / N, st = r - word * N;
        const plk_end in = inputs[e];
        const uint32_t* src = in.is_tip ? D.tip_planes : D.planes;
        pl[t] = src[((size_t)in.id * nw + w0 + word) * N + st];
    }
    for (int t = threadIdx.x; t < nin * W; t += kThreads) {
        const plk_end in = inputs[t / W];
        cnt[t] = in.is_tip ? 0.0 : D.counts[(size_t)in.id * nw + w0 + t % W];
    }
    for (int t = threadIdx.x; t < W * 32; t += kThreads) fq[t] = D.freq[32 * w0 + t];
    __syncthreads();
    constexpr int kLanes = kThreads / W;
    const int word = (int)threadIdx.x % W, lane = (int)threadIdx.x / W;
    auto planes_of = [&](int e) { return load_planes<N>(pl + ((size_t)e * W + word) * N); };
    auto weigh_lds = [&](uint32_t e) {
        double s = 0.0;
        while (e) {
            s += fq[32 * word + __ffs(e) - 1];
            e &= e - 1;
        }
        return s;
    };
    for (int l = 0; l < nlevels; l++) {
        for (int o = level_start[l] + lane; o < level_start[l + 1]; o += kLanes) {
            const int4 op = lops[o];
            const Planes<N> x = planes_of(op.x), y = planes_of(op.y);
            const uint32_t e = disjoint<N>(x, y);
            Planes<N> z;
#pragma unroll
            for (int k = 0; k < N / 4; k++) {
                const uint4 a = x.v[k], b = y.v[k];
                z.v[k].x = (a.x & b.x) | (e & (a.x | b.x));
                z.v[k].y = (a.y & b.y) | (e & (a.y | b.y));
                z.v[k].z = (a.z & b.z) | (e & (a.z | b.z));
                z.v[k].w = (a.w & b.w) | (e & (a.w | b.w));
            }
            const double c = cnt[op.x * W + word] + cnt[op.y * W + word] + weigh_lds(e);
            store_planes<N>(pl + ((size_t)op.z * W + word) * N, z);
            cnt[op.z * W + word] = c;
            store_planes<N>(D.planes + ((size_t)op.w * nw + w0 + word) * N, z);
            D.counts[(size_t)op.w * nw + w0 + word] = c;
        }
        __syncthreads();   // this level's sets visible to the next
    }
    double score = 0.0, bsub = 0.0;
    if (lane == 0) {
        bsub = weigh_lds(disjoint<N>(planes_of(bnode), planes_of(bdad)));
        score = cnt[bnode * W + word] + cnt[bdad * W + word] + bsub;
    }
    double v[2] = {score, bsub};
    block_sum<2, false>(v, red);
    if (threadIdx.x == 0) {
        block_sums[2 * blockIdx.x] = v[0];
        block_sums[2 * blockIdx.x + 1] = v[1];
    }
}

// LDS bytes of k_pars_lds
__host__ __device__ inline size_t pars_lds_bytes(int n, int W, int nends) {
    return (size_t)nends * W * n * 4 + (size_t)nends * W * 8 + (size_t)W * 32 * 8;
}

// Codons: one thread per pattern, one 64-bit set and the number of changes
// in the subtree per pattern, weighted by the frequency at the branch.
__device__ inline void pars_end64(const ParsDev& D, plk_end e, int p, uint64_t& set, int& c) {
    if (e.is_tip) {
        set = tip_set(D.n, D.SU, D.seq, D.states[(size_t)e.id * D.Ppad + p]);
        c = 0;
    } else {
        set = D.sets[(size_t)e.id * D.Ppad + p];
        c = D.changes[(size_t)e.id * D.Ppad + p];
    }
}

__global__ void __launch_bounds__(kThreads)
k_pars_sets(ParsDev D, const plk_op* __restrict__ ops, int nops, plk_end node, plk_end dad,
            double* __restrict__ block_sums) {
    __shared__ double red[2 * kThreads / 64];
    const int p = blockIdx.x * kThreads + threadIdx.x;
    double score = 0.0, bsub = 0.0;
    if (p < D.Ppad) {
        const double f = D.freq[p];
        for (int o = 0; o < nops; o++) {
            const plk_op op = ops[o];
            uint64_t a, b;
            int ca, cb;
            pars_end64(D, plk_end{op.child[0].is_tip, op.child[0].id}, p, a, ca);
            pars_end64(D, plk_end{op.child[1].is_tip, op.child[1].id}, p, b, cb);
            const uint64_t both = a & b;
            D.sets[(size_t)op.dad_slot * D.Ppad + p] = both ? both : (a | b);
            D.changes[(size_t)op.dad_slot * D.Ppad + p] = (uint16_t)(ca + cb + (both ? 0 : 1));
        }
        uint64_t a, b;
        int ca, cb;
        pars_end64(D, node, p, a, ca);
        pars_end64(D, dad, p, b, cb);
        bsub = (a & b) ? 0.0 : f;
        score = (double)(ca + cb) * f + bsub;
    }
    double v[2] = {score, bsub};
    block_sum<2>(v, red);
    if (threadIdx.x == 0) {
        block_sums[2 * blockIdx.x] = v[0];
        block_sums[2 * blockIdx.x + 1] = v[1];
    }
}

}  // namespace plkdev

namespace plkdev {

constexpr int kSankoffMaxSU = 61;   // tip-cost rows in LDS: DNA 18, protein 23, codon 61

// Sankoff parsimony (tree/phylotreepars.cpp:636-885): one thread per pattern
// runs the whole op list (a pattern's costs depend on its own children only),
// costs pattern-major [slot][Ppad][N] in HBM, the cost matrix and tip costs
// in LDS; then the branch: min over the node's states of the two ends' costs
// (through the cost matrix for an internal dad) and the cost on the branch at
// the first minimum, both weighted by the pattern frequency. Block sums
// [grid][2].
template <int N>
__global__ void __launch_bounds__(kThreads)
k_sankoff(const plk_op* __restrict__ ops, int nops, const uint32_t* __restrict__ cost_g,
          const uint32_t* __restrict__ tip_g, int SU, const uint8_t* __restrict__ states, int P, int Ppad,
          uint32_t* __restrict__ part, plk_end node, plk_end dad, const double* __restrict__ freq,
          double* __restrict__ block_sums) {
    __shared__ uint32_t cost[N * N];
    __shared__ uint32_t tip[(kSankoffMaxSU + 1) * N];
    constexpr int kUnrollI = N <= 20 ? N : 1, kUnrollJ = N <= 20 ? N : 4;   // codons: no full unroll
    __shared__ double sh[2 * (kThreads / 64)];
    for (int t = threadIdx.x; t < N * N; t += kThreads) cost[t] = cost_g[t];
    for (int t = threadIdx.x; t < (SU + 1) * N; t += kThreads) tip[t] = tip_g[t];
    __syncthreads();
    const int p = blockIdx.x * kThreads + threadIdx.x;
    double acc[2] = {0.0, 0.0};
    // the last KC dad slots' costs stay in registers: in post-order a child
    // is usually one of the latest dads, whose store would otherwise be read
    // back through the memory hierarchy before the op can start
    constexpr int KC = N <= 4 ? 4 : (N <= 20 ? 2 : 1);
    uint32_t cache[KC][N];
    int cslot[KC];
#pragma unroll
    for (int k = 0; k < KC; k++) cslot[k] = -1;
    int cnext = 0;
    if (p < P) {
        for (int o = 0; o < nops; o++) {
            const plk_op op = ops[o];
            int left = (!op.child[0].is_tip && op.child[1].is_tip) ? 1 : 0;
            const plk_child L = op.child[left], R = op.child[1 - left];
            uint32_t out[N];
#pragma unroll
            for (int i = 0; i < N; i++) out[i] = 0;
#pragma unroll
            for (int k = 0; k < 2; k++) {
                const plk_child ch = k == 0 ? L : R;
                if (ch.is_tip) {
                    const uint32_t* tp = tip + (int)states[(size_t)ch.id * Ppad + p] * N;
#pragma unroll
                    for (int i = 0; i < N; i++) out[i] += tp[i];
                } else {
                    uint32_t cv[N];
                    bool hit = false;
#pragma unroll
                    for (int kk = 0; kk < KC; kk++)
                        if (cslot[kk] == ch.id) {
                            hit = true;
#pragma unroll
                            for (int j = 0; j < N; j++) cv[j] = cache[kk][j];
                        }
                    if (!hit) {
                        const uint32_t* cp = part + ((size_t)ch.id * Ppad + p) * N;
#pragma unroll
                        for (int j = 0; j < N; j++) cv[j] = cp[j];
                    }
#pragma unroll kUnrollI
                    for (int i = 0; i < N; i++) {
                        uint32_t m = cv[0] + cost[i * N];
#pragma unroll kUnrollJ
                        for (int j = 1; j < N; j++) m = min(m, cv[j] + cost[i * N + j]);
                        out[i] += m;
                    }
                }
            }
            uint32_t* dp = part + ((size_t)op.dad_slot * Ppad + p) * N;
#pragma unroll
            for (int i = 0; i < N; i++) dp[i] = out[i];
            // a slot written again replaces its cached copy
#pragma unroll
            for (int kk = 0; kk < KC; kk++)
                if (cslot[kk] == op.dad_slot) cslot[kk] = -1;
#pragma unroll
            for (int kk = 0; kk < KC; kk++)
                if (kk == cnext) {
                    cslot[kk] = op.dad_slot;
#pragma unroll
                    for (int i = 0; i < N; i++) cache[kk][i] = out[i];
                }
            cnext = cnext + 1 == KC ? 0 : cnext + 1;
        }
        const uint32_t* nb = part + ((size_t)node.id * Ppad + p) * N;
        uint32_t best, br;
        if (dad.is_tip) {
            const uint32_t* tp = tip + (int)states[(size_t)dad.id * Ppad + p] * N;
            best = tp[0] + nb[0];
            br = tp[0];
            for (int i = 1; i < N; i++)
                if (tp[i] + nb[i] < best) {
                    best = tp[i] + nb[i];
                    br = tp[i];
                }
        } else {
            const uint32_t* db = part + ((size_t)dad.id * Ppad + p) * N;
            uint32_t dv[N];
#pragma unroll
            for (int j = 0; j < N; j++) dv[j] = db[j];
            best = 0xFFFFFFFFu;
            br = 0xFFFFFFFFu;
            for (int i = 0; i < N; i++) {
                uint32_t m = dv[0] + cost[i * N], bs = cost[i * N];
#pragma unroll
                for (int j = 1; j < N; j++)
                    if (dv[j] + cost[i * N + j] < m) {
                        m = dv[j] + cost[i * N + j];
                        bs = cost[i * N + j];
                    }
                m += nb[i];
                if (m < best) {
                    best = m;
                    br = bs;
                }
            }
        }
        acc[0] = (double)best * freq[p];
        acc[1] = (double)br * freq[p];
    }
    block_sum<2>(acc, sh);
    if (threadIdx.x == 0) {
        block_sums[(size_t)blockIdx.x * 3] = acc[0];
        block_sums[(size_t)blockIdx.x * 3 + 1] = acc[1];
        block_sums[(size_t)blockIdx.x * 3 + 2] = 0.0;
    }
}

}  // namespace plkdev
