// Batched likelihood over many instances (partitions of a PhyloSuperTree,
// tree/phylosupertree.cpp:701-735, or independent requests) in one
// persistent launch.
//
// Work list: every (instance, pattern tile) pair with real patterns, ordered
// by the host largest-cost first (computePartitionOrder,
// tree/phylosupertree.cpp:654-690). Workgroups take the next tile from an
// atomic counter; a tile runs its instance's whole op list (transition tables
// of all ops of all instances were built by one k_prep_batch launch) and the
// root lnL of its instance's branch, and stores its lnL sum in its own slot.
// k_batch_reduce then adds each instance's tile sums in tile order, so the
// result does not depend on which workgroup ran which tile.
#pragma once
#include "plk_mfma.h"

namespace plkdev {

struct BatchItem {
    InstDev I;
    ModelDev m;
    const plk_op* ops;     // this instance's ops inside the batch op array
    const void* tables;    // [nops][2][table_size] storage precision
    const double* root;    // val[C][n] then, for a tip dad, lut[SU+1][C][n]
    int nops;
    int table_size;
    int tile_begin;        // first global tile of this instance
    int ntiles;            // tiles with real patterns
    plk_end node, dad;
    int n, C, f32;
    int mode;              // BATCH_LNL: branch lnL; BATCH_THETA: theta buffer of the branch
};

enum { BATCH_LNL = 0, BATCH_THETA = 1 };

// Transition tables of a batch with k_prep_ops' arithmetic, and the root
// tables of k_prep_root. Workgroup (item, row) builds the tables of ops row,
// row + rows, ... of its item with the item's eigenvectors and tip vectors
// staged in LDS once; row 0 also writes the item's root tables and workgroup
// (0, 0) clears the work-list counters. LDS: evec n*n | tip SU*n | ex 2*C*n.
template <typename T>
__device__ inline void prep_child_table(const BatchItem& it, const plk_child ch, const double* ev_sh,
                                        const double* tp_sh, const double* ex, T* out) {
    const int n = it.n, C = it.C, SU = it.I.SU, nn = n * n;
    if (!ch.is_tip) {
        if (n == kProtN) {
            // the matrix-core tile's fragment order
            for (int t = threadIdx.x; t < C * kProtFrag; t += kThreads) {
                const int c = t / kProtFrag;
                int x, i;
                prot_frag_index(t - c * kProtFrag, x, i);
                out[t] = x < 0 ? (T)0 : (T)(ev_sh[x * n + i] * ex[c * n + i]);
            }
            return;
        }
        for (int t = threadIdx.x; t < C * nn; t += kThreads) {
            const int c = t / nn, r = t - c * nn, x = r / n, i = r - x * n;
            out[t] = (T)(ev_sh[x * n + i] * ex[c * n + i]);
        }
        return;
    }
    const int nv = C * n;
    for (int t = threadIdx.x; t < (SU + 1) * nv; t += kThreads) {
        const int s = t / nv, c = (t / n) % C, x = t % n;
        double v;
        if (s == SU) {
            v = 1.0;
        } else {
            const double* ev = ev_sh + x * n;
            const double* e = ex + c * n;
            const double* tp = tp_sh + s * n;
            v = ev[0] * e[0] * tp[0];
            for (int i = 1; i < n; i++) v += ev[i] * e[i] * tp[i];
        }
        out[t] = (T)v;
    }
}

__global__ void __launch_bounds__(kThreads)
k_prep_batch(const BatchItem* __restrict__ items, const double* __restrict__ lengths,
             double* __restrict__ roots, unsigned* __restrict__ counters, int ncounters) {
    extern __shared__ double sh[];
    if (blockIdx.x == 0 && blockIdx.y == 0)
        for (int t = threadIdx.x; t < ncounters; t += kThreads) counters[t] = 0u;
    const int i = blockIdx.x;
    const BatchItem& it = items[i];
    const int n = it.n, C = it.C, SU = it.I.SU, nn = n * n, nv = C * n;
    double* ev_sh = sh;
    double* tp_sh = ev_sh + nn;
    double* ex = tp_sh + SU * n;
    if ((int)blockIdx.y < it.nops) {
        for (int t = threadIdx.x; t < nn; t += kThreads) ev_sh[t] = it.m.evec[t];
        for (int t = threadIdx.x; t < SU * n; t += kThreads) tp_sh[t] = it.m.tip[t];
    }
    for (int o = blockIdx.y; o < it.nops; o += gridDim.y) {
        const plk_op op = it.ops[o];
        __syncthreads();   // staging visible, the previous op's ex consumed
        for (int t = threadIdx.x; t < 2 * nv; t += kThreads) {
            const int k = t / nv, c = (t / n) % C, q = t % n;
            ex[t] = exp(it.m.eval[q] * (it.m.rates[c] * op.child[k].length));
        }
        __syncthreads();
        for (int k = 0; k < 2; k++) {
            const size_t off = (size_t)(2 * o + k) * it.table_size;
            if (it.f32)
                prep_child_table<float>(it, op.child[k], ev_sh, tp_sh, ex + k * nv,
                                        reinterpret_cast<float*>(const_cast<void*>(it.tables)) + off);
            else
                prep_child_table<double>(it, op.child[k], ev_sh, tp_sh, ex + k * nv,
                                         reinterpret_cast<double*>(const_cast<void*>(it.tables)) + off);
        }
    }
    if (blockIdx.y != 0) return;
    double* out = const_cast<double*>(it.root);
    const double len = lengths[i];
    for (int t = threadIdx.x; t < nv; t += blockDim.x) {
        const int c = t / n, q = t % n;
        out[t] = exp(it.m.eval[q] * (it.m.rates[c] * len)) * it.m.props[c];
    }
    if (!it.dad.is_tip) return;
    for (int t = threadIdx.x; t < (SU + 1) * nv; t += blockDim.x) {
        const int q = t % n, c = (t / n) % C, s = t / nv;
        out[nv + t] = exp(it.m.eval[q] * (it.m.rates[c] * len)) * it.m.props[c] * it.m.tip[s * n + q];
    }
}

// Branch ends streamed one state at a time (20 states keep no per-thread
// state arrays), with the arithmetic of load_branch + the root / theta code
// of k_traverse and k_root_stream. MODE_BRANCH returns the pattern's
// category sum (every lane) and the scale minimum; MODE_THETA writes theta
// and buffer_scale (tip dads read tip_eig).
template <typename T, int N, int C>
__device__ inline double branch_end_stream(const InstDev& I, plk_end node, plk_end dad, int mode,
                                           const double* root_sh, int p, int c,
                                           const double* __restrict__ tip_eig, int* minsc_out = nullptr) {
    using ST = typename Prec<T>::scale_t;
    constexpr int block = N * C;
    const T* partial = reinterpret_cast<const T*>(I.partial);
    const ST* scale = reinterpret_cast<const ST*>(I.scale);
    const bool safe = I.safe != 0;
    const size_t sidx = safe ? (size_t)p * C + c : (size_t)p;
    const T* nv = partial + (size_t)node.id * I.slot_stride + (size_t)p * block + c * N;
    int sc = scale[(size_t)node.id * I.scale_stride + sidx];
    const T* dv = nullptr;
    int s = 0;
    if (dad.is_tip) {
        s = I.states[(size_t)dad.id * I.Ppad + p];
    } else {
        dv = partial + (size_t)dad.id * I.slot_stride + (size_t)p * block + c * N;
        sc += scale[(size_t)dad.id * I.scale_stride + sidx];
    }
    int minsc = sc;
    double adjust = 1.0;
    if (safe) {
        minsc = group_min_int<C>(sc);
        adjust = Prec<T>::safe_adjust(sc - minsc);
    }
    if (mode == MODE_THETA) {
        T* th = reinterpret_cast<T*>(I.theta) + (size_t)p * block + c * N;
        if (dad.is_tip) {
            const double* tp = tip_eig + s * N;
#pragma unroll 4
            for (int i = 0; i < N; i++) th[i] = (T)((double)nv[i] * tp[i] * adjust);
        } else {
#pragma unroll 4
            for (int i = 0; i < N; i++) th[i] = (T)((double)nv[i] * (double)dv[i] * adjust);
        }
        if (c == 0) I.buffer_scale[p] = (double)minsc * Prec<T>::log_unit;
        return 0.0;
    }
    double lhc;
    if (dad.is_tip) {
        const double* lut = root_sh + block + (s * C + c) * N;
        lhc = lut[0] * (double)nv[0];
#pragma unroll 1
        for (int i = 1; i < N; i++) lhc = fma(lut[i], (double)nv[i], lhc);
    } else {
        const double* val = root_sh + c * N;
        lhc = val[0] * (double)dv[0] * (double)nv[0];
#pragma unroll 1
        for (int i = 1; i < N; i++) lhc = fma(val[i] * (double)dv[i], (double)nv[i], lhc);
    }
    if (minsc_out) *minsc_out = minsc;
    return safe ? group_sum_sumvec<C>(lhc * adjust) : group_sum_ordered<C>(lhc);
}

// 16-pattern MFMA sub-tiles of a protein batch tile
constexpr int kBatchSub = 4;

// Dynamic LDS bytes batch_tile uses for one item shape: U^-1, the two child
// tables of a step and the root tables (which reuse the table area), or for
// 20 states the maxima of prot_ops_tile (2 x kBatchSub x C x 16) and the
// root tables.
__host__ __device__ inline size_t batch_smem_bytes(int n, int C, int SU, int table_size, size_t elem) {
    const size_t root = (size_t)(SU + 2) * n * C * 8;
    if (n == kCodonN) return (size_t)2 * n * kCodonStride * elem;   // E | U^-1; root rows stay in HBM
    if (n == kProtN) {
        const size_t mf = (size_t)2 * kBatchSub * C * 16 * elem;
        const size_t rt = (size_t)n * n * elem + root;
        return mf > rt ? mf : rt;
    }
    const size_t tab = 2 * (size_t)table_size * elem;
    return (size_t)n * n * elem + (tab > root ? tab : root);
}

// patterns per thread of a 4-state batch tile: two, so that each step's
// table staging and barriers serve twice the data and two patterns' loads
// are in flight together
constexpr int kDnaPatternsPerThread = 2;

// patterns per work-list tile: 4 states, kDnaPatternsPerThread patterns per
// C threads (128 with four categories); 20 and 61 states, one pattern per C
// threads (four 16-pattern protein matrix-core tiles, or one codon tile)
__host__ __device__ constexpr int batch_tile_patterns(int n, int C) {
    return (n == 4 ? kDnaPatternsPerThread : 1) * kThreads / C;
}

// all ops of one tile of one item, then its root lnL; returns the tile's
// lnL sum (thread 0)
template <typename T, int N, int C>
__device__ __forceinline__ double batch_tile(const BatchItem& it, int tile, unsigned char* smem) {
    constexpr int PB = batch_tile_patterns(N, C);
    constexpr int kPP = N == 4 ? kDnaPatternsPerThread : 1;   // patterns per thread
    constexpr int block = N * C;
    T* inv_sh = reinterpret_cast<T*>(smem);
    T* tab_sh = inv_sh + N * N;
    const InstDev& I = it.I;
    const int c = threadIdx.x % C;
    const int p = tile * PB + (int)threadIdx.x / C;
    const T* tables = reinterpret_cast<const T*>(it.tables);
    __syncthreads();
    if constexpr (N == kCodonN) {
        // codon: the lone kernel's matrix-core tile, one pass per rate
        // category (each category's scaling is its own)
        constexpr int S = kCodonStride;
        static_assert(PB == kCodonTile, "batch tile must be one codon tile");
        T* e_m = reinterpret_cast<T*>(smem);
        T* inv_m = e_m + N * S;
        stage_lds<N * N>(inv_m, it.m.inv_evec, N * N, [](int t) { return (t / N) * S + t % N; });
        __syncthreads();
        for (int cc = 0; cc < C; cc++)
            codon_ops_tile<T>(I, it.ops, it.nops, tables, it.table_size, cc, tile * PB, e_m, inv_m);
    } else if constexpr (N == 20) {
        // protein: the matrix-core tile of the lone traversal kernel, every
        // wave one category of four 16-pattern sub-tiles
        static_assert(PB == kBatchSub * kMfmaTile, "batch tile must be kBatchSub MFMA tiles");
        prot_ops_tile<T, C, kBatchSub>(I, it.m.inv_evec, it.ops, it.nops, tables, it.table_size,
                                       reinterpret_cast<T*>(smem), tile * PB);
    } else {
    for (int t = threadIdx.x; t < N * N; t += kThreads) inv_sh[t] = (T)it.m.inv_evec[t];
    for (int o = 0; o < it.nops; o++) {
        const plk_op op = it.ops[o];
        int left = 0, right = 1;
        if (!op.child[0].is_tip && op.child[1].is_tip) { left = 1; right = 0; }
        const plk_child L = op.child[left], R = op.child[right];
        const int szL = L.is_tip ? (I.SU + 1) * block : block * N;
        const int szR = R.is_tip ? (I.SU + 1) * block : block * N;
        __syncthreads();
        const T* gL = tables + (size_t)(2 * o + left) * it.table_size;
        const T* gR = tables + (size_t)(2 * o + right) * it.table_size;
        for (int t = threadIdx.x; t < szL; t += kThreads) tab_sh[t] = gL[t];
        for (int t = threadIdx.x; t < szR; t += kThreads) tab_sh[it.table_size + t] = gR[t];
        __syncthreads();
        // the barrier above also orders the previous step's scale counters
        // (written by the category-0 lane) before these loads
        OpIn<T, N> in[kPP];
#pragma unroll
        for (int j = 0; j < kPP; j++)
            op_load<T, N, C>(I, L.is_tip, L.id, R.is_tip, R.id, p + j * (kThreads / C), c, in[j]);
#pragma unroll
        for (int j = 0; j < kPP; j++)
            op_finish<T, N, C>(I, inv_sh, tab_sh, tab_sh + it.table_size, L.is_tip, R.is_tip, op.dad_slot,
                               p + j * (kThreads / C), c, in[j]);
    }
    }
    double* root_sh = reinterpret_cast<double*>(tab_sh);
    const plk_end node = it.node, dad = it.dad;
    __syncthreads();
    if (it.mode == BATCH_THETA) {
        // computeLikelihoodBufferSIMD's theta: tip dads read the tip vectors
        // (the streamed form reads them from HBM)
        if constexpr (N == 4) {
            if (dad.is_tip)
                for (int t = threadIdx.x; t < (I.SU + 1) * N; t += kThreads) root_sh[t] = it.m.tip[t];
            __syncthreads();
#pragma unroll
            for (int j = 0; j < kPP; j++) theta_apply<T, N, C>(I, node, dad, root_sh, p + j * (kThreads / C), c);
        } else {
            branch_end_stream<T, N, C>(I, node, dad, MODE_THETA, root_sh, p, c, it.m.tip);
        }
        return 0.0;
    }
    if constexpr (N == kCodonN) {
        // 61-state lookup rows (121 KB for a tip dad) stay in HBM, as k_root_stream reads them
        root_sh = const_cast<double*>(it.root);
    } else {
        const int root_size = block + (dad.is_tip ? (I.SU + 1) * block : 0);
        for (int t = threadIdx.x; t < root_size; t += kThreads) root_sh[t] = it.root[t];
        __syncthreads();
    }
    double contrib = 0.0;
#pragma unroll
    for (int j = 0; j < kPP; j++) {
        const int pj = p + j * (kThreads / C);
        double lh0;
        int minsc;
        if constexpr (N != 4) {
            // streamed (no per-thread state arrays), same arithmetic as below
            lh0 = branch_end_stream<T, N, C>(I, node, dad, MODE_BRANCH, root_sh, pj, c, it.m.tip, &minsc);
        } else {
            double nv[N], dv[N], adjust;
            load_branch<T, N, C>(I, node, dad, nullptr, pj, c, nv, dv, minsc, adjust);
            double lhc;
            if (dad.is_tip) {
                const int s = I.states[(size_t)dad.id * I.Ppad + pj];
                const double* lut = root_sh + block + (s * C + c) * N;
                lhc = lut[0] * nv[0];
#pragma unroll
                for (int i = 1; i < N; i++) lhc = fma(lut[i], nv[i], lhc);
            } else {
                const double* val = root_sh + c * N;
                lhc = val[0] * dv[0] * nv[0];
#pragma unroll
                for (int i = 1; i < N; i++) lhc = fma(val[i] * dv[i], nv[i], lhc);
            }
            lh0 = I.safe ? group_sum_sumvec<C>(lhc * adjust) : group_sum_ordered<C>(lhc);
        }
        if (c == 0) {
            const double lh = fabs(lh0) + I.invar[pj];
            const double lnp = log(lh) + (double)minsc * Prec<T>::log_unit;
            if (pj < I.P) {
                I.ptn_lh[pj] = lnp;
                contrib += lnp * I.freq[pj];
            }
        }
    }
    double v[1] = {contrib};
    block_sum<1>(v, reinterpret_cast<double*>(smem));
    return v[0];
}

// Instance classes a batch kernel is built for (bit mask). BC_ALL is the
// fused DNA + protein launch; codon tiles (63 KB of LDS) always run in a
// launch of their own class, so they do not lower the others' occupancy.
enum BatchClass {
    BC_DNA64 = 1, BC_PROT64 = 2, BC_DNA32 = 4, BC_PROT32 = 8, BC_ALL = 15,
    BC_COD64 = 16, BC_COD32 = 32
};
constexpr int kBatchClasses = 6;

__host__ __device__ inline int batch_class(int n, int f32) {
    if (n == 4) return f32 ? BC_DNA32 : BC_DNA64;
    if (n == 20) return f32 ? BC_PROT32 : BC_PROT64;
    return f32 ? BC_COD32 : BC_COD64;
}

// The tile list of one launch in kXcds segments, one per XCD of the device:
// the host places all of an item's tiles in one segment (its transition
// tables then sit in one XCD's L2 instead of every XCD fetching them),
// items bin-packed by cost. Segment q holds tiles [b[q], b[q + 1]).
constexpr int kXcds = 8;
struct XcdSplit {
    int b[kXcds + 1];
};

// Persistent workgroups over the tiles of `xs`. A workgroup takes tiles from
// its own XCD's segment (workgroups are dispatched to the XCDs round-robin,
// so blockIdx.x % kXcds), then from the others' in turn until every segment
// is drained; every workgroup ends after trying each segment's counter. MASK
// names the instance classes the launch may meet; a tile's code path is
// uniform across its workgroup.
// The protein classes are held to 3 (FP64, <= 168 registers) and 4 (FP32,
// <= 128) waves per SIMD, which they fit without spilling; the compiler's
// own choice left each one wave lower.
template <int MASK>
__global__ void __launch_bounds__(kThreads)
__attribute__((amdgpu_waves_per_eu(MASK == BC_PROT64 ? 3 : (MASK == BC_PROT32 ? 4 : 1))))
k_batch(const BatchItem* __restrict__ items, const int* __restrict__ tile_item, XcdSplit xs,
        unsigned* __restrict__ counters, double* __restrict__ tile_sums) {
    extern __shared__ __align__(16) unsigned char smem[];
    // thread 0's scheduling state lives in LDS, so no register stays live
    // across the tiles
    __shared__ int next_sh, tried_sh, bound_sh[kXcds + 1];
    if (threadIdx.x <= kXcds) bound_sh[threadIdx.x] = xs.b[threadIdx.x];
    if (threadIdx.x == 0) tried_sh = 0;   // segments found empty
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) {
            const int home = blockIdx.x % kXcds;
            int g = -1, tried = tried_sh;
            while (tried < kXcds) {
                const int q = (home + tried) % kXcds;
                const int t = bound_sh[q] + (int)atomicAdd(&counters[q], 1u);
                if (t < bound_sh[q + 1]) {
                    g = t;
                    break;
                }
                tried++;
            }
            tried_sh = tried;
            next_sh = g;
        }
        __syncthreads();
        // wave-uniform by construction; readfirstlane lets the item's fields
        // live in scalar registers
        const int g = __builtin_amdgcn_readfirstlane(next_sh);
        if (g < 0) break;
        const int i = __builtin_amdgcn_readfirstlane(tile_item[g]);
        const BatchItem& it = items[i];
        const int tile = g - it.tile_begin;
        const int cls = batch_class(it.n, it.f32);
        double s = 0.0;
        if ((MASK & BC_DNA64) && cls == BC_DNA64) s = batch_tile<double, 4, 4>(it, tile, smem);
        else if ((MASK & BC_PROT64) && cls == BC_PROT64) s = batch_tile<double, 20, 4>(it, tile, smem);
        else if ((MASK & BC_DNA32) && cls == BC_DNA32) s = batch_tile<float, 4, 4>(it, tile, smem);
        else if ((MASK & BC_PROT32) && cls == BC_PROT32) s = batch_tile<float, 20, 4>(it, tile, smem);
        else if ((MASK & BC_COD64) && cls == BC_COD64) s = batch_tile<double, kCodonN, 4>(it, tile, smem);
        else if ((MASK & BC_COD32) && cls == BC_COD32) s = batch_tile<float, kCodonN, 4>(it, tile, smem);
        if (threadIdx.x == 0) tile_sums[g] = s;
    }
}

// lnl[i] = sum of item i's tile sums in tile order (one thread per item)
__global__ void __launch_bounds__(kThreads)
k_batch_reduce(const BatchItem* __restrict__ items, int nitems, const double* __restrict__ tile_sums,
               double* __restrict__ lnl) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nitems) return;
    double s = 0.0;
    for (int t = 0; t < items[i].ntiles; t++) s += tile_sums[items[i].tile_begin + t];
    lnl[i] = s;
}

// ---------------------------------------------------------------------------
// Batched theta-buffer evaluations: one item = (instance, branch length), the
// derivative and lnL sums of computeLikelihoodDervSIMD for each. Every item
// is split into the same pattern tiles as a lone plk_likelihood_derv launch
// and its tile sums are reduced with the same fixed-order tree, so an item's
// result is bit-identical to the single-instance call whatever else shares
// the batch. One launch per (precision, states, categories) shape present.
struct DervItem {
    InstDev I;
    ModelDev m;
    double length;
    int tile_begin;   // first global tile
    int ntiles;       // Ppad / (kThreads / C)
};

template <typename T, int N, int C>
__global__ void __launch_bounds__(kThreads)
k_derv_batch(const DervItem* __restrict__ items, const int* __restrict__ tile_item, int tile_begin,
             double* __restrict__ tile_sums) {
    __shared__ double val[3 * N * C];
    __shared__ double red[3 * (kThreads / 64)];
    const int g = tile_begin + blockIdx.x;
    const DervItem& it = items[tile_item[g]];
    buffer_block<T, N, C>(it.I, it.m, it.length, BUF_DERV, g - it.tile_begin, val, red,
                          tile_sums + 3 * (size_t)g);
}

// out[3i..3i+2] = df, ddf, lnL of item i (one workgroup per item)
__global__ void __launch_bounds__(kThreads)
k_derv_reduce(const DervItem* __restrict__ items, const double* __restrict__ tile_sums,
              double* __restrict__ out) {
    __shared__ double sh[kThreads];
    const DervItem& it = items[blockIdx.x];
    reduce_block(tile_sums + 3 * (size_t)it.tile_begin, it.ntiles, 3, out + 3 * (size_t)blockIdx.x, sh);
}

}  // namespace plkdev
