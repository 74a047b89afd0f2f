// Partial-likelihood updates for small state spaces (protein, 20 states) on
// the matrix cores: tree/phylokernelnew.h:1632-2031 as two contractions per
// category and traversal step,
//   phase A  tmp[x][p] = prod_children sum_k E_child[c][x][k] * child[p][c][k]
//            (tip children: the lookup leaf[s_p][c][x])
//   phase B  dad[p][i] = sum_x tmp[p][x] * inv_evec[i][x]
// with 16x16x4 MFMA tiles: a workgroup owns 16 patterns (one N tile) and its
// four waves split the rate categories; states pad to 16-row tiles and
// 4-wide K steps (20 states: two tiles, five exact steps). Phase A puts
// states on M and patterns on N; its
// C/D layout (lane = pattern + 16*g, register r = state row(g, r)) is the A
// layout phase B needs with patterns on M when phase B's K order runs over
// (state tile, register) pairs, so tmp stays in registers; phase B's output
// has patterns on rows and 16 consecutive states across lanes, so every
// store writes contiguous runs of a pattern's block. The rescale decision is
// per pattern over the whole block (normal: the maxima of every wave's
// categories meet in LDS) or per category (safe), as in the reference.
#pragma once
#include "plk_codon.h"

namespace plkdev {

template <typename T, int N, int C>
struct MfmaSmall {
    static constexpr int S = N + 1;             // LDS row stride (elements)
    static constexpr int NT = (N + 15) / 16;    // 16-state tiles
    static constexpr int KS = (N + 3) / 4;      // 4-wide K steps over states
    static constexpr int kInv = N * S;          // U^-1
    static constexpr int kTab = C * N * S;      // one internal child, all categories
    static constexpr size_t kBytes = (size_t)(kInv + 2 * kTab) * sizeof(T);
    static_assert(kBytes <= 64 * 1024, "tables must fit the workgroup's LDS");
};

template <typename T, int N>
__device__ inline void stage_inv(const double* __restrict__ inv, T* inv_sh) {
    constexpr int S = N + 1;
    stage_lds<N * N>(inv_sh, inv, N * N, [](int t) { return (t / N) * S + t % N; });
}

// Patterns per workgroup: 16, one MFMA N tile; the four waves split the
// rate categories (wave w owns categories w, w + 4, ...), which gives four
// times as many workgroups as one wave per 16 patterns would.
constexpr int kMfmaTile = 16;

// All steps of an op list for the patterns p0 .. p0 + 16*SUB - 1. Under each
// step's staged tables every wave computes its categories of the SUB
// 16-pattern sub-tiles one after another and stores them unscaled, with the
// per-(pattern, category) maxima in LDS; after one barrier the patterns
// that need rescaling (rare: underflow below 2^-256 or the FP32 threshold)
// are read back, scaled and stored again, and the scale counters written.
// The stored values equal scaling the computed ones directly. U^-1 is staged
// in inv_sh (row stride N + 1); tab_sh receives the two child tables of each
// step (all categories); mx_sh holds SUB x C x 16 maxima. PREFETCH stages
// the tables with every thread's loads issued first (one L2 round trip
// instead of one per element; it costs registers).
template <typename T, int N, int C, int SUB, bool PREFETCH = true>
__device__ inline void mfma_ops_tile(const InstDev& I, const plk_op* __restrict__ ops, int nops,
                                     const T* __restrict__ tables, int table_size, const T* inv_sh,
                                     T* tab_sh, T* mx_sh, int p_first) {
    using ST = typename Prec<T>::scale_t;
    using M = Mfma<T>;
    using acc_t = typename M::acc_t;
    using L = MfmaSmall<T, N, C>;
    constexpr int S = L::S, NT = L::NT, KS = L::KS, block = N * C;
    constexpr int CW = C < 4 ? C : 4;          // waves with categories
    constexpr int CPW = (C + CW - 1) / CW;     // categories per wave
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int col = lane & 15, g = lane >> 4;
    const bool active = wave < CW;
    T* partial = reinterpret_cast<T*>(I.partial);
    ST* scale = reinterpret_cast<ST*>(I.scale);
    const bool safe = I.safe != 0;

    for (int o = 0; o < nops; o++) {
        const plk_op op = ops[o];
        int left = 0, right = 1;
        if (!op.child[0].is_tip && op.child[1].is_tip) { left = 1; right = 0; }
        const plk_child ch[2] = {op.child[left], op.child[right]};
        const T* gt[2] = {tables + (size_t)(2 * o + left) * table_size,
                          tables + (size_t)(2 * o + right) * table_size};
        const bool tiptip = ch[0].is_tip && ch[1].is_tip;
        T* dst = partial + (size_t)op.dad_slot * I.slot_stride;
        // E tables of the internal children, every category, row x, column k
        // (the barrier also orders the previous step's maxima reads before
        // this step's writes)
        __syncthreads();
        for (int k = 0; k < 2; k++) {
            if (ch[k].is_tip) continue;
            if constexpr (PREFETCH) {
                stage_lds<C * N * N>(tab_sh + k * L::kTab, gt[k], C * N * N, [](int t) {
                    const int c = t / (N * N), r = t - c * N * N;
                    return c * N * S + (r / N) * S + r % N;
                });
            } else {
                T* tdst = tab_sh + k * L::kTab;
                for (int t = threadIdx.x; t < C * N * N; t += kThreads) {
                    const int c = t / (N * N), r = t - c * N * N;
                    tdst[c * N * S + (r / N) * S + r % N] = gt[k][t];
                }
            }
        }
        __syncthreads();
        if (active) {
#pragma unroll 1
            for (int sub = 0; sub < SUB; sub++) {
                const int p0 = p_first + 16 * sub;
                const int pb = p0 + col;   // this lane's pattern as a B operand
#pragma unroll
                for (int j = 0; j < CPW; j++) {
                    const int c = wave + CW * j;
                    if (c >= C) break;
                    acc_t prod[NT];   // tmp[x = 16 xb + row(g, r)][pattern col]
#pragma unroll
                    for (int k = 0; k < 2; k++) {
                        acc_t a[NT];
                        if (ch[k].is_tip) {
                            const int s = I.states[(size_t)ch[k].id * I.Ppad + pb];
                            const T* lut = gt[k] + (size_t)(s * C + c) * N;
#pragma unroll
                            for (int xb = 0; xb < NT; xb++)
#pragma unroll
                                for (int r = 0; r < 4; r++) {
                                    const int x = 16 * xb + M::row(g, r);
                                    a[xb][r] = x < N ? lut[x] : (T)0;
                                }
                        } else {
                            const T* E = tab_sh + k * L::kTab + c * N * S;
                            const T* cv = partial + (size_t)ch[k].id * I.slot_stride + (size_t)pb * block + c * N;
                            T bv[KS];
#pragma unroll
                            for (int st = 0; st < KS; st++) {
                                const int kk = 4 * st + g;
                                bv[st] = kk < N ? cv[kk] : (T)0;
                            }
#pragma unroll
                            for (int xb = 0; xb < NT; xb++) a[xb] = acc_t{0, 0, 0, 0};
#pragma unroll
                            for (int st = 0; st < KS; st++) {
                                const int kk = 4 * st + g;
                                const bool kin = kk < N;
#pragma unroll
                                for (int xb = 0; xb < NT; xb++) {
                                    const int x = 16 * xb + col;
                                    const T av = (kin && x < N) ? E[x * S + kk] : (T)0;
                                    a[xb] = M::mma(av, bv[st], a[xb]);
                                }
                            }
                        }
                        if (k == 0) {
#pragma unroll
                            for (int xb = 0; xb < NT; xb++) prod[xb] = a[xb];
                        } else {
#pragma unroll
                            for (int xb = 0; xb < NT; xb++) prod[xb] = prod[xb] * a[xb];
                        }
                    }
                    // phase B, patterns on M: A[p = col][k = g] is prod[xb][r] for
                    // state x = 16 xb + row(g, r); B[k][i = col] = U^-1[i][x];
                    // d[ib][r] = dad[p0 + row(g, r)][c][16 ib + col]
                    acc_t d[NT];
#pragma unroll
                    for (int ib = 0; ib < NT; ib++) d[ib] = acc_t{0, 0, 0, 0};
#pragma unroll
                    for (int xb = 0; xb < NT; xb++)
#pragma unroll
                        for (int r = 0; r < 4; r++) {
                            if (16 * xb + M::row(0, r) >= N) continue;   // a step with no state inside
                            const int x = 16 * xb + M::row(g, r);
                            const T av = prod[xb][r];
#pragma unroll
                            for (int ib = 0; ib < NT; ib++) {
                                const int i = 16 * ib + col;
                                const T bv = (i < N && x < N) ? inv_sh[i * S + x] : (T)0;
                                d[ib] = M::mma(av, bv, d[ib]);
                            }
                        }
                    // unscaled store and the category maxima
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        T* row = dst + (size_t)(p0 + M::row(g, r)) * block + c * N;
                        T m = (T)0;
#pragma unroll
                        for (int ib = 0; ib < NT; ib++) {
                            const int i = 16 * ib + col;
                            if (i < N) {
                                row[i] = d[ib][r];
                                m = fmax(m, fabs(d[ib][r]));
                            }
                        }
                        m = row16_max(m);   // over the 16 lanes (states) of this pattern row
                        if (col == 0) mx_sh[(sub * C + c) * 16 + M::row(g, r)] = m;
                    }
                }
            }
        }
        __syncthreads();
        if (!active) continue;
        // rescale (rare) and scale counters; lane (g, col) handles the patterns
        // row(g, r) of each sub-tile, columns col, col + 16 of its categories
#pragma unroll 1
        for (int sub = 0; sub < SUB; sub++) {
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int q = M::row(g, r);
                const int p = p_first + 16 * sub + q;
                const bool may_scale = !tiptip && I.invar[p] == 0.0;
                int kblock = 0;
                if (!safe && may_scale) {
                    T bm = mx_sh[(sub * C) * 16 + q];
                    for (int c = 1; c < C; c++) bm = fmax(bm, mx_sh[(sub * C + c) * 16 + q]);
                    if (Prec<T>::underflow(bm)) kblock = Prec<T>::exponent(bm);
                }
#pragma unroll
                for (int j = 0; j < CPW; j++) {
                    const int c = wave + CW * j;
                    if (c >= C) break;
                    int kx = kblock;
                    if (safe && may_scale) {
                        const T cm = mx_sh[(sub * C + c) * 16 + q];
                        kx = Prec<T>::underflow(cm) ? Prec<T>::exponent(cm) : 0;
                    }
                    if (kx) {
                        T* row = dst + (size_t)p * block + c * N;
#pragma unroll
                        for (int ib = 0; ib < NT; ib++) {
                            const int i = 16 * ib + col;
                            if (i < N) row[i] = Prec<T>::apply(row[i], kx);
                        }
                    }
                    // counters: per category (safe) or once per pattern by wave 0
                    if (col == 0 && (safe || c == 0)) {
                        const size_t sidx = safe ? (size_t)p * C + c : (size_t)p;
                        int sc = kx;
                        if (!ch[0].is_tip) sc += scale[(size_t)ch[0].id * I.scale_stride + sidx];
                        if (!ch[1].is_tip) sc += scale[(size_t)ch[1].id * I.scale_stride + sidx];
                        scale[(size_t)op.dad_slot * I.scale_stride + sidx] = (ST)sc;
                    }
                }
            }
        }
    }
}

// SUB 16-pattern sub-tiles per workgroup share each step's staged tables
template <typename T, int N, int C, int SUB>
__global__ void __launch_bounds__(kThreads)
k_traverse_mfma(InstDev I, ModelDev m, const plk_op* __restrict__ ops, int nops,
                const T* __restrict__ tables, int table_size) {
    using L = MfmaSmall<T, N, C>;
    __shared__ T inv_sh[L::kInv];
    __shared__ T tab_sh[2 * L::kTab];
    __shared__ T mx_sh[SUB * C * 16];
    stage_inv<T, N>(m.inv_evec, inv_sh);
    mfma_ops_tile<T, N, C, SUB>(I, ops, nops, tables, table_size, inv_sh, tab_sh, mx_sh,
                                blockIdx.x * kMfmaTile * SUB);
}

}  // namespace plkdev
