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
// per pattern over the whole block (normal: the waves exchange their maxima
// through LDS) or per category (safe), as in the reference.
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

template <typename T>
__device__ inline void stage_inv(const double* __restrict__ inv, T* inv_sh, int N, int S) {
    for (int t = threadIdx.x; t < N * N; t += kThreads) inv_sh[(t / N) * S + t % N] = (T)inv[t];
}

// Patterns per workgroup: 16, one MFMA N tile; the four waves split the
// rate categories (wave w owns categories w, w + 4, ...), which gives four
// times as many workgroups as one wave per 16 patterns would.
constexpr int kMfmaTile = 16;

// All steps of an op list for the patterns p0 .. p0 + 16*SUB - 1, SUB
// 16-pattern tiles one after another under each step's staged tables. U^-1
// is staged in inv_sh (row stride N + 1); tab_sh receives the two child
// tables of each step (all categories); red_sh holds 2 x 4 x 16 per-wave
// maxima.
template <typename T, int N, int C, int SUB>
__device__ inline void mfma_ops_tile(const InstDev& I, const plk_op* __restrict__ ops, int nops,
                                     const T* __restrict__ tables, int table_size, const T* inv_sh,
                                     T* tab_sh, T* red_sh, int p_first) {
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
        // E tables of the internal children, every category, row x, column k
        __syncthreads();
        for (int k = 0; k < 2; k++) {
            if (ch[k].is_tip) continue;
            T* dst = tab_sh + k * L::kTab;
            for (int t = threadIdx.x; t < C * N * N; t += kThreads) {
                const int c = t / (N * N), r = t - c * N * N;
                dst[c * N * S + (r / N) * S + r % N] = gt[k][t];
            }
        }
        __syncthreads();
#pragma unroll 1
        for (int sub = 0; sub < SUB; sub++) {
        const int p0 = p_first + 16 * sub;
        const int pb = p0 + col;   // this lane's pattern as a B operand / A row
        T* red = red_sh + (sub & 1) * 64;   // alternating, so no barrier is needed before refilling
        acc_t d[CPW][NT];   // d[j][ib][r] = dad[p0 + row(g, r)][category wave + CW j][16 ib + col]
        T wmax[4] = {(T)0, (T)0, (T)0, (T)0};   // per register-row pattern, over this wave's categories
        if (active) {
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
                // state x = 16 xb + row(g, r); B[k][i = col] = U^-1[i][x]
#pragma unroll
                for (int ib = 0; ib < NT; ib++) d[j][ib] = acc_t{0, 0, 0, 0};
#pragma unroll
                for (int xb = 0; xb < NT; xb++)
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        if (16 * xb + M::row(0, r) >= N) continue;   // a step with no state inside
                        const int x = 16 * xb + M::row(g, r);
                        const T a = prod[xb][r];
#pragma unroll
                        for (int ib = 0; ib < NT; ib++) {
                            const int i = 16 * ib + col;
                            const T bv = (i < N && x < N) ? inv_sh[i * S + x] : (T)0;
                            d[j][ib] = M::mma(a, bv, d[j][ib]);
                        }
                    }
            }
        }
        // per-pattern maxima: register r holds pattern p0 + row(g, r)
        T cmax[CPW][4];
#pragma unroll
        for (int j = 0; j < CPW; j++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                T m = (T)0;
                if (active && wave + CW * j < C) {
#pragma unroll
                    for (int ib = 0; ib < NT; ib++)
                        if (16 * ib + col < N) m = fmax(m, fabs(d[j][ib][r]));
                }
#pragma unroll
                for (int dl = 1; dl < 16; dl <<= 1) m = fmax(m, __shfl_xor(m, dl));
                cmax[j][r] = m;
                wmax[r] = fmax(wmax[r], m);
            }
        if (!safe) {
            // the block maximum spans every wave's categories
            if (col == 0)
#pragma unroll
                for (int r = 0; r < 4; r++) red[wave * 16 + M::row(g, r)] = wmax[r];
            __syncthreads();
#pragma unroll
            for (int r = 0; r < 4; r++) {
                T m = red[M::row(g, r)];
                for (int w = 1; w < CW; w++) m = fmax(m, red[w * 16 + M::row(g, r)]);
                wmax[r] = m;
            }
        }
        if (!active) continue;
        T* dst = partial + (size_t)op.dad_slot * I.slot_stride;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int p = p0 + M::row(g, r);
            const bool may_scale = !tiptip && I.invar[p] == 0.0;
            const int kblock = !safe && may_scale && Prec<T>::underflow(wmax[r]) ? Prec<T>::exponent(wmax[r]) : 0;
            T* row = dst + (size_t)p * block;
#pragma unroll
            for (int j = 0; j < CPW; j++) {
                const int c = wave + CW * j;
                if (c >= C) break;
                const int kx = safe ? (may_scale && Prec<T>::underflow(cmax[j][r]) ? Prec<T>::exponent(cmax[j][r]) : 0)
                                    : kblock;
#pragma unroll
                for (int ib = 0; ib < NT; ib++) {
                    const int i = 16 * ib + col;
                    if (i < N) {
                        T v = d[j][ib][r];
                        if (kx) v = Prec<T>::apply(v, kx);
                        row[c * N + i] = v;
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
        }   // sub-tiles
    }
}

template <typename T, int N, int C>
__global__ void __launch_bounds__(kThreads)
k_traverse_mfma(InstDev I, ModelDev m, const plk_op* __restrict__ ops, int nops,
                const T* __restrict__ tables, int table_size) {
    using L = MfmaSmall<T, N, C>;
    __shared__ T inv_sh[L::kInv];
    __shared__ T tab_sh[2 * L::kTab];
    __shared__ T red_sh[2 * 4 * 16];
    stage_inv<T>(m.inv_evec, inv_sh, N, L::S);
    mfma_ops_tile<T, N, C, 1>(I, ops, nops, tables, table_size, inv_sh, tab_sh, red_sh,
                              blockIdx.x * kMfmaTile);
}

}  // namespace plkdev
