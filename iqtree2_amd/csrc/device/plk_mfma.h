// Partial-likelihood updates for protein (20 states) on the matrix cores:
// tree/phylokernelnew.h:1632-2031 as two contractions per category and
// traversal step,
//   phase A  tmp[x][p] = prod_children sum_k E_child[c][x][k] * child[p][c][k]
//            (tip children: the lookup leaf[s_p][c][x])
//   phase B  dad[p][i] = sum_x tmp[p][x] * inv_evec[i][x]
// with 16x16x4 MFMA tiles. A wave owns one rate category of 16*SUB patterns
// (SUB sub-tiles of one MFMA N tile each); the four waves of a workgroup
// split the categories. Phase A puts states on M (20 states: two M tiles)
// and patterns on N, with K = 20 in five exact 4-wide steps; its C/D layout
// (lane = pattern + 16*g, register r = state row(g, r)) is the A layout
// phase B needs with patterns on M when phase B's K order runs over
// (state tile, register) pairs, so tmp stays in registers; phase B's output
// has patterns on rows and 16 consecutive states across lanes, so a store
// writes contiguous runs of a pattern's block.
//
// Operands come straight from HBM/L2 into registers, with no LDS staging and
// no barrier for them:
//   - an internal child's table is stored in MFMA fragment order (below), so
//     each phase-A A operand is one contiguous 64-lane read;
//   - in step st, lane (col, g) takes K index 5g + st, so a lane's five B
//     operands (child values) are five consecutive elements of its pattern;
//   - U^-1 (phase B's B operand) is staged in LDS in lane order once per tile;
//   - tip states are read one step ahead, so a tip child's lookup rows are
//     read together with the internal child's values.
// The only exchange between waves is the per-(pattern, category) maximum that
// the normal-mode rescale decision needs over all categories: one LDS write
// and one barrier per step (double-buffered, so no second barrier). Values are stored unscaled;
// the rare pattern that needs rescaling is re-read, scaled and rewritten by
// the wave that wrote it, which also writes the scale counters
// (tree/phylokernelnew.h:1993-2030).
#pragma once
#include "plk_codon.h"

namespace plkdev {

constexpr int kMfmaTile = 16;              // patterns per MFMA N tile

// All steps of an op list for the patterns p_first .. p_first + 16*SUB - 1.
// mx_sh holds 2 x SUB x C x 16 maxima (one buffer per step parity).
template <typename T, int C, int SUB>
__device__ inline void prot_ops_tile(const InstDev& I, const double* __restrict__ inv_evec,
                                     const plk_op* __restrict__ ops, int nops, const T* __restrict__ tables,
                                     int table_size, T* mx_sh, int p_first) {
    static_assert(SUB >= 1 && SUB <= 4, "16 to 64 patterns per wave");
    using ST = typename Prec<T>::scale_t;
    using M = Mfma<T>;
    using acc_t = typename M::acc_t;
    constexpr int N = kProtN, block = N * C, KS = kProtKSteps;
    constexpr int CW = C < 4 ? C : 4;          // waves with categories
    constexpr int CPW = (C + CW - 1) / CW;     // categories per wave
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int col = lane & 15, g = lane >> 4;
    // phase B's B operand, B[k = g][n = col] = U^-1[16 ib + col][x] for the
    // K steps (xb, r), x = 16 xb + row(g, r), in LDS in lane order:
    // ub_sh[((ib * 2 + xb) * 4 + r) * 64 + lane] (in registers it costs the
    // FP64 tile a wave per SIMD); staged by every thread before the waves
    // without categories leave
    __shared__ T ub_sh[2 * 2 * 4 * 64];
    for (int t = threadIdx.x; t < 2 * 2 * 4 * 64; t += kThreads) {
        const int tb = t >> 9, txb = (t >> 8) & 1, tr = (t >> 6) & 3, tl = t & 63;
        const int i = 16 * tb + (tl & 15), x = 16 * txb + M::row(tl >> 4, tr);
        ub_sh[t] = (i < N && x < N) ? (T)inv_evec[i * N + x] : (T)0;
    }
    __syncthreads();
    if (wave >= CW) {
        // no categories: only the per-step barriers
        for (int o = 0; o < nops; o++) __syncthreads();
        return;
    }
    T* partial = reinterpret_cast<T*>(I.partial);
    ST* scale = reinterpret_cast<ST*>(I.scale);
    const bool safe = I.safe != 0;
    // counters and rescale decisions: lane l owns pattern p_first + l
    const bool plane = lane < 16 * SUB;
    const int pl = p_first + lane;
    const bool no_invar = plane && I.invar[pl] == 0.0;
    const bool counters = safe || wave == 0;   // normal mode: one counter per pattern, wave 0

    // tip states of the next step's children (this lane's pattern column of
    // sub-tile s in byte s)
    unsigned s_next[2];
    auto order = [](const plk_op& op, plk_child (&ch)[2], int& left) {
        left = (!op.child[0].is_tip && op.child[1].is_tip) ? 1 : 0;
        ch[0] = op.child[left];
        ch[1] = op.child[1 - left];
    };
    auto load_states = [&](int o) {
        plk_child ch[2];
        int left;
        order(ops[o], ch, left);
#pragma unroll
        for (int k = 0; k < 2; k++) {
            s_next[k] = 0u;
            if (ch[k].is_tip) {
                const uint8_t* sp = I.states + (size_t)ch[k].id * I.Ppad + p_first + col;
#pragma unroll
                for (int s = 0; s < SUB; s++) s_next[k] |= (unsigned)sp[16 * s] << (8 * s);
            }
        }
    };
    if (nops > 0) load_states(0);

    for (int o = 0; o < nops; o++) {
        const plk_op op = ops[o];
        plk_child ch[2];
        int left;
        order(op, ch, left);
        const T* gt[2] = {tables + (size_t)(2 * o + left) * table_size,
                          tables + (size_t)(2 * o + 1 - left) * table_size};
        const bool tiptip = ch[0].is_tip && ch[1].is_tip;
        const unsigned s_cur[2] = {s_next[0], s_next[1]};
        if (o + 1 < nops) load_states(o + 1);
        T* dst = partial + (size_t)op.dad_slot * I.slot_stride;
        T* mx = mx_sh + (o & 1) * (SUB * C * 16);

        // the children's scale counters of the lane's pattern (issued early)
        int sc_child[CPW];
#pragma unroll
        for (int j = 0; j < CPW; j++) {
            const int c = wave + CW * j;
            sc_child[j] = 0;
            if (c < C && counters && plane && (safe || j == 0)) {
                const size_t sidx = safe ? (size_t)pl * C + c : (size_t)pl;
#pragma unroll
                for (int k = 0; k < 2; k++)
                    if (!ch[k].is_tip) sc_child[j] += scale[(size_t)ch[k].id * I.scale_stride + sidx];
            }
        }

#pragma unroll
        for (int j = 0; j < CPW; j++) {
            const int c = wave + CW * j;
            if (c >= C) break;
            // phase-A A operands of the internal children (one fragment run each)
            T ea[2][2][KS];
#pragma unroll
            for (int k = 0; k < 2; k++) {
                if (ch[k].is_tip) continue;
                const T* fr = gt[k] + (size_t)c * kProtFrag + lane;
#pragma unroll
                for (int xb = 0; xb < 2; xb++)
#pragma unroll
                    for (int st = 0; st < KS; st++) ea[k][xb][st] = fr[(xb * KS + st) * 64];
            }
            // one sub-tile at a time: unrolled, the sub-tiles' loads are all
            // hoisted and hold about 80 more registers (FP64), costing occupancy
#pragma unroll 1
            for (int s = 0; s < SUB; s++) {
                const int p0 = p_first + 16 * s;
                acc_t prod[2];   // tmp[x = 16 xb + row(g, r)][pattern p0 + col]
#pragma unroll
                for (int k = 0; k < 2; k++) {
                    acc_t a[2];
                    if (ch[k].is_tip) {
                        const int st8 = (int)((s_cur[k] >> (8 * s)) & 0xffu);
                        const T* lut = gt[k] + (size_t)(st8 * C + c) * N;
#pragma unroll
                        for (int xb = 0; xb < 2; xb++)
#pragma unroll
                            for (int r = 0; r < 4; r++) {
                                const int x = 16 * xb + M::row(g, r);
                                a[xb][r] = x < N ? lut[x] : (T)0;
                            }
                    } else {
                        const T* cv = partial + (size_t)ch[k].id * I.slot_stride + (size_t)(p0 + col) * block +
                                      c * N + 5 * g;
                        T bv[KS];
#pragma unroll
                        for (int st = 0; st < KS; st++) bv[st] = cv[st];
#pragma unroll
                        for (int xb = 0; xb < 2; xb++) a[xb] = acc_t{0, 0, 0, 0};
#pragma unroll
                        for (int st = 0; st < KS; st++)
#pragma unroll
                            for (int xb = 0; xb < 2; xb++) a[xb] = M::mma(ea[k][xb][st], bv[st], a[xb]);
                    }
                    if (k == 0) {
                        prod[0] = a[0];
                        prod[1] = a[1];
                    } else {
                        prod[0] = prod[0] * a[0];
                        prod[1] = prod[1] * a[1];
                    }
                }
                // phase B, patterns on M: A[p = col][k = g] is prod[xb][r];
                // d[ib][r] = dad[p0 + row(g, r)][c][16 ib + col]
                acc_t d[2] = {acc_t{0, 0, 0, 0}, acc_t{0, 0, 0, 0}};
#pragma unroll
                for (int xb = 0; xb < 2; xb++)
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        if (16 * xb + M::row(0, r) >= N) continue;   // a K step with no state
#pragma unroll
                        for (int ib = 0; ib < 2; ib++) d[ib] = M::mma(prod[xb][r], ub_sh[((ib * 2 + xb) * 4 + r) * 64 + lane], d[ib]);
                    }
                // unscaled store and the per-(pattern, category) maxima. Only
                // whether a maximum is below the rescale threshold matters
                // unless it is (and then its value, for the FP32 exponent):
                // when every row of the register holds a value at or above
                // the threshold (one ballot), each row records 1, which gives
                // the same decisions; otherwise the rows are reduced exactly.
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    T* row = dst + (size_t)(p0 + M::row(g, r)) * block + c * N;
                    row[col] = d[0][r];
                    T m = fabs(d[0][r]);
                    if (col < N - 16) {
                        row[16 + col] = d[1][r];
                        m = fmax(m, fabs(d[1][r]));
                    }
                    const unsigned long long big = __builtin_amdgcn_ballot_w64(m >= Prec<T>::threshold);
                    const bool rows_big = (big & 0xFFFFull) && (big & 0xFFFF0000ull) &&
                                          (big & 0xFFFF00000000ull) && (big >> 48);
                    m = rows_big ? (T)1 : row16_max(m);
                    if (col == 0) mx[(s * C + c) * 16 + M::row(g, r)] = m;
                }
            }
        }
        // every wave's maxima (normal mode decides over all categories; in
        // safe mode the barrier only orders a wave's own LDS writes and reads)
        __syncthreads();
        // rescale decisions, rare rewrites and counters; lane l owns pattern
        // p_first + l (sub-tile l / 16, row l % 16)
        const bool may_scale = no_invar && !tiptip;
        const int ms = (lane >> 4) * C * 16 + (lane & 15);
        int kblock = 0;
        if (!safe && may_scale) {
            T bm = mx[ms];
#pragma unroll
            for (int c = 1; c < C; c++) bm = fmax(bm, mx[ms + c * 16]);
            if (Prec<T>::underflow(bm)) kblock = Prec<T>::exponent(bm);
        }
#pragma unroll
        for (int j = 0; j < CPW; j++) {
            const int c = wave + CW * j;
            if (c >= C) break;
            int kx = kblock;
            if (safe) {
                kx = 0;
                if (may_scale) {
                    const T cm = mx[ms + c * 16];
                    if (Prec<T>::underflow(cm)) kx = Prec<T>::exponent(cm);
                }
            }
            if (__builtin_amdgcn_ballot_w64(kx != 0)) {
                if (kx) {
                    T* row = dst + (size_t)pl * block + c * N;
#pragma unroll 4
                    for (int i = 0; i < N; i++) row[i] = Prec<T>::apply(row[i], kx);
                }
            }
            if (counters && plane && (safe || j == 0)) {
                const size_t sidx = safe ? (size_t)pl * C + c : (size_t)pl;
                scale[(size_t)op.dad_slot * I.scale_stride + sidx] = (ST)(kx + sc_child[j]);
            }
        }
    }
}

// LDS elements of prot_ops_tile's maxima
template <int C, int SUB>
constexpr int prot_mx_elems() { return 2 * SUB * C * 16; }

// SUB 16-pattern sub-tiles per wave
template <typename T, int N, int C, int SUB>
__global__ void __launch_bounds__(kThreads)
k_traverse_mfma(InstDev I, ModelDev m, const plk_op* __restrict__ ops, int nops,
                const T* __restrict__ tables, int table_size) {
    static_assert(N == kProtN, "20-state tile");
    __shared__ T mx_sh[prot_mx_elems<C, SUB>()];
    prot_ops_tile<T, C, SUB>(I, m.inv_evec, ops, nops, tables, table_size, mx_sh,
                             blockIdx.x * kMfmaTile * SUB);
}

}  // namespace plkdev
