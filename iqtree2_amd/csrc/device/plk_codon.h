// Codon models (61 states) on the matrix cores.
//
// IQ-TREE runs every model whose state count is not 4 or 20 with
// per-category scaling (tree/phylotreesse.cpp:95-96), so this path is safe
// mode only. The update of one traversal step (tree/phylokernelnew.h:1632-2031)
// is two dense contractions per category:
//   phase A  tmp[x][p] = prod_children sum_k E_child[c][x][k] * child[p][c][k]
//            (tip children: tmp[x][p] = leaf[s_p][c][x])
//   phase B  dad[i][p] = sum_x inv_evec[i][x] * tmp[x][p]
// followed by the per-(pattern, category) rescale when max_i |dad| < 2^-256
// (FP32 storage: the exact power-of-two rescale of the float form).
//
// Work split: a workgroup of four waves owns 64 patterns of one rate
// category (safe mode keeps a scale counter per category, so categories are
// independent), each wave 16 patterns as the N dimension of 16x16x4 MFMA tiles; the 61 states pad to 64 =
// four M tiles and sixteen K steps. E_child[c] and U^-1 sit in LDS (row
// stride 65 elements); child vectors are read straight from HBM into the B
// operand. In the C/D layout of the phase-A product a lane holds one
// pattern and four states (which four depends on the element type, see
// Mfma::row); that is exactly the B layout phase B needs when its K order
// runs over (x tile, register) pairs, so tmp never leaves registers. The new
// partials are staged through LDS so that their HBM write is row-contiguous.
#pragma once
#include "plk_kernels.h"

namespace plkdev {

// Workgroup copy of n <= NMAX elements, src[t] -> dst[map(t)]: every thread
// issues all its loads before its LDS stores (a strided loop with a
// data-dependent trip count pays one L2 round trip per element otherwise).
// CHUNK bounds the loads a thread holds at once (registers).
template <int NMAX, int CHUNK = 64, typename TD, typename TS, typename Map>
__device__ inline void stage_lds(TD* dst, const TS* __restrict__ src, int n, Map map) {
    constexpr int IT = (NMAX + kThreads - 1) / kThreads;
    constexpr int CH = IT < CHUNK ? IT : CHUNK;
#pragma unroll
    for (int j0 = 0; j0 < IT; j0 += CH) {
        TS v[CH];
#pragma unroll
        for (int j = 0; j < CH; j++) {
            const int t = (int)threadIdx.x + (j0 + j) * kThreads;
            if (j0 + j < IT && t < n) v[j] = src[t];
        }
#pragma unroll
        for (int j = 0; j < CH; j++) {
            const int t = (int)threadIdx.x + (j0 + j) * kThreads;
            if (j0 + j < IT && t < n) dst[map(t)] = (TD)v[j];
        }
    }
}


typedef double mfma_d4 __attribute__((ext_vector_type(4)));
typedef float mfma_f4 __attribute__((ext_vector_type(4)));

template <typename T> struct Mfma;
template <> struct Mfma<double> {
    using acc_t = mfma_d4;
    // D[16x16] += A[16x4] * B[4x16]: lane l supplies A[l%16][l/16] and
    // B[l/16][l%16]; register r of lane l holds D[row(l/16, r)][l%16]
    __device__ static acc_t mma(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    __device__ static constexpr int row(int g, int r) { return g + 4 * r; }
};
template <> struct Mfma<float> {
    using acc_t = mfma_f4;
    __device__ static acc_t mma(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    __device__ static constexpr int row(int g, int r) { return 4 * g + r; }
};

constexpr int kCodonN = 61;        // sense codons of the standard code
constexpr int kCodonStride = 65;   // LDS row stride in elements
constexpr int kCodonTile = 64;     // patterns per workgroup

// All steps of an op list for the kCodonTile patterns from p0 in rate
// category c (per-category scaling makes categories independent). inv_sh
// holds U^-1 (row i, column x, row stride kCodonStride); e_sh receives one
// internal child's E[c] (row x, column k) during phase A and the dad tile
// [64 patterns][61] while it is written out. Ends with a barrier.
template <typename T>
__device__ inline void codon_ops_tile(const InstDev& I, const plk_op* __restrict__ ops, int nops,
                                      const T* __restrict__ tables, int table_size, int c, int p0,
                                      T* e_sh, const T* inv_sh) {
    using ST = typename Prec<T>::scale_t;
    using M = Mfma<T>;
    using acc_t = typename M::acc_t;
    constexpr int N = kCodonN, S = kCodonStride;
    const int C = I.C;
    const int block = N * C;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int col = lane & 15, g = lane >> 4;
    const int pl = wave * 16 + col;   // this lane's pattern within the tile
    const int pw = p0 + pl;
    T* partial = reinterpret_cast<T*>(I.partial);
    ST* scale = reinterpret_cast<ST*>(I.scale);

    for (int o = 0; o < nops; o++) {
        const plk_op op = ops[o];
        int left = 0, right = 1;
        if (!op.child[0].is_tip && op.child[1].is_tip) { left = 1; right = 0; }
        const plk_child ch[2] = {op.child[left], op.child[right]};
        const T* gt[2] = {tables + (size_t)(2 * o + left) * table_size,
                          tables + (size_t)(2 * o + right) * table_size};
        const bool tiptip = ch[0].is_tip && ch[1].is_tip;
        T* dst = partial + (size_t)op.dad_slot * I.slot_stride;
        {
            acc_t prod[4];   // tmp[x = 16*xb + row(g, r)][pattern pl] in prod[xb][r]
#pragma unroll
            for (int k = 0; k < 2; k++) {
                acc_t a[4];
                if (ch[k].is_tip) {
                    const int s = I.states[(size_t)ch[k].id * I.Ppad + pw];
                    const T* lut = gt[k] + (size_t)(s * C + c) * N;
#pragma unroll
                    for (int xb = 0; xb < 4; xb++)
#pragma unroll
                        for (int r = 0; r < 4; r++) {
                            const int x = 16 * xb + M::row(g, r);
                            a[xb][r] = x < N ? lut[x] : (T)0;
                        }
                } else {
                    // the child's values first (they do not wait for the
                    // tables), then the tables staged while they arrive
                    const T* cv = partial + (size_t)ch[k].id * I.slot_stride + (size_t)pw * block + c * N;
                    T bv[16];
#pragma unroll
                    for (int st = 0; st < 16; st++) {
                        const int kk = 4 * st + g;
                        bv[st] = kk < N ? cv[kk] : (T)0;
                    }
                    __syncthreads();   // earlier readers of e_sh are done
                    const T* e = gt[k] + (size_t)c * N * N;
                    stage_lds<N * N>(e_sh, e, N * N, [](int t) { return (t / N) * S + t % N; });
                    __syncthreads();
#pragma unroll
                    for (int xb = 0; xb < 4; xb++) a[xb] = acc_t{0, 0, 0, 0};
#pragma unroll
                    for (int st = 0; st < 16; st++) {
                        const int kk = 4 * st + g;
                        const bool kin = kk < N;
                        const T b = bv[st];
#pragma unroll
                        for (int xb = 0; xb < 4; xb++) {
                            const int x = 16 * xb + col;
                            const T av = (kin && x < N) ? e_sh[x * S + kk] : (T)0;
                            a[xb] = M::mma(av, b, a[xb]);
                        }
                    }
                }
                if (k == 0) {
#pragma unroll
                    for (int xb = 0; xb < 4; xb++) prod[xb] = a[xb];
                } else {
#pragma unroll
                    for (int xb = 0; xb < 4; xb++) prod[xb] = prod[xb] * a[xb];
                }
            }
            // phase B: K runs over (xb, r); lane (col, g) holds tmp[16xb + row(g, r)][col]
            acc_t d[4];
#pragma unroll
            for (int ib = 0; ib < 4; ib++) d[ib] = acc_t{0, 0, 0, 0};
#pragma unroll
            for (int xb = 0; xb < 4; xb++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int x = 16 * xb + M::row(g, r);
                    const T b = prod[xb][r];
#pragma unroll
                    for (int ib = 0; ib < 4; ib++) {
                        const int i = 16 * ib + col;
                        const T av = (i < N && x < N) ? inv_sh[i * S + x] : (T)0;
                        d[ib] = M::mma(av, b, d[ib]);
                    }
                }
            // d[ib][r] = dad[i = 16ib + row(g, r)][pattern pl]; per-(pattern, category) rescale
            T mx = (T)0;
#pragma unroll
            for (int ib = 0; ib < 4; ib++)
#pragma unroll
                for (int r = 0; r < 4; r++)
                    if (16 * ib + M::row(g, r) < N) mx = fmax(mx, fabs(d[ib][r]));
            mx = fmax(mx, __shfl_xor(mx, 16));
            mx = fmax(mx, __shfl_xor(mx, 32));
            const size_t sidx = (size_t)pw * C + c;
            int sc = 0, kx = 0;
            if (!ch[0].is_tip) sc += scale[(size_t)ch[0].id * I.scale_stride + sidx];
            if (!ch[1].is_tip) sc += scale[(size_t)ch[1].id * I.scale_stride + sidx];
            if (!tiptip && Prec<T>::underflow(mx) && I.invar[pw] == 0.0) {
                kx = Prec<T>::exponent(mx);
                sc += kx;
            }
            if (g == 0) scale[(size_t)op.dad_slot * I.scale_stride + sidx] = (ST)sc;
            __syncthreads();   // phase-A readers of e_sh and the previous write-out are done
            T* stage = e_sh;   // [64][N]
#pragma unroll
            for (int ib = 0; ib < 4; ib++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int i = 16 * ib + M::row(g, r);
                    if (i < N) {
                        T v = d[ib][r];
                        if (kx) v = Prec<T>::apply(v, kx);
                        stage[pl * N + i] = v;
                    }
                }
            __syncthreads();
            for (int t = threadIdx.x; t < kCodonTile * N; t += kThreads) {
                const int pp = t / N, i = t - pp * N;
                dst[(size_t)(p0 + pp) * block + c * N + i] = stage[t];
            }
            // the write-out stores other threads' patterns: the next step's
            // child loads (issued before its first barrier) must follow it
            __syncthreads();
        }
    }
}

template <typename T>
__global__ void __launch_bounds__(kThreads)
k_traverse_codon(InstDev I, ModelDev m, const plk_op* __restrict__ ops, int nops,
                 const T* __restrict__ tables, int table_size) {
    constexpr int N = kCodonN, S = kCodonStride;
    __shared__ T e_sh[N * S];
    __shared__ T inv_sh[N * S];
    stage_lds<N * N>(inv_sh, m.inv_evec, N * N, [](int t) { return (t / N) * S + t % N; });
    __syncthreads();
    codon_ops_tile<T>(I, ops, nops, tables, table_size, blockIdx.y, blockIdx.x * kCodonTile, e_sh, inv_sh);
}

// Branch ends for the matrix-core kernels, one thread per (pattern,
// category) with the C lanes of a pattern adjacent, streaming the
// N-element vectors. MODE_BRANCH: pattern lnL (tree/phylokernelnew.h:
// 2753-3104; safe mode aligns the per-category counters to the least-scaled
// category) and the workgroup sum; MODE_THETA: theta and buffer_scale
// (tree/phylokernelnew.h:2091-2230).
template <typename T, int N, int C>
__global__ void __launch_bounds__(kThreads)
k_root_stream(InstDev I, ModelDev m, int mode, plk_end node, plk_end dad,
              const double* __restrict__ root_tab, double* __restrict__ block_sums) {
    using ST = typename Prec<T>::scale_t;
    __shared__ double red[kThreads / 64];
    constexpr int PB = kThreads / C;
    constexpr int block = N * C;
    const int c = threadIdx.x % C;
    const int p = blockIdx.x * PB + threadIdx.x / C;
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
    // safe mode: per-category counters aligned to the least-scaled category
    int minsc = sc;
    double adjust = 1.0;
    if (safe) {
        minsc = group_min_int<C>(sc);
        adjust = Prec<T>::safe_adjust(sc - minsc);
    }
    if (mode == MODE_THETA) {
        T* th = reinterpret_cast<T*>(I.theta) + (size_t)p * block + c * N;
        if (dad.is_tip) {
            const double* tp = m.tip + s * N;
            for (int i = 0; i < N; i++) th[i] = (T)((double)nv[i] * tp[i] * adjust);
        } else {
            for (int i = 0; i < N; i++) th[i] = (T)((double)nv[i] * (double)dv[i] * adjust);
        }
        if (c == 0) I.buffer_scale[p] = (double)minsc * Prec<T>::log_unit;
        return;
    }
    double lhc = 0.0;
    if (dad.is_tip) {
        const double* lut = root_tab + block + (size_t)(s * C + c) * N;
        lhc = lut[0] * (double)nv[0];
        for (int i = 1; i < N; i++) lhc = fma(lut[i], (double)nv[i], lhc);
    } else {
        const double* val = root_tab + c * N;
        lhc = val[0] * (double)dv[0] * (double)nv[0];
        for (int i = 1; i < N; i++) lhc = fma(val[i] * (double)dv[i], (double)nv[i], lhc);
    }
    const double lh0 = safe ? group_sum_sumvec<C>(lhc * adjust) : group_sum_ordered<C>(lhc);
    double contrib = 0.0;
    if (c == 0) {
        const double lh = fabs(lh0) + I.invar[p];
        const int pobs = I.P - I.nasc;
        if (p < pobs) {
            const double lnp = log(lh) + (double)minsc * Prec<T>::log_unit;
            I.ptn_lh[p] = lnp;
            contrib = lnp * I.freq[p];
        } else if (p < I.P) {
            I.asc[4 * (p - pobs)] = Prec<T>::asc_scale(lh, minsc);   // +ASC
        }
    }
    double v[1] = {contrib};
    block_sum<1>(v, red);
    if (threadIdx.x == 0) block_sums[blockIdx.x] = v[0];
}

}  // namespace plkdev
