// State-basis kernels for non-reversible and mixture models: the kernel set
// of tree/phylokernelnonrev.h. Partials hold state-basis likelihoods; each
// child's table is its transition matrix P(r_c t) = Re(V exp(L r_c t) V^-1)
// of the category's mixture component (ModelMarkov::computeTransMatrixNonrev,
// model/modelmarkov.cpp:480-492), transposed for a child whose direction is
// toward the root, with the root leaf's table the component's frequencies
// (computePartialInfo's !useRevKernel branch, tree/phylokernelnew.h:934-986).
// Written in the reversible tables' layouts, these run through the same
// traversal kernels with the identity as the output transform: the products
// (P_l l) * (P_r r) of the reference's nonrev partial kernel. The branch
// likelihood and its derivatives apply P, QP and QQP between the two ends.
#pragma once
#include "plk_kernels.h"

namespace plkdev {

// child and end flags carried in is_tip
constexpr int kNrTip = 1, kNrTowardRoot = 2, kNrRootLeaf = 4;

struct NrDev {
    const double* ceval;   // [nmix][n] complex (re, im)
    const double* cevec;   // [nmix][n][n] complex
    const double* cinv;    // [nmix][n][n] complex
    const double* Q;       // [nmix][n][n]
    const double* pi;      // [nmix][n]
    const double* rate;    // [C] site rate of each category
    const double* wt;      // [C] getProp * getMixtureWeight
    const int* mix;        // [C] mixture component
    int n, C, rooted;
};

// exp(lambda_i len) of component m into ex (re, im pairs), then
// P = Re(V diag(ex) V^-1) into P; every thread of the block takes part
__device__ inline void nr_trans_block(const NrDev& D, int m, double len, double* ex, double* P) {
    const int n = D.n;
    const double* ev = D.ceval + (size_t)m * 2 * n;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const double a = exp(ev[2 * i] * len), b = ev[2 * i + 1] * len;
        ex[2 * i] = a * cos(b);
        ex[2 * i + 1] = a * sin(b);
    }
    __syncthreads();
    const double* V = D.cevec + (size_t)m * 2 * n * n;
    const double* W = D.cinv + (size_t)m * 2 * n * n;
    for (int t = threadIdx.x; t < n * n; t += blockDim.x) {
        const int x = t / n, y = t % n;
        double s = 0.0;
        for (int i = 0; i < n; i++) {
            const double vr = V[2 * (x * n + i)], vi = V[2 * (x * n + i) + 1];
            const double dr = vr * ex[2 * i] - vi * ex[2 * i + 1];
            const double di = vr * ex[2 * i + 1] + vi * ex[2 * i];
            s += dr * W[2 * (i * n + y)] - di * W[2 * (i * n + y) + 1];
        }
        P[t] = s;
    }
    __syncthreads();
}

// Per (op, child, category) table, in the layouts of k_prep_ops: an
// internal child's P (fragment order for 20 states), a tip child's lookup
// leaf[s][c][x] = sum_y P[x][y] tip[s][y] (leaf[SU] = 1), the root leaf's
// the frequencies for every s. LDS: ex[2n], P[n*n].
template <typename T>
__global__ void __launch_bounds__(kThreads)
k_nr_prep_ops(const plk_op* __restrict__ ops, const uint8_t* __restrict__ flags, int nops, NrDev D,
              const double* __restrict__ tip, int SU, int table_size, T* __restrict__ tables) {
    extern __shared__ double sh[];
    const int o = blockIdx.x >> 1, k = blockIdx.x & 1, c = blockIdx.y;
    if (o >= nops) return;
    const int n = D.n, C = D.C, m = D.mix[c];
    const plk_child ch = ops[o].child[k];
    const int fl = flags[2 * o + k];
    double* ex = sh;
    double* P = sh + 2 * n;
    nr_trans_block(D, m, D.rate[c] * ch.length, ex, P);
    const bool tr = (fl & kNrTowardRoot) != 0;
    T* out = tables + (size_t)(2 * o + k) * table_size;
    if (!ch.is_tip) {
        if (n == kProtN) {
            T* oc = out + (size_t)c * kProtFrag;
            for (int t = threadIdx.x; t < kProtFrag; t += blockDim.x) {
                int x, i;
                prot_frag_index(t, x, i);
                oc[t] = x < 0 ? (T)0 : (T)(tr ? P[i * n + x] : P[x * n + i]);
            }
            return;
        }
        T* oc = out + (size_t)c * n * n;
        for (int t = threadIdx.x; t < n * n; t += blockDim.x) {
            const int x = t / n, i = t % n;
            oc[t] = (T)(tr ? P[i * n + x] : P[x * n + i]);
        }
        return;
    }
    for (int t = threadIdx.x; t < (SU + 1) * n; t += blockDim.x) {
        const int x = t % n, s = t / n;
        double v;
        if (fl & kNrRootLeaf) {
            v = D.pi[(size_t)m * n + x];
        } else if (s == SU) {
            v = 1.0;
        } else {
            v = 0.0;
            for (int y = 0; y < n; y++) v += (tr ? P[y * n + x] : P[x * n + y]) * tip[s * n + y];
        }
        out[(size_t)(s * C + c) * n + x] = (T)v;
    }
}

// Root tables at `length`, one block per category (tree/phylokernelnonrev.h:
// 630-666, 1107-1131, 1146-1175): T0 = prop P, T1 = prop r Q P, T2 = prop r^2
// Q Q P (computeTransDerv, model/modelmarkov.cpp:735-766), rows times the
// frequencies when unrooted, into T[3][C][n][n]; for a tip dad the lookups
// lut[k][s][c][x] = sum_y Tk[x][y] tip[s][y], for the root leaf the
// frequencies times prop (no derivative). LDS: ex[2n], P, D1, D2 [n*n].
__global__ void __launch_bounds__(kThreads)
k_nr_prep_root(NrDev D, const double* __restrict__ tip, int SU, double length, int derv, int dad_tip,
               int dad_root, double* __restrict__ T, double* __restrict__ lut) {
    extern __shared__ double sh[];
    const int c = blockIdx.x, n = D.n, C = D.C, m = D.mix[c];
    const int nn = n * n;
    double* ex = sh;
    double* P = sh + 2 * n;
    double* D1 = P + nn;
    double* D2 = D1 + nn;
    const double r = D.rate[c], w = D.wt[c];
    nr_trans_block(D, m, r * length, ex, P);
    const double* Q = D.Q + (size_t)m * nn;
    if (derv) {
        for (int t = threadIdx.x; t < nn; t += blockDim.x) {
            const int x = t / n, y = t % n;
            double s = 0.0;
            for (int k = 0; k < n; k++) s += Q[x * n + k] * P[k * n + y];
            D1[t] = s;
        }
        __syncthreads();
        for (int t = threadIdx.x; t < nn; t += blockDim.x) {
            const int x = t / n, y = t % n;
            double s = 0.0;
            for (int k = 0; k < n; k++) s += Q[x * n + k] * D1[k * n + y];
            D2[t] = s;
        }
        __syncthreads();
    }
    // each thread rescales its own entries in place; the lookups below read
    // across threads after the barrier
    for (int t = threadIdx.x; t < nn; t += blockDim.x) {
        const double f = D.rooted ? 1.0 : D.pi[(size_t)m * n + t / n];
        const double t0 = P[t] * w * f;
        const double t1 = derv ? D1[t] * (w * r) * f : 0.0;
        const double t2 = derv ? D2[t] * (w * r * r) * f : 0.0;
        T[(size_t)c * nn + t] = t0;
        T[((size_t)C + c) * nn + t] = t1;
        T[((size_t)2 * C + c) * nn + t] = t2;
        P[t] = t0;
        D1[t] = t1;
        D2[t] = t2;
    }
    __syncthreads();
    if (!dad_tip) return;
    const size_t ls = (size_t)(SU + 1) * C * n;
    for (int t = threadIdx.x; t < (SU + 1) * n; t += blockDim.x) {
        const int x = t % n, s = t / n;
        const size_t o = ((size_t)s * C + c) * n + x;
        if (dad_root) {
            lut[o] = D.pi[(size_t)m * n + x] * w;
            lut[ls + o] = 0.0;
            lut[2 * ls + o] = 0.0;
            continue;
        }
        double v0 = 0.0, v1 = 0.0, v2 = 0.0;
        for (int y = 0; y < n; y++) {
            const double tv = tip[s * n + y];
            v0 += P[x * n + y] * tv;
            v1 += D1[x * n + y] * tv;
            v2 += D2[x * n + y] * tv;
        }
        lut[o] = v0;
        lut[ls + o] = v1;
        lut[2 * ls + o] = v2;
    }
}

struct NrRoot {
    const double* T;     // [3][C][n][n]
    const double* lut;   // [3][SU+1][C][n]
    int node_slot, dad_slot, dad_tip, dad_taxon, derv;
};

// The branch lnL (derv = 0) or df, ddf and lnL (derv = 1): sum_c sum_x
// N[c][x] sum_y Tk[c][x][y] D[c][y], one thread per (pattern, category), the
// C threads of a pattern adjacent in the wavefront and combined by lane
// shuffles. Safe scaling: the lnL takes per-category sums of both ends'
// counters against their minimum (tree/phylokernelnonrev.h:1422-1442), the
// derivatives each end against its own minimum (893-947). Block sums
// [grid][3]: lnL * freq, or df * freq, ddf terms, lnL * freq.
template <int C>
__device__ inline double group_sum(double v) {
#pragma unroll
    for (int m = 1; m < C; m <<= 1) v += __shfl_xor(v, m, C);
    return v;
}
template <int C>
__device__ inline int group_min(int v) {
#pragma unroll
    for (int m = 1; m < C; m <<= 1) v = min(v, __shfl_xor(v, m, C));
    return v;
}

template <typename T, int N, int C>
__global__ void __launch_bounds__(kThreads)
k_nr_root(InstDev I, NrRoot R, double* __restrict__ block_sums) {
    using S = typename Prec<T>::scale_t;
    __shared__ double sh[3 * (kThreads / 64)];
    const int tid = blockIdx.x * kThreads + threadIdx.x;
    const int p = tid / C, c = tid % C;
    const int nobs = I.P - I.nasc;
    const bool live = p < I.P;   // observed, or an unobserved pattern of +ASC
    double u0 = 0.0, u1 = 0.0, u2 = 0.0;
    int a = 0, b = 0;
    if (live) {
        const T* pn = reinterpret_cast<const T*>(I.partial) + R.node_slot * I.slot_stride + ((size_t)p * C + c) * N;
        if (R.dad_tip) {
            const int s = I.states[(size_t)R.dad_taxon * I.Ppad + p];
            const size_t ls = (size_t)(I.SU + 1) * C * N;
            const double* l0 = R.lut + ((size_t)s * C + c) * N;
#pragma unroll
            for (int x = 0; x < N; x++) {
                const double nv = (double)pn[x];
                u0 = fma(nv, l0[x], u0);
                if (R.derv) {
                    u1 = fma(nv, l0[ls + x], u1);
                    u2 = fma(nv, l0[2 * ls + x], u2);
                }
            }
        } else {
            const T* pd = reinterpret_cast<const T*>(I.partial) + R.dad_slot * I.slot_stride + ((size_t)p * C + c) * N;
            double dv[N];
#pragma unroll
            for (int y = 0; y < N; y++) dv[y] = (double)pd[y];
            const double* T0 = R.T + (size_t)c * N * N;
            const double* T1 = T0 + (size_t)C * N * N;
            const double* T2 = T1 + (size_t)C * N * N;
            for (int x = 0; x < N; x++) {
                double v0 = 0.0, v1 = 0.0, v2 = 0.0;
#pragma unroll
                for (int y = 0; y < N; y++) {
                    v0 = fma(T0[x * N + y], dv[y], v0);
                    if (R.derv) {
                        v1 = fma(T1[x * N + y], dv[y], v1);
                        v2 = fma(T2[x * N + y], dv[y], v2);
                    }
                }
                const double nv = (double)pn[x];
                u0 = fma(nv, v0, u0);
                u1 = fma(nv, v1, u1);
                u2 = fma(nv, v2, u2);
            }
        }
        if (I.safe) {
            a = reinterpret_cast<const S*>(I.scale)[R.node_slot * I.scale_stride + (size_t)p * C + c];
            b = R.dad_tip ? 0 : (int)reinterpret_cast<const S*>(I.scale)[R.dad_slot * I.scale_stride + (size_t)p * C + c];
        }
    }
    // every lane takes part in the shuffles
    int minsc, minsd;
    double fl = 1.0, fd = 1.0;
    if (I.safe) {
        minsc = group_min<C>(a + b);
        const int mn = group_min<C>(a), md = group_min<C>(b);
        minsd = mn + md;
        fl = Prec<T>::safe_adjust(a + b - minsc);
        fd = Prec<T>::safe_adjust(a - mn) * Prec<T>::safe_adjust(b - md);
    } else {
        minsc = 0;
        if (live) {
            const S* sn = reinterpret_cast<const S*>(I.scale) + R.node_slot * I.scale_stride;
            minsc = (int)sn[p];
            if (!R.dad_tip) minsc += (int)reinterpret_cast<const S*>(I.scale)[R.dad_slot * I.scale_stride + p];
        }
        minsd = minsc;
    }
    const double lh = group_sum<C>(u0 * fl);
    double l = 0.0, d1 = 0.0, d2 = 0.0;
    if (R.derv) {
        l = group_sum<C>(u0 * fd);
        d1 = group_sum<C>(u1 * fd);
        d2 = group_sum<C>(u2 * fd);
    }
    double acc[3] = {0.0, 0.0, 0.0};
    if (live && c == 0 && p >= nobs) {
        // an unobserved pattern (+ASC, tree/phylokernelnonrev.h:1000-1013,
        // 1455-1469): lh at asc[4k], and for derivatives the lh, d1, d2 the
        // derivative sums count at 4k+1..3, each rescaled once when scaled
        double* o = I.asc + 4 * (size_t)(p - nobs);
        o[0] = Prec<T>::asc_scale(lh + I.invar[p], minsc);
        if (R.derv) {
            o[1] = Prec<T>::asc_scale(l + I.invar[p], minsd);
            o[2] = Prec<T>::asc_scale(d1, minsd);
            o[3] = Prec<T>::asc_scale(d2, minsd);
        }
    } else if (live && c == 0) {
        const double lp = log(lh + I.invar[p]) + minsc * Prec<T>::log_unit;
        const double fq = I.freq[p];
        if (!R.derv) {
            I.ptn_lh[p] = lp;
            acc[0] = lp * fq;
        } else {
            const double inv = 1.0 / (l + I.invar[p]);
            const double df_frac = d1 * inv, ddf_frac = d2 * inv;
            const double tmp1 = df_frac * fq, tmp2 = ddf_frac * fq;
            acc[0] = tmp1;
            acc[1] = tmp2 - tmp1 * df_frac;
            acc[2] = lp * fq;
        }
    }
    if (p >= nobs && c == 0 && p < I.Ppad && !R.derv) I.ptn_lh[p] = 0.0;
    block_sum<3>(acc, sh);
    if (threadIdx.x == 0) {
        block_sums[(size_t)blockIdx.x * 3 + 0] = acc[0];
        block_sums[(size_t)blockIdx.x * 3 + 1] = acc[1];
        block_sums[(size_t)blockIdx.x * 3 + 2] = acc[2];
    }
}

}  // namespace plkdev
