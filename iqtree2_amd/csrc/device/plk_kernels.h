// HIP kernels of the likelihood engine (gfx950, wave64).
//
// Layout in HBM (one instance):
//   partial[slot][p][c][x]   T (double or float), pattern-major, eigen basis
//   scale[slot][p]           uint16 (FP64 form) / int32 exponent (FP32 form),
//                            [slot][p][c] in safe mode
//   theta[p][c][x]           T; buffer_scale[p] double
//   states[taxon][p]         uint8
// Thread mapping: one thread per (pattern, category); the C threads of a
// pattern are adjacent lanes of one wavefront, so a pattern's block of C*n
// values is read and written as one contiguous run, and cross-category
// reductions are lane shuffles. A workgroup owns a tile of 256/C patterns
// and walks the whole post-order op list for that tile: patterns are
// independent, so no synchronisation crosses workgroups.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "plk.h"

namespace plkdev {

constexpr int kThreads = 256;
constexpr double kScaleThreshold64 = 0x1p-256;          // tree/phylotree.h:74-80
constexpr int kScaleExp64 = 256;
constexpr double kLogScaleThreshold = -177.4456782233459932741;
constexpr float kScaleThreshold32 = 0x1p-40f;           // FP32-storage rescale threshold
constexpr double kLn2 = 0.6931471805599453094172321;

// Precision traits: storage type, scale counter type and the rescale rule.
template <typename T> struct Prec;
template <> struct Prec<double> {
    using scale_t = uint16_t;
    // one rescale by 2^256 per underflowing node (tree/phylokernelnew.h:1993-2030)
    __device__ static bool underflow(double m) { return m < kScaleThreshold64; }
    static constexpr double threshold = kScaleThreshold64;   // a block max at or above it never rescales
    __device__ static int exponent(double) { return 1; }
    __device__ static double apply(double v, int) { return ldexp(v, kScaleExp64); }
    // safe mode: a category scaled once more than the least-scaled one is
    // multiplied by 2^-256, more than once is dropped (tree/phylokernelnew.h:2891-2915)
    __device__ static double safe_adjust(int d) {
        return d == 0 ? 1.0 : (d == 1 ? kScaleThreshold64 : 0.0);
    }
    static constexpr double log_unit = kLogScaleThreshold;   // ln(2^-256)
    // an unobserved constant pattern (+ASC): a rescaled one counts once at the
    // scaling threshold, whatever its counter (tree/phylokernelnew.h:2940-2947)
    __device__ static double asc_scale(double v, int minsc) { return minsc != 0 ? v * kScaleThreshold64 : v; }
    __device__ static double asc_scale_log(double v, double bs) { return bs != 0.0 ? v * kScaleThreshold64 : v; }
};
template <> struct Prec<float> {
    using scale_t = int32_t;
    // exact power-of-two rescale that brings the block max into [0.5, 1)
    __device__ static bool underflow(float m) { return m < kScaleThreshold32 && m > 0.0f; }
    static constexpr float threshold = kScaleThreshold32;
    __device__ static int exponent(float m) { return -(ilogbf(m) + 1); }
    __device__ static float apply(float v, int k) { return ldexpf(v, k); }
    // counters are exponents, so categories differ by an exact power of two
    __device__ static double safe_adjust(int d) { return ldexp(1.0, -d); }
    static constexpr double log_unit = -kLn2;
    // counters are exponents: the exact scale back
    __device__ static double asc_scale(double v, int minsc) { return ldexp(v, -minsc); }
    __device__ static double asc_scale_log(double v, double bs) { return v * exp(bs); }
};

struct ModelDev {
    const double* eval;      // [n]
    const double* evec;      // [n][n]
    const double* inv_evec;  // [n][n]
    const double* tip;       // [SU+1][n]
    const double* rates;     // [C]
    const double* props;     // [C]
};

struct InstDev {
    int P, Ppad, n, C, SU, safe;
    const uint8_t* states;   // [ntaxa][Ppad]
    const double* freq;      // [Ppad]
    const double* invar;     // [Ppad]
    void* partial;           // [nslots][Ppad*block]
    void* scale;             // [nslots][Ppad*(safe?C:1)]
    void* theta;             // [Ppad*block]
    double* buffer_scale;    // [Ppad]
    double* ptn_lh;          // [Ppad]
    size_t slot_stride;      // Ppad*block
    size_t scale_stride;     // Ppad*(safe?C:1)
    // +ASC (Lewis): the last nasc of the P patterns are the unobserved
    // constant patterns; pattern k's lh goes to asc[4k], and a derivative
    // evaluation adds lh as the derivative sum counts it, d1 and d2 at 4k+1..3
    int nasc;
    double* asc;
};

// Transition tables of protein (20-state) internal children are stored in
// the operand order of the matrix-core tile (plk_mfma.h).
constexpr int kProtN = 20;                 // protein states
constexpr int kProtKSteps = 5;             // 4-wide K steps over 20 states
constexpr int kProtFrag = 2 * kProtKSteps * 64;   // fragment elements per category

// Elements of one (op, child) table for 20 states: an internal child in
// fragment order (C categories), a tip child's lookup [SU+1][C][20].
__host__ __device__ inline int prot_table_size(int C, int SU) {
    const int frag = C * kProtFrag, lut = (SU + 1) * C * kProtN;
    return frag > lut ? frag : lut;
}

// Fragment order of an internal child's table, per category:
// frag[xb][st][lane] = E[x = 16 xb + lane % 16][k = 5 (lane / 16) + st],
// zero for x >= 20. Returns the (x, k) of element t, x = -1 for padding.
__host__ __device__ inline void prot_frag_index(int t, int& x, int& k) {
    const int xb = t / (kProtKSteps * 64), st = (t / 64) % kProtKSteps, l = t & 63;
    x = 16 * xb + (l & 15);
    k = 5 * (l >> 4) + st;
    if (x >= kProtN) x = -1;
}

// A plain device copy of a known size (each launch reads and writes n*32
// bytes): the HBM bandwidth a streaming kernel reaches, and the reference
// size that calibrates the hardware byte counters.
__global__ void __launch_bounds__(kThreads)
k_copy(const double4* __restrict__ src, double4* __restrict__ dst, size_t n) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

enum TraverseMode { MODE_PARTIALS = 0, MODE_BRANCH = 1, MODE_THETA = 2 };
enum BufferMode { BUF_DERV = 0, BUF_LNL = 1 };

// ---------------------------------------------------------------------------
// computePartialInfo (tree/phylokernelnew.h:1008-1109): per op and child, the
// eigen-basis transition rows echild[c][x][i] = evec[x][i] exp(eval[i] r_c t)
// for an internal child, or the tip table leaf[s][c][x] =
// sum_i echild[c][x][i] tip[s][i] (leaf[SU] = 1) for a tip child.
// One workgroup per (op, child, category); tables are written in storage
// precision.
template <typename T>
__global__ void __launch_bounds__(kThreads)
k_prep_ops(const plk_op* __restrict__ ops, int nops, ModelDev m, int n, int C, int SU,
           int table_size, T* __restrict__ tables) {
    // one block per (op, child, category); LDS: ex[n], then for a tip child
    // U[n][n] and the tip vectors [SU][n]
    extern __shared__ double sh[];
    const int o = blockIdx.x >> 1, k = blockIdx.x & 1, c = blockIdx.y;
    if (o >= nops) return;
    const plk_child ch = ops[o].child[k];
    double* ex = sh;
    for (int i = threadIdx.x; i < n; i += blockDim.x) ex[i] = exp(m.eval[i] * (m.rates[c] * ch.length));
    T* out = tables + (size_t)(2 * o + k) * table_size;
    if (!ch.is_tip) {
        __syncthreads();
        if (n == kProtN) {
            // the matrix-core tile's fragment order
            T* oc = out + (size_t)c * kProtFrag;
            for (int t = threadIdx.x; t < kProtFrag; t += blockDim.x) {
                int x, i;
                prot_frag_index(t, x, i);
                oc[t] = x < 0 ? (T)0 : (T)(m.evec[x * n + i] * ex[i]);
            }
            return;
        }
        T* oc = out + (size_t)c * n * n;
        for (int t = threadIdx.x; t < n * n; t += blockDim.x) {
            const int i = t % n, x = t / n;
            oc[t] = (T)(m.evec[x * n + i] * ex[i]);
        }
        return;
    }
    double* ev_sh = ex + n;
    double* tp_sh = ev_sh + n * n;
    for (int t = threadIdx.x; t < n * n; t += blockDim.x) ev_sh[t] = m.evec[t];
    for (int t = threadIdx.x; t < SU * n; t += blockDim.x) tp_sh[t] = m.tip[t];
    __syncthreads();
    for (int t = threadIdx.x; t < (SU + 1) * n; t += blockDim.x) {
        const int x = t % n, s = t / n;
        double v;
        if (s == SU) {
            v = 1.0;
        } else {
            const double* ev = ev_sh + x * n;
            const double* tp = tp_sh + s * n;
            v = ev[0] * ex[0] * tp[0];
            for (int i = 1; i < n; i++) v += ev[i] * ex[i] * tp[i];
        }
        out[(size_t)(s * C + c) * n + x] = (T)v;
    }
}

// Root tables for the branch likelihood (tree/phylokernelnew.h:2713-2806):
// val[c][i] = exp(eval[i] r_c t) p_c, and for a tip dad
// lut[s][c][i] = val[c][i] * tip[s][i] for every s in 0..SU.
__global__ void __launch_bounds__(kThreads)
k_prep_root(ModelDev m, int n, int C, int SU, double length, int tip_dad, double* __restrict__ out) {
    const int nv = C * n;
    for (int t = threadIdx.x; t < nv; t += blockDim.x) {
        const int c = t / n, i = t % n;
        out[t] = exp(m.eval[i] * (m.rates[c] * length)) * m.props[c];
    }
    if (!tip_dad) return;
    const int total = (SU + 1) * nv;
    for (int t = threadIdx.x; t < total; t += blockDim.x) {
        const int i = t % n, c = (t / n) % C, s = t / nv;
        out[nv + t] = exp(m.eval[i] * (m.rates[c] * length)) * m.props[c] * m.tip[s * n + i];
    }
}

// ---------------------------------------------------------------------------
// A pattern's C category values sit in C adjacent lanes (C = 1, 2, 4, 8).
// lane_get<C, K>(v): the value of lane K of this lane's group. For C <= 4 the
// group lies inside one quad and a DPP quad permutation moves the value (no
// LDS round trip); 8 lanes use a bpermute.
template <int C, int K>
__device__ inline int lane_get_i32(int v) {
    if constexpr (C == 1) {
        return v;
    } else if constexpr (C <= 4) {
        constexpr int a = K, b = C == 4 ? K : 2 + K;   // C = 2: lanes {0,1} and {2,3}
        constexpr int ctrl = a | (a << 2) | (b << 4) | (b << 6);
        return __builtin_amdgcn_update_dpp(0, v, ctrl, 0xf, 0xf, false);
    } else {
        const int base = (threadIdx.x & 63) & ~(C - 1);
        return __shfl(v, base + K);
    }
}
template <int C, int K>
__device__ inline double lane_get(double v) {
    const long long x = __double_as_longlong(v);
    const int lo = lane_get_i32<C, K>((int)(unsigned)(x & 0xffffffffll));
    const int hi = lane_get_i32<C, K>((int)(x >> 32));
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
// max over the 16 lanes of this lane's DPP row (every lane gets it): quad
// swaps, then half-row and row mirrors pair each quad with the others
template <int CTRL>
__device__ inline float dpp_f32(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xf, 0xf, false));
}
template <int CTRL>
__device__ inline double dpp_f64(double v) {
    const long long x = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(x & 0xffffffffll), CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(x >> 32), CTRL, 0xf, 0xf, false);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
template <int CTRL>
__device__ inline float dpp_move(float v) { return dpp_f32<CTRL>(v); }
template <int CTRL>
__device__ inline double dpp_move(double v) { return dpp_f64<CTRL>(v); }
template <typename T>
__device__ inline T row16_max(T m) {
    m = fmax(m, dpp_move<0xB1>(m));    // quad_perm [1,0,3,2]
    m = fmax(m, dpp_move<0x4E>(m));    // quad_perm [2,3,0,1]
    m = fmax(m, dpp_move<0x141>(m));   // row_half_mirror
    m = fmax(m, dpp_move<0x140>(m));   // row_mirror
    return m;
}

// the C values of this lane's group, in category order
template <int C, int K = 0>
__device__ inline void lane_gather(double v, double* out) {
    if constexpr (K < C) {
        out[K] = lane_get<C, K>(v);
        lane_gather<C, K + 1>(v, out);
    }
}

template <int C>
__device__ inline double group_max(double v) {
    double vals[C];
    lane_gather<C>(v, vals);
    double m = vals[0];
#pragma unroll
    for (int c = 1; c < C; c++) m = fmax(m, vals[c]);
    return m;
}
template <int C, int K = 0>
__device__ inline int group_min_int(int v) {
    if constexpr (K + 1 < C) return min(lane_get_i32<C, K>(v), group_min_int<C, K + 1>(v));
    else return lane_get_i32<C, K>(v);
}
// sum over the C lanes of a pattern in category order 0..C-1 (the order the
// reference accumulates lh_cat), result in every lane of the group
template <int C>
__device__ inline double group_sum_ordered(double v) {
    double vals[C];
    lane_gather<C>(v, vals);
    double s = vals[0];
#pragma unroll
    for (int c = 1; c < C; c++) s += vals[c];
    return s;
}
// sumVec<VectorClass, true> grouping over categories (tree/phylokernelnew.h:43-71)
template <int C>
__device__ inline double group_sum_sumvec(double v) {
    double vals[C];
    lane_gather<C>(v, vals);
    double x = 0.0;
    constexpr int cruft = C & 3;
    if (cruft < C) {
        double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
#pragma unroll
        for (int c = 0; c < C - cruft; c += 4) {
            a0 += vals[c]; a1 += vals[c + 1]; a2 += vals[c + 2]; a3 += vals[c + 3];
        }
        a0 += a2;
        a1 += a3;
        a0 += a1;
        x += a0;
    }
#pragma unroll
    for (int c = C - cruft; c < C; c++) x += vals[c];
    return x;
}

// Sum over the 64 lanes of a wavefront with the same bits in every lane:
// DPP quad swaps and row mirrors (each step adds the same two operands in
// both partner lanes) give every lane its 16-lane row sum, then the four row
// sums are read as scalars and added in a fixed order.
__device__ inline double wave_sum(double v) {
    v += dpp_f64<0xB1>(v);    // quad_perm [1,0,3,2]
    v += dpp_f64<0x4E>(v);    // quad_perm [2,3,0,1]
    v += dpp_f64<0x141>(v);   // row_half_mirror
    v += dpp_f64<0x140>(v);   // row_mirror
    const long long x = __double_as_longlong(v);
    const int lo = (int)(unsigned)(x & 0xffffffffll), hi = (int)(x >> 32);
    double r[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const unsigned l = (unsigned)__builtin_amdgcn_readlane(lo, 16 * k);
        const unsigned h = (unsigned)__builtin_amdgcn_readlane(hi, 16 * k);
        r[k] = __longlong_as_double((long long)(((unsigned long long)h << 32) | l));
    }
    return (r[0] + r[1]) + (r[2] + r[3]);
}

// Deterministic workgroup sum: wave_sum inside each wavefront, then the
// wavefront partials in index order. Result valid in thread 0. GUARD = false
// skips the barrier that protects sh from a previous reader, for callers
// whose last use of sh already ended with a barrier.
template <int NV, bool GUARD = true>
__device__ inline void block_sum(double (&v)[NV], double* sh) {
#pragma unroll
    for (int j = 0; j < NV; j++) v[j] = wave_sum(v[j]);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (GUARD) __syncthreads();
    if (lane == 0)
#pragma unroll
        for (int j = 0; j < NV; j++) sh[wave * NV + j] = v[j];
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nw = blockDim.x >> 6;
#pragma unroll
        for (int j = 0; j < NV; j++) {
            double s = sh[j];
            for (int w = 1; w < nw; w++) s += sh[w * NV + j];
            v[j] = s;
        }
    }
}

// ---------------------------------------------------------------------------
// One binary traversal step for the pattern/category (p, c) of this thread
// (computePartialLikelihoodSIMD, tree/phylokernelnew.h:1632-2031):
// tmp[x] = prod_children (tip: leaf[s][c][x]; internal: sum_i E[c][x][i] child[i]),
// dad = U^-1 tmp, then the reference's rescaling rule. tL / tR are the
// children's tables (tip lookup or transition rows), a tip child goes left.
// Split into the loads (op_load) and the arithmetic and stores (op_finish)
// so that a caller can have the loads of several patterns in flight.
template <typename T, int N>
struct OpIn {
    T l[N], r[N];   // internal children's partials (unused for a tip child)
    int sl, sr;     // tip children's states
    int scl, scr;   // internal children's scale counters
};

template <typename T, int N, int C>
__device__ inline void op_load(const InstDev& I, int l_tip, int l_id, int r_tip, int r_id, int p, int c,
                               OpIn<T, N>& in) {
    using ST = typename Prec<T>::scale_t;
    constexpr int block = N * C;
    const T* partial = reinterpret_cast<const T*>(I.partial);
    const ST* scale = reinterpret_cast<const ST*>(I.scale);
    const size_t sidx = I.safe ? (size_t)p * C + c : (size_t)p;
    in.scl = in.scr = 0;
    if (l_tip) {
        in.sl = I.states[(size_t)l_id * I.Ppad + p];
    } else {
        const T* src = partial + (size_t)l_id * I.slot_stride + (size_t)p * block + c * N;
#pragma unroll
        for (int i = 0; i < N; i++) in.l[i] = src[i];
        in.scl = scale[(size_t)l_id * I.scale_stride + sidx];
    }
    if (r_tip) {
        in.sr = I.states[(size_t)r_id * I.Ppad + p];
    } else {
        const T* src = partial + (size_t)r_id * I.slot_stride + (size_t)p * block + c * N;
#pragma unroll
        for (int i = 0; i < N; i++) in.r[i] = src[i];
        in.scr = scale[(size_t)r_id * I.scale_stride + sidx];
    }
}

template <typename T, int N, int C>
__device__ inline void op_finish(const InstDev& I, const T* inv_sh, const T* tL, const T* tR, int l_tip,
                                 int r_tip, int dad_slot, int p, int c, const OpIn<T, N>& in) {
    using ST = typename Prec<T>::scale_t;
    constexpr int block = N * C;
    T* partial = reinterpret_cast<T*>(I.partial);
    ST* scale = reinterpret_cast<ST*>(I.scale);
    const bool safe = I.safe != 0;
    T tmp[N];
    const size_t sidx = safe ? (size_t)p * C + c : (size_t)p;
    if (l_tip) {
        const T* lut = tL + (in.sl * C + c) * N;
#pragma unroll
        for (int x = 0; x < N; x++) tmp[x] = lut[x];
    } else {
        const T* E = tL + c * N * N;
#pragma unroll
        for (int x = 0; x < N; x++) {
            T acc = E[x * N] * in.l[0];
#pragma unroll
            for (int i = 1; i < N; i++) acc = fma(E[x * N + i], in.l[i], acc);
            tmp[x] = acc;
        }
    }
    if (r_tip) {
        const T* lut = tR + (in.sr * C + c) * N;
#pragma unroll
        for (int x = 0; x < N; x++) tmp[x] *= lut[x];
    } else {
        const T* E = tR + c * N * N;
#pragma unroll
        for (int x = 0; x < N; x++) {
            T acc = E[x * N] * in.r[0];
#pragma unroll
            for (int i = 1; i < N; i++) acc = fma(E[x * N + i], in.r[i], acc);
            tmp[x] *= acc;
        }
    }
    // back to the eigen basis: dad[i] = sum_x inv_evec[i][x] tmp[x]
    T dadv[N];
    T lmax = (T)0;
#pragma unroll
    for (int i = 0; i < N; i++) {
        T acc = inv_sh[i * N] * tmp[0];
#pragma unroll
        for (int x = 1; x < N; x++) acc = fma(inv_sh[i * N + x], tmp[x], acc);
        dadv[i] = acc;
        lmax = fmax(lmax, fabs(acc));
    }
    int sc = in.scl + in.scr;
    if (!(l_tip && r_tip)) {   // tip-tip steps are never rescaled
        const bool no_invar = I.invar[p] == 0.0;
        const T m = safe ? lmax : (T)group_max<C>((double)lmax);
        if (Prec<T>::underflow(m) && no_invar) {
            const int k = Prec<T>::exponent(m);
#pragma unroll
            for (int i = 0; i < N; i++) dadv[i] = Prec<T>::apply(dadv[i], k);
            sc += k;
        }
    }
    T* dst = partial + (size_t)dad_slot * I.slot_stride + (size_t)p * block + c * N;
#pragma unroll
    for (int i = 0; i < N; i++) dst[i] = dadv[i];
    if (safe || c == 0) scale[(size_t)dad_slot * I.scale_stride + sidx] = (ST)sc;
}

template <typename T, int N, int C>
__device__ inline void op_apply(const InstDev& I, const T* inv_sh, const T* tL, const T* tR,
                                int l_tip, int l_id, int r_tip, int r_id, int dad_slot, int p,
                                int c) {
    OpIn<T, N> in;
    op_load<T, N, C>(I, l_tip, l_id, r_tip, r_id, p, c, in);
    op_finish<T, N, C>(I, inv_sh, tL, tR, l_tip, r_tip, dad_slot, p, c, in);
}

// Branch ends for (p, c): node partial nv, dad partial (or tip vector from
// tip_sh) dv, the summed scale counter and, in safe mode, the per-category
// adjustment relative to the least-scaled category.
template <typename T, int N, int C>
__device__ inline void load_branch(const InstDev& I, plk_end node, plk_end dad, const double* tip_sh,
                                   int p, int c, double* nv, double* dv, int& minsc,
                                   double& adjust) {
    using ST = typename Prec<T>::scale_t;
    constexpr int block = N * C;
    const T* partial = reinterpret_cast<const T*>(I.partial);
    const ST* scale = reinterpret_cast<const ST*>(I.scale);
    const bool safe = I.safe != 0;
    const size_t sidx = safe ? (size_t)p * C + c : (size_t)p;
    const T* nsrc = partial + (size_t)node.id * I.slot_stride + (size_t)p * block + c * N;
#pragma unroll
    for (int i = 0; i < N; i++) nv[i] = (double)nsrc[i];
    int sc = scale[(size_t)node.id * I.scale_stride + sidx];
    if (dad.is_tip) {
        if (tip_sh) {   // callers that use a val-weighted lookup pass nullptr
            const int s = I.states[(size_t)dad.id * I.Ppad + p];
            const double* tp = tip_sh + s * N;
#pragma unroll
            for (int i = 0; i < N; i++) dv[i] = tp[i];
        }
    } else {
        const T* dsrc = partial + (size_t)dad.id * I.slot_stride + (size_t)p * block + c * N;
#pragma unroll
        for (int i = 0; i < N; i++) dv[i] = (double)dsrc[i];
        sc += scale[(size_t)dad.id * I.scale_stride + sidx];
    }
    minsc = sc;
    adjust = 1.0;
    if (safe) {
        minsc = group_min_int<C>(sc);
        adjust = Prec<T>::safe_adjust(sc - minsc);
    }
}

// theta = node * dad (tip dad: its eigen-basis tip vector) and buffer_scale
// (computeLikelihoodBufferSIMD, tree/phylokernelnew.h:2091-2230)
template <typename T, int N, int C>
__device__ inline void theta_apply(const InstDev& I, plk_end node, plk_end dad, const double* tip_sh,
                                   int p, int c) {
    double nv[N], dv[N], adjust;
    int minsc;
    load_branch<T, N, C>(I, node, dad, tip_sh, p, c, nv, dv, minsc, adjust);
    T* th = reinterpret_cast<T*>(I.theta) + (size_t)p * (N * C) + c * N;
#pragma unroll
    for (int i = 0; i < N; i++) th[i] = (T)(nv[i] * dv[i] * adjust);
    if (c == 0) I.buffer_scale[p] = (double)minsc * Prec<T>::log_unit;
}

// Fused post-order traversal over the op list followed, per mode, by the
// branch lnL (computeLikelihoodBranchSIMD 2753-3104) or the theta buffer of
// the branch (node, dad).
template <typename T, int N, int C>
__global__ void __launch_bounds__(kThreads)
k_traverse(InstDev I, ModelDev m, const plk_op* __restrict__ ops, int nops,
           const T* __restrict__ tables, int table_size, int mode, plk_end node, plk_end dad,
           const double* __restrict__ root_tab, double* __restrict__ block_sums) {
    extern __shared__ __align__(16) unsigned char smem[];
    T* inv_sh = reinterpret_cast<T*>(smem);                 // [N*N]
    T* tab_sh = inv_sh + N * N;                              // [2][table_size]
    constexpr int PB = kThreads / C;
    constexpr int block = N * C;
    const int c = threadIdx.x % C;
    const int p = blockIdx.x * PB + threadIdx.x / C;
    const bool live = p < I.Ppad;

    for (int t = threadIdx.x; t < N * N; t += kThreads) inv_sh[t] = (T)m.inv_evec[t];

    for (int o = 0; o < nops; o++) {
        const plk_op op = ops[o];
        int left = 0, right = 1;
        if (!op.child[0].is_tip && op.child[1].is_tip) { left = 1; right = 0; }
        const plk_child L = op.child[left], R = op.child[right];
        const int szL = L.is_tip ? (I.SU + 1) * block : block * N;
        const int szR = R.is_tip ? (I.SU + 1) * block : block * N;
        __syncthreads();
        const T* gL = tables + (size_t)(2 * o + left) * table_size;
        const T* gR = tables + (size_t)(2 * o + right) * table_size;
        for (int t = threadIdx.x; t < szL; t += kThreads) tab_sh[t] = gL[t];
        for (int t = threadIdx.x; t < szR; t += kThreads) tab_sh[table_size + t] = gR[t];
        __syncthreads();
        if (live)
            op_apply<T, N, C>(I, inv_sh, tab_sh, tab_sh + table_size, L.is_tip, L.id, R.is_tip, R.id,
                              op.dad_slot, p, c);
    }

    if (mode == MODE_PARTIALS) return;

    // ---- root stage ---------------------------------------------------------
    // MODE_BRANCH: root_tab = val[c][i] then, for a tip dad, lut[s][c][i] = val * tip
    // MODE_THETA:  the eigen-basis tip vectors
    double* root_sh = reinterpret_cast<double*>(tab_sh);
    __syncthreads();
    if (mode == MODE_BRANCH) {
        const int root_size = block + (dad.is_tip ? (I.SU + 1) * block : 0);
        for (int t = threadIdx.x; t < root_size; t += kThreads) root_sh[t] = root_tab[t];
    } else if (dad.is_tip) {
        for (int t = threadIdx.x; t < (I.SU + 1) * N; t += kThreads) root_sh[t] = m.tip[t];
    }
    __syncthreads();
    double contrib = 0.0;
    if (live) {
        if (mode == MODE_THETA) {
            theta_apply<T, N, C>(I, node, dad, root_sh, p, c);
        } else {
            double nv[N], dv[N], adjust;
            int minsc;
            // a tip dad reads the val-weighted lookup lut[s][c][i] below instead of dv
            load_branch<T, N, C>(I, node, dad, nullptr, p, c, nv, dv, minsc, adjust);
            double lhc;
            if (dad.is_tip) {
                const int s = I.states[(size_t)dad.id * I.Ppad + p];
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
            const double lh0 = I.safe ? group_sum_sumvec<C>(lhc * adjust) : group_sum_ordered<C>(lhc);
            if (c == 0) {
                const double lh = fabs(lh0) + I.invar[p];
                const int pobs = I.P - I.nasc;
                if (p < pobs) {
                    const double lnp = log(lh) + (double)minsc * Prec<T>::log_unit;
                    I.ptn_lh[p] = lnp;
                    contrib = lnp * I.freq[p];
                } else if (p < I.P) {
                    I.asc[4 * (p - pobs)] = Prec<T>::asc_scale(lh, minsc);
                }
            }
        }
    }
    if (mode == MODE_BRANCH) {
        double v[1] = {contrib};
        block_sum<1>(v, reinterpret_cast<double*>(smem));
        if (threadIdx.x == 0) block_sums[blockIdx.x] = v[0];
    }
}

// ---------------------------------------------------------------------------
// Derivatives and lnL on the theta buffer (computeLikelihoodDervSIMD
// 2312-2551, computeLikelihoodFromBufferSIMD 3270-3343). Per pattern:
// lh = |sum_{c,i} val0 theta| + invar, d1 = sum val1 theta, d2 = sum val2 theta;
// df += f d1/lh, ddf += f (d2/lh - (d1/lh)^2), lnL += f (log lh + scale).
// One workgroup's share (patterns tile*PB ..) of the theta-buffer sums;
// out[0..2] = df, ddf and lnL partial sums (thread 0). val: LDS [3][N*C].
template <typename T, int N, int C>
__device__ inline void buffer_block(const InstDev& I, const ModelDev& m, double length, int mode,
                                    int tile, double* val, double* red, double* out) {
    constexpr int PB = kThreads / C;
    constexpr int block = N * C;
    for (int t = threadIdx.x; t < N * C; t += kThreads) {
        const int cc = t / N, i = t % N;
        const double cof = m.eval[i] * m.rates[cc];
        const double v0 = exp(cof * length) * m.props[cc];
        const double v1 = cof * v0;
        val[t] = v0;
        val[N * C + t] = v1;
        val[2 * N * C + t] = cof * v1;
    }
    __syncthreads();
    // With an odd state count and an even block, the reference's derivative
    // dot product (dotProductTriple, tree/phylokernelnew.h:668-716, called
    // with N = block at 2489-2491) adds the last (category, state) term twice
    // to lh, d1 and d2; the derivatives reproduce that for identical Newton
    // steps, while the lnL returned beside them uses the plain sum.
    constexpr bool kTwiceLast = (N & 1) && !((N * C) & 1);
    const int c = threadIdx.x % C;
    const int p = tile * PB + threadIdx.x / C;
    double lh = 0.0, d1 = 0.0, d2 = 0.0, lhq = 0.0;
    if (p < I.Ppad) {
        const T* th = reinterpret_cast<const T*>(I.theta) + (size_t)p * block + c * N;
        const double* v0 = val + c * N;
        const double* v1 = val + N * C + c * N;
        const double* v2 = val + 2 * N * C + c * N;
        constexpr int kUnroll = N <= 20 ? N : 4;   // 61 states: bounded register use
#pragma unroll kUnroll
        for (int i = 0; i < N; i++) {
            const double t = (double)th[i];
            lh = fma(v0[i], t, lh);
            if (mode == BUF_DERV) {
                d1 = fma(v1[i], t, d1);
                d2 = fma(v2[i], t, d2);
            }
        }
        lhq = lh;
        if (kTwiceLast && mode == BUF_DERV && c == C - 1) {
            const double t = (double)th[N - 1];
            lhq = fma(v0[N - 1], t, lh);
            d1 = fma(v1[N - 1], t, d1);
            d2 = fma(v2[N - 1], t, d2);
        }
    }
    lh = group_sum_ordered<C>(lh);
    if (mode == BUF_DERV) {
        d1 = group_sum_ordered<C>(d1);
        d2 = group_sum_ordered<C>(d2);
        lhq = kTwiceLast ? group_sum_ordered<C>(lhq) : lh;
    }
    double acc[3] = {0.0, 0.0, 0.0};
    const int pobs = I.P - I.nasc;
    if (c == 0 && p >= pobs && p < I.P) {
        // an unobserved constant pattern (+ASC): its probability and
        // derivatives, no weight of its own (tree/phylokernelnew.h:2497-2530, 3340-3358)
        const double bs = I.buffer_scale[p];
        double* a = I.asc + 4 * (p - pobs);
        a[0] = Prec<T>::asc_scale_log(fabs(lh) + I.invar[p], bs);
        if (mode == BUF_DERV) {
            a[1] = Prec<T>::asc_scale_log(fabs(lhq) + I.invar[p], bs);
            a[2] = Prec<T>::asc_scale_log(d1, bs);
            a[3] = Prec<T>::asc_scale_log(d2, bs);
        }
    }
    if (c == 0 && p < pobs) {
        lh = fabs(lh) + I.invar[p];
        const double f = I.freq[p];
        const double lnp = log(lh) + I.buffer_scale[p];
        if (mode == BUF_DERV) {
            const double inv = 1.0 / (kTwiceLast ? fabs(lhq) + I.invar[p] : lh);
            const double df_frac = d1 * inv, ddf_frac = d2 * inv;
            const double tmp1 = df_frac * f, tmp2 = ddf_frac * f;
            acc[0] = tmp1;
            acc[1] = tmp2 - tmp1 * df_frac;
            acc[2] = lnp * f;
        } else {
            I.ptn_lh[p] = lnp;
            acc[2] = lnp * f;
        }
    }
    block_sum<3>(acc, red);
    if (threadIdx.x == 0) {
        out[0] = acc[0];
        out[1] = acc[1];
        out[2] = acc[2];
    }
}

template <typename T, int N, int C>
__global__ void __launch_bounds__(kThreads)
k_buffer(InstDev I, ModelDev m, double length, int mode, double* __restrict__ block_sums) {
    __shared__ double val[3 * N * C];
    __shared__ double red[3 * (kThreads / 64)];
    buffer_block<T, N, C>(I, m, length, mode, blockIdx.x, val, red, block_sums + 3 * blockIdx.x);
}

// +ASC: pattern lnL of the observed patterns shifted by -log(1 - P(const)),
// the unobserved patterns' entries zeroed
__global__ void __launch_bounds__(kThreads)
k_asc_shift(double* __restrict__ ptn_lh, int pobs, int P, double corr) {
    const int p = blockIdx.x * kThreads + threadIdx.x;
    if (p < pobs) ptn_lh[p] -= corr;
    else if (p < P) ptn_lh[p] = 0.0;
}

// Fixed-order final reduction of per-workgroup partial sums: thread t adds
// entries t, t+256, ... in index order, then a fixed tree over threads.
__device__ inline void reduce_block(const double* __restrict__ in, int nblocks, int nvals,
                                    double* __restrict__ out, double* sh) {
    for (int j = 0; j < nvals; j++) {
        double s = 0.0;
        for (int b = threadIdx.x; b < nblocks; b += kThreads) s += in[(size_t)b * nvals + j];
        sh[threadIdx.x] = s;
        __syncthreads();
        for (int w = kThreads / 2; w >= 1; w >>= 1) {
            if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[j] = sh[0];
        __syncthreads();
    }
}

__global__ void __launch_bounds__(kThreads)
k_reduce(const double* __restrict__ in, int nblocks, int nvals, double* __restrict__ out) {
    __shared__ double sh[kThreads];
    reduce_block(in, nblocks, nvals, out, sh);
}

// RELL: out[b] = sum_p w[b][p] * ptn_lh[p], one workgroup per replicate.
__global__ void __launch_bounds__(kThreads)
k_rell(const float* __restrict__ w, const double* __restrict__ ptn_lh, int P, double* __restrict__ out) {
    __shared__ double sh[kThreads / 64];
    const float* wb = w + (size_t)blockIdx.x * P;
    double s = 0.0;
    for (int p = threadIdx.x; p < P; p += kThreads) s = fma((double)wb[p], ptn_lh[p], s);
    double v[1] = {s};
    block_sum<1>(v, sh);
    if (threadIdx.x == 0) out[blockIdx.x] = v[0];
}

}  // namespace plkdev
