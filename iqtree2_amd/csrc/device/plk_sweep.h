// Device-resident branch-length optimisation: one cooperative launch runs an
// optimizeAllBranches sweep (tree/phylotree.cpp:2698-2773) - for every branch
// in order, the stale partials (transition tables built in-kernel from the
// device copy of the branch lengths), theta, and minimizeNewton
// (utils/optimization.cpp:422-499) with its exact control flow.
//
// Every workgroup owns a fixed set of pattern tiles for the whole launch, so
// partial updates and theta only read values the same thread wrote; the only
// grid-wide exchange is one all-reduce per likelihood evaluation, with no
// flags: the partial sums themselves are polled.
//   - exchange e uses slot buffer e % 3; a slot holds a sentinel bit pattern
//     (a signalling NaN that no arithmetic produces) until its owner stores
//     its partial sum there;
//   - every workgroup polls the slots of all workgroups together until none
//     holds the sentinel and adds them in one fixed order, so all workgroups
//     hold the same bits and take identical Newton decisions;
//   - right after exchange e a workgroup restores the sentinel in its own
//     slots of buffer (e+2) % 3, last used by exchange e-1: every workgroup
//     has read e-1, since it published e only after reading e-1. Before its
//     exchange-(e+1) store a workgroup waits for its own earlier stores to be
//     acknowledged (s_waitcnt; by then long done), so anyone who sees its e+1
//     value also sees the restored sentinel of buffer (e+2) % 3.
// Values go through device-scope atomic loads and stores, coherent across the
// XCDs' L2 caches without writing caches back. (Measured alternatives that
// were slower: a master workgroup reducing and broadcasting; arrival flags
// next to the values, one round trip more per exchange.)
// Every spin is bounded: a workgroup that waits too long raises a shared
// abort flag, every spinning workgroup checks it, and all of them leave the
// kernel, which then reports an error.
//
// 61-state (codon) models: a step's transition tables (C x 61 x 61 per
// child) do not fit LDS, so the workgroups build them together into a
// double-buffered HBM area, meet at one grid-wide exchange, and run the
// step on their tiles with the codon matrix-core tile (plk_codon.h, one pass
// per rate category); theta stays in HBM, and each thread streams its
// (pattern, category)'s 61 states.
#pragma once
#include "plk_codon.h"

namespace plkdev {

constexpr long kSpinLimit = 1L << 24;   // ~0.5 s of polling before giving up
// pattern tiles whose branch-end loads a workgroup keeps in flight together
// in the theta pass (more did not pay for the partial updates)
template <int N, int C>
constexpr int kThetaTiles = N <= 4 ? 2 : 1;
// tiles of LDS-resident theta read together by a derivative evaluation
template <int N>
constexpr int kEvalTiles = N <= 4 ? 4 : 1;
constexpr int kFlagStride = 32;         // one 128-byte line per flag
constexpr unsigned long long kSlotEmpty = 0x7FF4DEAD7FF4DEADull;   // signalling NaN
constexpr unsigned kSlotEmptyWord = 0x7FF4DEADu;                   // its 32-bit halves

struct SweepArgs {
    InstDev I;
    ModelDev m;
    const plk_sweep_op* ops;
    const plk_sweep_branch* br;
    int nbranches;
    int nedges;
    double* len;          // [nedges], written by workgroup 0
    double min_len, max_len;
    int max_nr;
    int table_size;
    double* part;         // [3][gridDim.x][2] partial-sum slots (kSlotEmpty per launch)
    unsigned* release;    // [1] abort flag (zeroed per launch)
    int* nderv;           // [nbranches]
    double* br_len;       // [nbranches] optimised length of each branch
    double* lnl_out;      // [nlnl] lnL after branches with lnl_index >= 0
    int nlnl;
    double* out;          // [0] lnL from buffer after the sweep, [1] status
    long long* timing;    // nullable: wall-clock ticks of workgroup 0 per phase
    int lds_tiles;        // > 0: theta, buffer_scale, freq and invar of the workgroup's
                          // tiles (at most lds_tiles) live in LDS for the whole launch
    void* tabs;           // codon: [2][2][table_size] tables of a step, by step parity
};

// phases accumulated in SweepArgs::timing (workgroup 0, thread 0)
enum SweepPhase { PH_OPS = 0, PH_THETA, PH_EVAL_WORK, PH_EXCHANGE, PH_TOTAL, PH_COUNT };

__device__ inline unsigned load_flag(const unsigned* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void store_flag(unsigned* p, unsigned v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void store_value(double* p, double v) {
    __hip_atomic_store(reinterpret_cast<unsigned long long*>(p),
                       (unsigned long long)__double_as_longlong(v), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline double load_value(const double* p) {
    return __longlong_as_double((long long)__hip_atomic_load(
        reinterpret_cast<const unsigned long long*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}
// every store this thread issued has been acknowledged
__device__ inline void stores_done() { __builtin_amdgcn_s_waitcnt(0); }

__device__ inline unsigned long long load_bits(const double* p) {
    return __hip_atomic_load(reinterpret_cast<const unsigned long long*>(p), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void store_bits(double* p, unsigned long long v) {
    __hip_atomic_store(reinterpret_cast<unsigned long long*>(p), v, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}

// Bounded wait until the n (<= 4) slots hold values, all polled together;
// false on timeout or abort.
__device__ inline bool wait_slots(const double* const* slot, int n, double* out, unsigned* abort_flag) {
    for (long it = 0; it < kSpinLimit; it++) {
        bool all = true;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (k < n) {
                const unsigned long long b = load_bits(slot[k]);
                out[k] = __longlong_as_double((long long)b);
                all = all && b != kSlotEmpty;
            }
        }
        if (all) return true;
        if ((it & 31) == 31 && __hip_atomic_load(abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            return false;
        __builtin_amdgcn_s_sleep(1);
    }
    __hip_atomic_store(abort_flag, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return false;
}

// Fixed-order sum over the workgroup's 256 partial sums; res[0..1] in LDS.
// wave_sum gives every lane of a wave the same bits, the four wave sums are
// added in wave order.
__device__ inline void tree_sum2(double s0, double s1, double* red_sh, double* res) {
    s0 = wave_sum(s0);
    s1 = wave_sum(s1);
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red_sh[2 * wave] = s0;
        red_sh[2 * wave + 1] = s1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double r0 = red_sh[0], r1 = red_sh[1];
        for (int w = 1; w < kThreads / 64; w++) {
            r0 += red_sh[2 * w];
            r1 += red_sh[2 * w + 1];
        }
        res[0] = r0;
        res[1] = r1;
    }
    __syncthreads();
}

// Grid-wide fixed-order sum of two per-workgroup values (valid in thread 0
// on entry); every workgroup computes the same res[0..1]. Returns false when
// the exchange was aborted.
__device__ inline bool grid_sum2(const SweepArgs& a, double v0, double v1, unsigned epoch,
                                 double* red_sh, double* res) {
    __shared__ int ok_sh;
    const int G = gridDim.x;
    double* buf = a.part + (size_t)(epoch % 3) * 2 * G;
    if (threadIdx.x == 0) {
        stores_done();   // the sentinel restored after the previous exchange is visible
        store_value(&buf[2 * blockIdx.x + 0], v0);
        store_value(&buf[2 * blockIdx.x + 1], v1);
        ok_sh = 1;
    }
    __syncthreads();
    unsigned* abort_flag = &a.release[1];
    // this thread's workgroups b = tid, tid + 256 (gridDim.x <= 512)
    const int b0 = threadIdx.x, b1 = threadIdx.x + kThreads;
    const double* slot[4] = {&buf[2 * b0], &buf[2 * b0 + 1], &buf[2 * b1], &buf[2 * b1 + 1]};
    const int n = b1 < G ? 4 : (b0 < G ? 2 : 0);
    double x[4] = {0.0, 0.0, 0.0, 0.0};
    const bool ok = wait_slots(slot, n, x, abort_flag);
    double s0 = 0.0, s1 = 0.0;
    if (ok) {
        if (n >= 2) { s0 += x[0]; s1 += x[1]; }
        if (n == 4) { s0 += x[2]; s1 += x[3]; }
    }
    if (!ok) ok_sh = 0;
    tree_sum2(s0, s1, red_sh, res);
    if (threadIdx.x == 0) {
        // every workgroup has read exchange e-1: re-arm this workgroup's
        // slots of its buffer for exchange e+2
        double* again = a.part + (size_t)((epoch + 2) % 3) * 2 * G;
        store_bits(&again[2 * blockIdx.x + 0], kSlotEmpty);
        store_bits(&again[2 * blockIdx.x + 1], kSlotEmpty);
    }
    return ok_sh != 0;
}

// LDS layout shared by the host (size) and the kernel (offsets)
struct SweepSmem {
    int tip, ex, val, red, lens, ctl, inv, tab, th, bs, fq, iv, total;
    __host__ __device__ SweepSmem(int N, int C, int SU, int table_size, int nedges, int tsize,
                                  int lds_tiles) {
        const int pats = lds_tiles * (kThreads / C);   // patterns resident per workgroup
        int off = 0;
        if (N == kCodonN) {
            // codon: the codon tile's E and U^-1 (row stride kCodonStride),
            // no tip or table copies (the tables are in HBM) and no resident theta
            tip = ex = bs = fq = iv = th = 0;
            val = off;  off += 3 * C * N * 8;
            red = off;  off += 2 * kThreads * 8;
            lens = off; off += nedges * 8;
            ctl = off;  off += 8 * 8;
            inv = off;  off += (N * kCodonStride * tsize + 15) / 16 * 16;
            tab = off;  off += (N * kCodonStride * tsize + 15) / 16 * 16;
            total = (off + 15) / 16 * 16;
            return;
        }
        tip = off;  off += (SU + 1) * N * 8;
        ex = off;   off += 2 * C * N * 8;
        val = off;  off += 3 * C * N * 8;
        red = off;  off += 2 * kThreads * 8;
        lens = off; off += nedges * 8;
        ctl = off;  off += 8 * 8;
        inv = off;  off += (N * N * tsize + 15) / 16 * 16;
        tab = off;  off += (2 * table_size * tsize + 15) / 16 * 16;
        bs = off;   off += pats * 8;
        fq = off;   off += pats * 8;
        iv = off;   off += pats * 8;
        th = off;   off += pats * N * C * tsize;
        total = (off + 15) / 16 * 16;
    }
};

// Where a workgroup's per-pattern state lives: LDS copies for its own tiles
// (resident pattern pl = local tile * PB + pattern in tile) when on, the
// instance's global arrays otherwise. The pointers always point into the
// workgroup's LDS (so every access through them is an LDS access) and are
// only used when on. LDS theta is state-major, th[i][local tile][thread],
// so a wave's reads of one state are contiguous.
template <typename T>
using Lds = __attribute__((address_space(3))) T;   // LDS-qualified (ds_* accesses only)

template <typename T>
struct SweepRes {
    bool on;
    Lds<T>* th;
    Lds<double>* bs;
    Lds<double>* fq;
    Lds<double>* iv;
    int stride;     // elements between states: resident tiles * kThreads
};

// Transition tables of both children of an op, built by the whole workgroup
// with the same arithmetic as k_prep_ops.
template <typename T, int N, int C>
__device__ inline void build_tables(const ModelDev& m, int SU, int table_size, const int* is_tip,
                                    const double* len, T* tab_sh, double* ex_sh) {
    for (int t = threadIdx.x; t < 2 * C * N; t += kThreads) {
        const int k = t / (C * N), r = t % (C * N), c = r / N, i = r % N;
        ex_sh[t] = exp(m.eval[i] * (m.rates[c] * len[k]));
    }
    __syncthreads();
    for (int k = 0; k < 2; k++) {
        T* out = tab_sh + k * table_size;
        const double* ex = ex_sh + k * C * N;
        if (!is_tip[k]) {
            for (int t = threadIdx.x; t < C * N * N; t += kThreads) {
                const int i = t % N, x = (t / N) % N, c = t / (N * N);
                out[t] = (T)(m.evec[x * N + i] * ex[c * N + i]);
            }
        } else {
            for (int t = threadIdx.x; t < (SU + 1) * C * N; t += kThreads) {
                const int x = t % N, c = (t / N) % C, s = t / (N * C);
                double v;
                if (s == SU) {
                    v = 1.0;
                } else {
                    const double* ev = m.evec + x * N;
                    const double* tp = m.tip + s * N;
                    const double* e = ex + c * N;
                    v = ev[0] * e[0] * tp[0];
                    for (int i = 1; i < N; i++) v += ev[i] * e[i] * tp[i];
                }
                out[t] = (T)v;
            }
        }
    }
    __syncthreads();
}

// Grid-wide evaluation on theta. mode 0: (df, ddf) at x; mode 1: lnL at x and
// at x2 (x2 < 0: only x), writing pattern lnL for x when store_ptn. Results
// land in res[0..1] of every workgroup.
template <typename T, int N, int C>
__device__ inline bool sweep_eval(const SweepArgs& a, const SweepRes<T>& R, int mode, double x, double x2,
                                  bool store_ptn, unsigned& epoch, double* val_sh, double* red_sh,
                                  double* res) {
    const long long t0 = wall_clock64();
    constexpr int PB = kThreads / C;
    constexpr int block = N * C;
    const InstDev& I = a.I;
    for (int t = threadIdx.x; t < N * C; t += kThreads) {
        const int cc = t / N, i = t % N;
        const double cof = a.m.eval[i] * a.m.rates[cc];
        if (mode == 0) {
            const double v0 = exp(cof * x) * a.m.props[cc];
            const double v1 = cof * v0;
            val_sh[t] = v0;
            val_sh[N * C + t] = v1;
            val_sh[2 * N * C + t] = cof * v1;
        } else {
            val_sh[t] = exp(cof * x) * a.m.props[cc];
            val_sh[N * C + t] = x2 >= 0.0 ? exp(cof * x2) * a.m.props[cc] : 0.0;
        }
    }
    __syncthreads();
    const int c = threadIdx.x % C;
    const int ntiles = I.Ppad / PB;
    double acc[2] = {0.0, 0.0};
    if constexpr (N <= 4 && C <= 4) {
        if (R.on) {
            // one thread per resident pattern: the thread adds its pattern's
            // category sums itself, in category order (as group_sum_ordered)
            const int mine = blockIdx.x < (unsigned)ntiles ? (ntiles - 1 - (int)blockIdx.x) / (int)gridDim.x + 1 : 0;
            // the evaluation's coefficients in registers when they are few
            constexpr bool kRegW = C * N <= 16;
            constexpr int NW = kRegW ? C * N : 1;
            double w[3][NW];
            if constexpr (kRegW) {
#pragma unroll
                for (int k = 0; k < NW; k++) {
                    w[0][k] = val_sh[k];
                    w[1][k] = val_sh[N * C + k];
                    w[2][k] = mode == 0 ? val_sh[2 * N * C + k] : 0.0;
                }
            }
            for (int pl = threadIdx.x; pl < mine * PB; pl += kThreads) {
                const int lt = pl / PB, q = pl - lt * PB;
                const int p = ((int)blockIdx.x + lt * (int)gridDim.x) * PB + q;
                const Lds<T>* th = R.th + lt * kThreads + q * C;
                double S0 = 0.0, S1 = 0.0, S2 = 0.0;
#pragma unroll
                for (int cc = 0; cc < C; cc++) {
                    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
#pragma unroll
                    for (int i = 0; i < N; i++) {
                        const double t = (double)th[i * R.stride + cc];
                        const int k = cc * N + i;
                        s0 = fma(kRegW ? w[0][kRegW ? k : 0] : val_sh[k], t, s0);
                        s1 = fma(kRegW ? w[1][kRegW ? k : 0] : val_sh[N * C + k], t, s1);
                        if (mode == 0) s2 = fma(kRegW ? w[2][kRegW ? k : 0] : val_sh[2 * N * C + k], t, s2);
                    }
                    if (cc == 0) { S0 = s0; S1 = s1; S2 = s2; }
                    else { S0 += s0; S1 += s1; S2 += s2; }
                }
                if (p < I.P) {
                    const double f = R.fq[pl], pinv = R.iv[pl];
                    if (mode == 0) {
                        const double lh = fabs(S0) + pinv;
                        const double inv = 1.0 / lh;
                        const double df_frac = S1 * inv, ddf_frac = S2 * inv;
                        const double tmp1 = df_frac * f, tmp2 = ddf_frac * f;
                        acc[0] += tmp1;
                        acc[1] += tmp2 - tmp1 * df_frac;
                    } else {
                        const double bsc = R.bs[pl];
                        const double lnp = log(fabs(S0) + pinv) + bsc;
                        if (store_ptn) I.ptn_lh[p] = lnp;
                        acc[0] += lnp * f;
                        if (x2 >= 0.0) acc[1] += (log(fabs(S1) + pinv) + bsc) * f;
                    }
                }
            }
            goto reduce;
        }
    }
    {
    // this thread's category coefficients in registers (small state spaces)
    constexpr bool kHoist = N <= 4;
    double cf[3][kHoist ? N : 1];
    if constexpr (kHoist) {
#pragma unroll
        for (int i = 0; i < N; i++) {
            cf[0][i] = val_sh[c * N + i];
            cf[1][i] = val_sh[N * C + c * N + i];
            cf[2][i] = mode == 0 ? val_sh[2 * N * C + c * N + i] : 0.0;
        }
    }
    // the theta reads of TM tiles issued together, then each tile in order
    constexpr int TM = kEvalTiles<N>;
    for (int t0 = blockIdx.x, lt0 = 0; t0 < ntiles; t0 += TM * gridDim.x, lt0 += TM) {
    T tvs[TM][N];
#pragma unroll
    for (int m = 0; m < TM; m++) {
        const int tile = t0 + m * gridDim.x;
        if (tile >= ntiles) break;
        if (R.on) {
            const Lds<T>* th = R.th + (lt0 + m) * kThreads + threadIdx.x;
#pragma unroll
            for (int i = 0; i < N; i++) tvs[m][i] = th[i * R.stride];
        } else {
            const T* th = reinterpret_cast<const T*>(I.theta) + (size_t)(tile * PB + threadIdx.x / C) * block + c * N;
#pragma unroll
            for (int i = 0; i < N; i++) tvs[m][i] = th[i];
        }
    }
#pragma unroll
    for (int m = 0; m < TM; m++) {
        const int tile = t0 + m * gridDim.x, lt = lt0 + m;
        if (tile >= ntiles) break;
        const int p = tile * PB + threadIdx.x / C;
        const int pl = lt * PB + threadIdx.x / C;
        double s0 = 0.0, s1 = 0.0, s2 = 0.0;
#pragma unroll
        for (int i = 0; i < N; i++) {
            const double t = (double)tvs[m][i];
            const double w0 = kHoist ? cf[0][kHoist ? i : 0] : val_sh[c * N + i];
            const double w1 = kHoist ? cf[1][kHoist ? i : 0] : val_sh[N * C + c * N + i];
            s0 = fma(w0, t, s0);
            s1 = fma(w1, t, s1);
            if (mode == 0) {
                const double w2 = kHoist ? cf[2][kHoist ? i : 0] : val_sh[2 * N * C + c * N + i];
                s2 = fma(w2, t, s2);
            }
        }
        s0 = group_sum_ordered<C>(s0);
        s1 = group_sum_ordered<C>(s1);
        if (mode == 0) s2 = group_sum_ordered<C>(s2);
        if (c == 0 && p < I.P) {
            const double f = R.on ? R.fq[pl] : I.freq[p];
            const double pinv = R.on ? R.iv[pl] : I.invar[p];
            const double bsc = mode == 0 ? 0.0 : (R.on ? R.bs[pl] : I.buffer_scale[p]);
            if (mode == 0) {
                const double lh = fabs(s0) + pinv;
                const double inv = 1.0 / lh;
                const double df_frac = s1 * inv, ddf_frac = s2 * inv;
                const double tmp1 = df_frac * f, tmp2 = ddf_frac * f;
                acc[0] += tmp1;
                acc[1] += tmp2 - tmp1 * df_frac;
            } else {
                const double lnp = log(fabs(s0) + pinv) + bsc;
                if (store_ptn) I.ptn_lh[p] = lnp;
                acc[0] += lnp * f;
                if (x2 >= 0.0) acc[1] += (log(fabs(s1) + pinv) + bsc) * f;
            }
        }
    }
    }
    }
reduce:
    block_sum<2, false>(acc, red_sh);   // red_sh: last read before tree_sum2's final barrier
    epoch++;
    const long long t1 = wall_clock64();
    const bool ok = grid_sum2(a, acc[0], acc[1], epoch, red_sh, res);
    if (a.timing && blockIdx.x == 0 && threadIdx.x == 0) {
        const long long t2 = wall_clock64();
        a.timing[PH_EVAL_WORK] += t1 - t0;
        a.timing[PH_EXCHANGE] += t2 - t1;
    }
    return ok;
}

// theta of branch (node, dad) and, from the same registers, the first
// Newton evaluation (df, ddf) at x: the values summed are the stored theta
// values, so the result equals a separate theta pass followed by sweep_eval
template <typename T, int N, int C>
__device__ inline bool theta_first_eval(const SweepArgs& a, const SweepRes<T>& R, bool write_global,
                                        plk_end node, plk_end dad, double x, unsigned& epoch,
                                        const double* tip_sh, double* val_sh, double* red_sh,
                                        double* res) {
    const long long t0 = wall_clock64();
    constexpr int PB = kThreads / C;
    constexpr int block = N * C;
    const InstDev& I = a.I;
    for (int t = threadIdx.x; t < N * C; t += kThreads) {
        const int cc = t / N, i = t % N;
        const double cof = a.m.eval[i] * a.m.rates[cc];
        const double v0 = exp(cof * x) * a.m.props[cc];
        const double v1 = cof * v0;
        val_sh[t] = v0;
        val_sh[N * C + t] = v1;
        val_sh[2 * N * C + t] = cof * v1;
    }
    __syncthreads();
    const int c = threadIdx.x % C;
    const int ntiles = I.Ppad / PB;
    double acc[2] = {0.0, 0.0};
    const bool to_global = !R.on || write_global;
    using ST = typename Prec<T>::scale_t;
    const T* partial = reinterpret_cast<const T*>(I.partial);
    const ST* scale = reinterpret_cast<const ST*>(I.scale);
    constexpr int TM = kThetaTiles<N, C>;
    for (int t0 = blockIdx.x, lt0 = 0; t0 < ntiles; t0 += TM * gridDim.x, lt0 += TM) {
    // the loads of TM tiles in flight first (load_branch's reads)
    T nr[TM][N], dr[TM][N];
    int scr[TM], sr[TM];
#pragma unroll
    for (int m = 0; m < TM; m++) {
        const int tile = t0 + m * gridDim.x;
        if (tile >= ntiles) break;
        const int p = tile * PB + threadIdx.x / C;
        const size_t sidx = I.safe ? (size_t)p * C + c : (size_t)p;
        const T* nsrc = partial + (size_t)node.id * I.slot_stride + (size_t)p * block + c * N;
#pragma unroll
        for (int i = 0; i < N; i++) nr[m][i] = nsrc[i];
        scr[m] = scale[(size_t)node.id * I.scale_stride + sidx];
        if (dad.is_tip) {
            sr[m] = I.states[(size_t)dad.id * I.Ppad + p];
        } else {
            const T* dsrc = partial + (size_t)dad.id * I.slot_stride + (size_t)p * block + c * N;
#pragma unroll
            for (int i = 0; i < N; i++) dr[m][i] = dsrc[i];
            scr[m] += scale[(size_t)dad.id * I.scale_stride + sidx];
        }
    }
#pragma unroll
    for (int m = 0; m < TM; m++) {
        const int tile = t0 + m * gridDim.x, lt = lt0 + m;
        if (tile >= ntiles) break;
        const int p = tile * PB + threadIdx.x / C;
        const int pl = lt * PB + threadIdx.x / C;
        double nv[N], dv[N], adjust = 1.0;
        int minsc = scr[m];
#pragma unroll
        for (int i = 0; i < N; i++) {
            nv[i] = (double)nr[m][i];
            dv[i] = dad.is_tip ? tip_sh[sr[m] * N + i] : (double)dr[m][i];
        }
        if (I.safe) {
            minsc = group_min_int<C>(scr[m]);
            adjust = Prec<T>::safe_adjust(scr[m] - minsc);
        }
        T* thg = reinterpret_cast<T*>(I.theta) + (size_t)p * block + c * N;
        Lds<T>* thl = R.th + lt * kThreads + threadIdx.x;
        const double* v0 = val_sh + c * N;
        const double* v1 = val_sh + N * C + c * N;
        const double* v2 = val_sh + 2 * N * C + c * N;
        double s0 = 0.0, s1 = 0.0, s2 = 0.0;
#pragma unroll
        for (int i = 0; i < N; i++) {
            const T tv = (T)(nv[i] * dv[i] * adjust);
            if (R.on) thl[i * R.stride] = tv;
            if (to_global) thg[i] = tv;
            const double t = (double)tv;
            s0 = fma(v0[i], t, s0);
            s1 = fma(v1[i], t, s1);
            s2 = fma(v2[i], t, s2);
        }
        if (c == 0) {
            const double bsc = (double)minsc * Prec<T>::log_unit;
            if (R.on) R.bs[pl] = bsc;
            if (to_global) I.buffer_scale[p] = bsc;
        }
        s0 = group_sum_ordered<C>(s0);
        s1 = group_sum_ordered<C>(s1);
        s2 = group_sum_ordered<C>(s2);
        if (c == 0 && p < I.P) {
            const double f = R.on ? R.fq[pl] : I.freq[p];
            const double lh = fabs(s0) + (R.on ? R.iv[pl] : I.invar[p]);
            const double inv = 1.0 / lh;
            const double df_frac = s1 * inv, ddf_frac = s2 * inv;
            const double tmp1 = df_frac * f, tmp2 = ddf_frac * f;
            acc[0] += tmp1;
            acc[1] += tmp2 - tmp1 * df_frac;
        }
    }
    }
    block_sum<2, false>(acc, red_sh);   // red_sh: last read before tree_sum2's final barrier
    epoch++;
    const long long t1 = wall_clock64();
    const bool ok = grid_sum2(a, acc[0], acc[1], epoch, red_sh, res);
    if (a.timing && blockIdx.x == 0 && threadIdx.x == 0) {
        const long long t2 = wall_clock64();
        a.timing[PH_THETA] += t1 - t0;
        a.timing[PH_EXCHANGE] += t2 - t1;
    }
    return ok;
}

// ---- 61 states ----

// 61 states and an even block (C even): the reference's derivative sums count
// the last (category, state) term twice (buffer_block); so does the sweep's
// derivative, whose lh then only serves the derivative ratios
template <int C>
constexpr bool kCodonTwiceLast = ((kCodonN * C) & 1) == 0;

// Both children's tables of sweep op `op` into tb (k_prep_ops' arithmetic),
// the element range split over the whole grid.
template <typename T, int C>
__device__ inline void codon_sweep_tables(const SweepArgs& a, const plk_sweep_op& op, const double* lens,
                                          T* tb) {
    constexpr int N = kCodonN;
    const int SU = a.I.SU, ts = a.table_size;
    const size_t stride = (size_t)gridDim.x * kThreads;
    for (size_t t = (size_t)blockIdx.x * kThreads + threadIdx.x; t < 2 * (size_t)ts; t += stride) {
        const int k = (int)(t / ts), r = (int)(t % ts);
        const double len = lens[op.child[k].edge];
        T* out = tb + (size_t)k * ts;
        if (!op.child[k].is_tip) {
            if (r >= C * N * N) continue;
            const int i = r % N, x = (r / N) % N, c = r / (N * N);
            out[r] = (T)(a.m.evec[x * N + i] * exp(a.m.eval[i] * (a.m.rates[c] * len)));
        } else {
            if (r >= (SU + 1) * C * N) continue;
            const int x = r % N, c = (r / N) % C, s = r / (N * C);
            double v;
            if (s == SU) {
                v = 1.0;
            } else {
                const double* ev = a.m.evec + x * N;
                const double* tp = a.m.tip + s * N;
                v = ev[0] * exp(a.m.eval[0] * (a.m.rates[c] * len)) * tp[0];
                for (int i = 1; i < N; i++) v += ev[i] * exp(a.m.eval[i] * (a.m.rates[c] * len)) * tp[i];
            }
            out[r] = (T)v;
        }
    }
}

// theta of branch (node, dad) into HBM and the first (df, ddf) at x from the
// stored values: one thread per (pattern, category), the 61 states streamed
template <typename T, int C>
__device__ inline bool codon_theta_first_eval(const SweepArgs& a, plk_end node, plk_end dad, double x,
                                              unsigned& epoch, double* val_sh, double* red_sh, double* res) {
    constexpr int N = kCodonN, PB = kThreads / C, block = N * C;
    const InstDev& I = a.I;
    for (int t = threadIdx.x; t < N * C; t += kThreads) {
        const int cc = t / N, i = t % N;
        const double cof = a.m.eval[i] * a.m.rates[cc];
        const double v0 = exp(cof * x) * a.m.props[cc];
        const double v1 = cof * v0;
        val_sh[t] = v0;
        val_sh[N * C + t] = v1;
        val_sh[2 * N * C + t] = cof * v1;
    }
    __syncthreads();
    using ST = typename Prec<T>::scale_t;
    const T* partial = reinterpret_cast<const T*>(I.partial);
    const ST* scale = reinterpret_cast<const ST*>(I.scale);
    const int c = threadIdx.x % C;
    const int ntiles = I.Ppad / PB;
    double acc[2] = {0.0, 0.0};
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int p = tile * PB + threadIdx.x / C;
        const size_t sidx = I.safe ? (size_t)p * C + c : (size_t)p;
        const T* nv = partial + (size_t)node.id * I.slot_stride + (size_t)p * block + c * N;
        int sc = scale[(size_t)node.id * I.scale_stride + sidx];
        const T* dv = nullptr;
        const double* tp = nullptr;
        if (dad.is_tip) {
            tp = a.m.tip + (size_t)I.states[(size_t)dad.id * I.Ppad + p] * N;
        } else {
            dv = partial + (size_t)dad.id * I.slot_stride + (size_t)p * block + c * N;
            sc += scale[(size_t)dad.id * I.scale_stride + sidx];
        }
        int minsc = sc;
        double adjust = 1.0;
        if (I.safe) {
            minsc = group_min_int<C>(sc);
            adjust = Prec<T>::safe_adjust(sc - minsc);
        }
        T* th = reinterpret_cast<T*>(I.theta) + (size_t)p * block + c * N;
        const double* v0 = val_sh + c * N;
        const double* v1 = val_sh + N * C + c * N;
        const double* v2 = val_sh + 2 * N * C + c * N;
        double s0 = 0.0, s1 = 0.0, s2 = 0.0;
        T last = (T)0;
        for (int i = 0; i < N; i++) {
            const double d = dad.is_tip ? tp[i] : (double)dv[i];
            const T tv = (T)((double)nv[i] * d * adjust);
            th[i] = tv;
            last = tv;
            const double t = (double)tv;
            s0 = fma(v0[i], t, s0);
            s1 = fma(v1[i], t, s1);
            s2 = fma(v2[i], t, s2);
        }
        // the reference's derivative counts the last (category, state) term
        // twice when the block is even (dotProductTriple; see buffer_block)
        if (kCodonTwiceLast<C> && c == C - 1) {
            const double t = (double)last;
            s0 = fma(v0[N - 1], t, s0);
            s1 = fma(v1[N - 1], t, s1);
            s2 = fma(v2[N - 1], t, s2);
        }
        if (c == 0) I.buffer_scale[p] = (double)minsc * Prec<T>::log_unit;
        s0 = group_sum_ordered<C>(s0);
        s1 = group_sum_ordered<C>(s1);
        s2 = group_sum_ordered<C>(s2);
        if (c == 0 && p < I.P) {
            const double f = I.freq[p];
            const double lh = fabs(s0) + I.invar[p];
            const double inv = 1.0 / lh;
            const double df_frac = s1 * inv, ddf_frac = s2 * inv;
            const double tmp1 = df_frac * f, tmp2 = ddf_frac * f;
            acc[0] += tmp1;
            acc[1] += tmp2 - tmp1 * df_frac;
        }
    }
    block_sum<2, false>(acc, red_sh);
    epoch++;
    return grid_sum2(a, acc[0], acc[1], epoch, red_sh, res);
}

// sweep_eval for 61 states: theta read from HBM, the states streamed
template <typename T, int C>
__device__ inline bool codon_sweep_eval(const SweepArgs& a, int mode, double x, double x2, bool store_ptn,
                                        unsigned& epoch, double* val_sh, double* red_sh, double* res) {
    constexpr int N = kCodonN, PB = kThreads / C, block = N * C;
    const InstDev& I = a.I;
    for (int t = threadIdx.x; t < N * C; t += kThreads) {
        const int cc = t / N, i = t % N;
        const double cof = a.m.eval[i] * a.m.rates[cc];
        if (mode == 0) {
            const double v0 = exp(cof * x) * a.m.props[cc];
            const double v1 = cof * v0;
            val_sh[t] = v0;
            val_sh[N * C + t] = v1;
            val_sh[2 * N * C + t] = cof * v1;
        } else {
            val_sh[t] = exp(cof * x) * a.m.props[cc];
            val_sh[N * C + t] = x2 >= 0.0 ? exp(cof * x2) * a.m.props[cc] : 0.0;
        }
    }
    __syncthreads();
    const int c = threadIdx.x % C;
    const int ntiles = I.Ppad / PB;
    double acc[2] = {0.0, 0.0};
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int p = tile * PB + threadIdx.x / C;
        const T* th = reinterpret_cast<const T*>(I.theta) + (size_t)p * block + c * N;
        const double* w0 = val_sh + c * N;
        const double* w1 = val_sh + N * C + c * N;
        const double* w2 = val_sh + 2 * N * C + c * N;
        double s0 = 0.0, s1 = 0.0, s2 = 0.0;
        for (int i = 0; i < N; i++) {
            const double t = (double)th[i];
            s0 = fma(w0[i], t, s0);
            s1 = fma(w1[i], t, s1);
            if (mode == 0) s2 = fma(w2[i], t, s2);
        }
        if (mode == 0 && kCodonTwiceLast<C> && c == C - 1) {   // as in codon_theta_first_eval
            const double t = (double)th[N - 1];
            s0 = fma(w0[N - 1], t, s0);
            s1 = fma(w1[N - 1], t, s1);
            s2 = fma(w2[N - 1], t, s2);
        }
        s0 = group_sum_ordered<C>(s0);
        s1 = group_sum_ordered<C>(s1);
        if (mode == 0) s2 = group_sum_ordered<C>(s2);
        if (c == 0 && p < I.P) {
            const double f = I.freq[p], pinv = I.invar[p];
            if (mode == 0) {
                const double lh = fabs(s0) + pinv;
                const double inv = 1.0 / lh;
                const double df_frac = s1 * inv, ddf_frac = s2 * inv;
                const double tmp1 = df_frac * f, tmp2 = ddf_frac * f;
                acc[0] += tmp1;
                acc[1] += tmp2 - tmp1 * df_frac;
            } else {
                const double bsc = I.buffer_scale[p];
                const double lnp = log(fabs(s0) + pinv) + bsc;
                if (store_ptn) I.ptn_lh[p] = lnp;
                acc[0] += lnp * f;
                if (x2 >= 0.0) acc[1] += (log(fabs(s1) + pinv) + bsc) * f;
            }
        }
    }
    block_sum<2, false>(acc, red_sh);
    epoch++;
    return grid_sum2(a, acc[0], acc[1], epoch, red_sh, res);
}

// the state count's theta / evaluation forms
template <typename T, int N, int C>
__device__ inline bool any_first_eval(const SweepArgs& a, const SweepRes<T>& R, bool write_global, plk_end node,
                                      plk_end dad, double x, unsigned& epoch, const double* tip_sh, double* val_sh,
                                      double* red_sh, double* res) {
    if constexpr (N == kCodonN)
        return codon_theta_first_eval<T, C>(a, node, dad, x, epoch, val_sh, red_sh, res);
    else
        return theta_first_eval<T, N, C>(a, R, write_global, node, dad, x, epoch, tip_sh, val_sh, red_sh, res);
}

template <typename T, int N, int C>
__device__ inline bool any_eval(const SweepArgs& a, const SweepRes<T>& R, int mode, double x, double x2,
                                bool store_ptn, unsigned& epoch, double* val_sh, double* red_sh, double* res) {
    if constexpr (N == kCodonN)
        return codon_sweep_eval<T, C>(a, mode, x, x2, store_ptn, epoch, val_sh, red_sh, res);
    else
        return sweep_eval<T, N, C>(a, R, mode, x, x2, store_ptn, epoch, val_sh, red_sh, res);
}

template <typename T, int N, int C>
__global__ void __launch_bounds__(kThreads) k_sweep(SweepArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    const InstDev& I = a.I;
    const SweepSmem L(N, C, I.SU, a.table_size, a.nedges, (int)sizeof(T), a.lds_tiles);
    double* tip_sh = reinterpret_cast<double*>(smem + L.tip);
    double* ex_sh = reinterpret_cast<double*>(smem + L.ex);
    double* val_sh = reinterpret_cast<double*>(smem + L.val);
    double* red_sh = reinterpret_cast<double*>(smem + L.red);
    double* lens = reinterpret_cast<double*>(smem + L.lens);
    double* ctl = reinterpret_cast<double*>(smem + L.ctl);
    T* inv_sh = reinterpret_cast<T*>(smem + L.inv);
    T* tab_sh = reinterpret_cast<T*>(smem + L.tab);
    constexpr int PB = kThreads / C;
    const int c = threadIdx.x % C;
    const int ntiles = I.Ppad / PB;

    if constexpr (N == kCodonN) {
        // the codon tile's U^-1 layout (row i, column x, stride kCodonStride)
        for (int t = threadIdx.x; t < N * N; t += kThreads)
            inv_sh[(t / N) * kCodonStride + t % N] = (T)a.m.inv_evec[t];
    } else {
        for (int t = threadIdx.x; t < N * N; t += kThreads) inv_sh[t] = (T)a.m.inv_evec[t];
        for (int t = threadIdx.x; t < (I.SU + 1) * N; t += kThreads) tip_sh[t] = a.m.tip[t];
    }
    for (int t = threadIdx.x; t < a.nedges; t += kThreads) lens[t] = a.len[t];
    SweepRes<T> R{a.lds_tiles > 0, (Lds<T>*)(smem + L.th), (Lds<double>*)(smem + L.bs),
                  (Lds<double>*)(smem + L.fq), (Lds<double>*)(smem + L.iv), a.lds_tiles * kThreads};
    if (R.on) {
        // weights and +I mass of the resident patterns, constant for the launch
        for (int tile = blockIdx.x, lt = 0; tile < ntiles; tile += gridDim.x, lt++)
            for (int t = threadIdx.x; t < PB; t += kThreads) {
                R.fq[lt * PB + t] = I.freq[tile * PB + t];
                R.iv[lt * PB + t] = I.invar[tile * PB + t];
            }
    }
    __syncthreads();

    unsigned epoch = 0;
    int status = 0;
    const bool timed = a.timing && blockIdx.x == 0 && threadIdx.x == 0;
    const long long t_start = wall_clock64();
    const double x1 = a.min_len, x2 = a.max_len, xacc = a.min_len;
    for (int j = 0; j < a.nbranches && status == 0; j++) {
        const plk_sweep_branch B = a.br[j];
        long long tp = wall_clock64();
        // stale partials of this branch
        if constexpr (N == kCodonN) {
            // the step's tables built by the whole grid into the HBM buffer of
            // the step's parity (the other buffer may still be read by a
            // workgroup finishing the previous step), published with a
            // device-scope fence before the exchange that every workgroup
            // passes, and acquired after it; then the codon tile on the
            // workgroup's own patterns, 64 at a time, one pass per category
            __shared__ plk_op op_sh;
            for (int o = B.op_begin; o < B.op_begin + B.op_count && status == 0; o++) {
                const plk_sweep_op op = a.ops[o];
                T* tb = reinterpret_cast<T*>(a.tabs) + (size_t)(o & 1) * 2 * a.table_size;
                codon_sweep_tables<T, C>(a, op, lens, tb);
                if (threadIdx.x == 0) {
                    op_sh.dad_slot = op.dad_slot;
                    op_sh.nchildren = 2;
                    for (int k = 0; k < 2; k++) {
                        op_sh.child[k].is_tip = op.child[k].is_tip;
                        op_sh.child[k].id = op.child[k].id;
                        op_sh.child[k].length = lens[op.child[k].edge];
                    }
                }
                __threadfence();
                __syncthreads();
                epoch++;
                if (!grid_sum2(a, 0.0, 0.0, epoch, red_sh, ctl)) {
                    status = PLK_ERR_DEVICE;
                    break;
                }
                __threadfence();
                for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x)
                    for (int q = 0; q < PB / kCodonTile; q++)
                        for (int cc = 0; cc < C; cc++)
                            codon_ops_tile<T>(I, &op_sh, 1, tb, a.table_size, cc, tile * PB + q * kCodonTile,
                                              tab_sh, inv_sh);
            }
            if (status != 0) break;
        } else {
            for (int o = B.op_begin; o < B.op_begin + B.op_count; o++) {
                const plk_sweep_op op = a.ops[o];
                int left = 0, right = 1;
                if (!op.child[0].is_tip && op.child[1].is_tip) { left = 1; right = 0; }
                const int tips[2] = {op.child[left].is_tip, op.child[right].is_tip};
                const double ln[2] = {lens[op.child[left].edge], lens[op.child[right].edge]};
                build_tables<T, N, C>(a.m, I.SU, a.table_size, tips, ln, tab_sh, ex_sh);
                for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
                    const int p = tile * PB + threadIdx.x / C;
                    op_apply<T, N, C>(I, inv_sh, tab_sh, tab_sh + a.table_size, tips[0],
                                      op.child[left].id, tips[1], op.child[right].id, op.dad_slot, p, c);
                }
                __syncthreads();
            }
        }
        if (timed) a.timing[PH_OPS] += wall_clock64() - tp;
        // minimizeNewton(min_len, len, max_len, xacc = min_len, max_nr) on
        // f = -dlnL/dt (PhyloTree::computeFuncDerv); every thread holds the
        // same scalars. The first evaluation comes with theta.
        const double len0 = lens[B.edge];
        int nd = 0;
        double f, df, dx, dxold, xl, xh, rts, rts_old, result = 0.0;
        bool done = false;
        rts = len0;
        if (rts < x1) rts = x1;
        if (rts > x2) rts = x2;
        // the instance keeps the last branch's theta (plk_likelihood_derv
        // after the sweep); earlier ones live in LDS only
        const bool last_branch = j == a.nbranches - 1;
        if (!any_first_eval<T, N, C>(a, R, last_branch, B.node, B.dad, rts, epoch, tip_sh, val_sh,
                                       red_sh, ctl)) {
            status = PLK_ERR_DEVICE;
            break;
        }
        nd++;
        double xlast = rts;   // length the last derivative evaluation used
        f = -ctl[0];
        df = -ctl[1];
        // a non-finite derivative: the safe kernel continues with df = ddf = 0
        // (tree/phylokernelnew.h:2629-2632), the normal kernel stops
        if (!isfinite(f) && I.safe) f = df = -0.0;
        if (!isfinite(f) || !isfinite(df)) {
            status = PLK_ERR_NONFINITE;
            break;
        }
        if (df >= 0.0 && fabs(f) < xacc) {
            result = rts;
            done = true;
        }
        if (!done) {
            if (f < 0.0) { xl = rts; xh = x2; } else { xh = rts; xl = x1; }
            dx = dxold = fabs(xh - xl);
            for (int it = 1; it <= a.max_nr; it++) {
                rts_old = rts;
                if ((df <= 0.0) || (((rts - xh) * df - f) * ((rts - xl) * df - f) >= 0.0)) {
                    dxold = dx;
                    dx = 0.5 * (xh - xl);
                    rts = xl + dx;
                    if (xl == rts) { result = rts; done = true; break; }
                } else {
                    dxold = dx;
                    dx = f / df;
                    const double temp = rts;
                    rts -= dx;
                    if (temp == rts) { result = rts; done = true; break; }
                }
                if (fabs(dx) < xacc || it == a.max_nr) { result = rts_old; done = true; break; }
                if (!any_eval<T, N, C>(a, R, 0, rts, -1.0, false, epoch, val_sh, red_sh, ctl)) {
                    status = PLK_ERR_DEVICE;
                    break;
                }
                nd++;
                xlast = rts;
                f = -ctl[0];
                df = -ctl[1];
                if (!isfinite(f) && I.safe) f = df = -0.0;
                if (!isfinite(f) || !isfinite(df)) { status = PLK_ERR_NONFINITE; break; }
                if (df > 0.0 && fabs(f) < xacc) { result = rts; done = true; break; }
                if (f < 0.0) xl = rts;
                else if (f > 0.0) xh = rts;
            }
            (void)dxold;
            if (status != 0) break;
            if (!done) { status = PLK_ERR_NEWTON; break; }
        }
        double optx = result;
        // divergence guard of optimizeOneBranch (tree/phylotree.cpp:2616-2625):
        // lnL at the last evaluated length against lnL at the old one
        if (optx > x2 * 0.95) {
            if (!any_eval<T, N, C>(a, R, 1, xlast, len0, false, epoch, val_sh, red_sh, ctl)) {
                status = PLK_ERR_DEVICE;
                break;
            }
            if (ctl[1] > ctl[0]) optx = len0;
        }
        __syncthreads();
        if (threadIdx.x == 0) lens[B.edge] = optx;
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            a.len[B.edge] = optx;
            a.nderv[j] = nd;
            a.br_len[j] = optx;
        }
        __syncthreads();
        // computeLikelihoodFromBuffer at the new length (current_it->length)
        if (B.lnl_index >= 0 && B.lnl_index < a.nlnl) {
            if (!any_eval<T, N, C>(a, R, 1, optx, -1.0, false, epoch, val_sh, red_sh, ctl)) {
                status = PLK_ERR_DEVICE;
                break;
            }
            if (!isfinite(ctl[0])) {
                status = PLK_ERR_NONFINITE;
                break;
            }
            if (blockIdx.x == 0 && threadIdx.x == 0) a.lnl_out[B.lnl_index] = ctl[0];
        }
    }
    // computeLikelihoodFromBuffer at the last branch
    if (status == 0 && a.nbranches > 0) {
        const double last = lens[a.br[a.nbranches - 1].edge];
        if (!any_eval<T, N, C>(a, R, 1, last, -1.0, true, epoch, val_sh, red_sh, ctl))
            status = PLK_ERR_DEVICE;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        a.out[0] = status == 0 ? ctl[0] : 0.0;
        a.out[1] = (double)status;
        if (a.timing) a.timing[PH_TOTAL] += wall_clock64() - t_start;
    }
}

}  // namespace plkdev
