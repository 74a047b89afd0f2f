/*
 * oracle.c - FP64 CPU restatement of IQ-TREE 2's reversible likelihood
 * kernels. TEST INFRASTRUCTURE ONLY (parity oracle and CPU baseline); the
 * product library never links or calls this file.
 *
 * Every function cites the reference code it restates. Arithmetic follows the
 * reference helpers' grouping (tree/phylokernelnew.h:81-796) on 4-wide
 * pattern vectors, so results agree with the reference FMA kernel to a few
 * ulps; exp/log come from libm instead of VCL 1.30.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double v4d __attribute__((vector_size(32), aligned(32)));
#define VS 4
#define SCALING_THRESHOLD 0x1p-256           /* tree/phylotree.h:74-80 */
#define SCALING_THRESHOLD_EXP 256
#define LOG_SCALING_THRESHOLD (-177.4456782233459932741)
#define INLINE static inline __attribute__((always_inline))

struct orc_tree {
    int ntaxa, n, C, P, Ppad, nslots, SU, safe, nthreads, npackets;
    size_t block;
    uint8_t* states;       /* [ntaxa][Ppad], padding = SU */
    double *freq, *invar;  /* [Ppad] */
    double *eval, *evec, *inv_evec, *tip, *rates, *props;
    double* partial;       /* [nslots][Ppad*block] interleaved */
    uint16_t* scale;       /* [nslots][Ppad*(safe?C:1)] */
    double* theta;         /* [Ppad*block] */
    double* buffer_scale;  /* [Ppad] */
    double* pattern_lh;    /* [Ppad] */
    int theta_ready;
    int nasc;              /* +ASC (Lewis): the last nasc patterns are the unobserved constant ones */
    double nsite;          /* sites of the observed patterns (aln->getNSite()) */
    double last_x;         /* length of the last derivative evaluation (Newton) */
    double* info;          /* per-op echild + leaf tables */
    size_t info_cap;
    double* tipbuf;        /* per-packet tip staging [npackets][2*block*VS] */
    uint64_t* pars_set;    /* Fitch state sets [nslots][Ppad] */
    uint16_t* pars_changes;/* changes below each slot [nslots][Ppad] */
    /* bit-plane form (computePartialParsimonyFast): sites expanded by
     * frequency, blocks of 256 sites, [block][state][8] words */
    int pf_nblocks;        /* 0: not built */
    uint32_t* pf_tip;      /* [ntaxa][nblocks][n][8] */
    uint32_t* pf_node;     /* [nslots][nblocks][n][8] */
    int64_t* pf_score;     /* [nslots] subtree score */
    /* non-reversible and mixture models in the state basis
     * (tree/phylokernelnonrev.h): per mixture component the complex
     * eigensystem (re, im pairs), the rate matrix and the frequencies */
    int nonrev, rooted, nmix;
    double *nr_ceval, *nr_cevec, *nr_cinv, *nr_Q, *nr_pi;
    int* nr_mix;           /* [C] mixture component of each category */
    uint8_t* nr_flags;     /* [2*nops] child flags of the staged ops */
    int nr_flags_cap;
    orc_end nr_node, nr_dad;
    /* Sankoff parsimony: cost matrix [n][n], tip costs [SU+1][n], partial
     * costs [nslots][Ppad][n] */
    uint32_t *sk_cost, *sk_tip, *sk_partial;
};

static void* amalloc(size_t bytes) {
    void* p = NULL;
    if (bytes == 0) bytes = 64;
    if (posix_memalign(&p, 64, (bytes + 63) & ~(size_t)63) != 0) return NULL;
    memset(p, 0, bytes);
    return p;
}

INLINE v4d vset1(double a) { return (v4d){a, a, a, a}; }
INLINE v4d vabs(v4d a) {
    v4d r;
    for (int k = 0; k < VS; k++) r[k] = fabs(a[k]);
    return r;
}
INLINE v4d vmax(v4d a, v4d b) {
    v4d r;
    for (int k = 0; k < VS; k++) r[k] = a[k] > b[k] ? a[k] : b[k];
    return r;
}
INLINE double hadd(v4d a) { return (a[0] + a[1]) + (a[2] + a[3]); }

/* ---- helpers restating tree/phylokernelnew.h:81-796 ----------------- */

/* dotProductVec, JAMES_VERSION (81-126): A scalar row, B pattern vectors */
INLINE v4d dot_vec_s(const double* A, const v4d* B, const int N) {
    if (N == 4) return (vset1(A[0]) * B[0] + vset1(A[1]) * B[1]) + (vset1(A[2]) * B[2] + vset1(A[3]) * B[3]);
    const int cruft = N & 3;
    if (N == cruft) {
        v4d X = vset1(0.0);
        for (int i = 0; i < cruft; i++) X = vset1(A[i]) * B[i] + X;
        return X;
    }
    v4d X = vset1(A[0]) * B[0], v1 = vset1(A[1]) * B[1], v2 = vset1(A[2]) * B[2],
        v3 = vset1(A[3]) * B[3];
    int i = 4;
    for (; i < N - cruft; i += 4) {
        X = vset1(A[i]) * B[i] + X;
        v1 = vset1(A[i + 1]) * B[i + 1] + v1;
        v2 = vset1(A[i + 2]) * B[i + 2] + v2;
        v3 = vset1(A[i + 3]) * B[i + 3] + v3;
    }
    if (cruft > 0) {
        X = vset1(A[i]) * B[i] + X;
        if (cruft > 1) {
            v1 = vset1(A[i + 1]) * B[i + 1] + v1;
            if (cruft > 2) v2 = vset1(A[i + 2]) * B[i + 2] + v2;
        }
    }
    return (X + v1) + (v2 + v3);
}

/* same with both operands pattern vectors */
INLINE v4d dot_vec_v(const v4d* A, const v4d* B, const int N) {
    if (N == 4) return (A[0] * B[0] + A[1] * B[1]) + (A[2] * B[2] + A[3] * B[3]);
    const int cruft = N & 3;
    if (N == cruft) {
        v4d X = vset1(0.0);
        for (int i = 0; i < cruft; i++) X = A[i] * B[i] + X;
        return X;
    }
    v4d X = A[0] * B[0], v1 = A[1] * B[1], v2 = A[2] * B[2], v3 = A[3] * B[3];
    int i = 4;
    for (; i < N - cruft; i += 4) {
        X = A[i] * B[i] + X;
        v1 = A[i + 1] * B[i + 1] + v1;
        v2 = A[i + 2] * B[i + 2] + v2;
        v3 = A[i + 3] * B[i + 3] + v3;
    }
    if (cruft > 0) {
        X = A[i] * B[i] + X;
        if (cruft > 1) {
            v1 = A[i + 1] * B[i + 1] + v1;
            if (cruft > 2) v2 = A[i + 2] * B[i + 2] + v2;
        }
    }
    return (X + v1) + (v2 + v3);
}

/* dotProductDualVec (214-314): (A.B) * (C.D), A and C scalar rows */
INLINE v4d dot_dual(const double* A, const v4d* B, const double* Cm, const v4d* D, const int N) {
    if (N == 4) {
        return ((vset1(A[1]) * B[1] + vset1(A[0]) * B[0]) + (vset1(A[3]) * B[3] + vset1(A[2]) * B[2])) *
               ((vset1(Cm[1]) * D[1] + vset1(Cm[0]) * D[0]) + (vset1(Cm[3]) * D[3] + vset1(Cm[2]) * D[2]));
    }
    if (N < 4) {
        v4d ab = vset1(A[0]) * B[0], cd = vset1(Cm[0]) * D[0];
        for (int i = 1; i < N; i++) {
            ab = vset1(A[i]) * B[i] + ab;
            cd = vset1(Cm[i]) * D[i] + cd;
        }
        return ab * cd;
    }
    switch (N % 4) {
    case 0: {
        v4d AB[4], CD[4];
        for (int j = 0; j < 4; j++) { AB[j] = vset1(A[j]) * B[j]; CD[j] = vset1(Cm[j]) * D[j]; }
        for (int i = 4; i < N; i += 4)
            for (int j = 0; j < 4; j++) {
                AB[j] = vset1(A[i + j]) * B[i + j] + AB[j];
                CD[j] = vset1(Cm[i + j]) * D[i + j] + CD[j];
            }
        return ((AB[0] + AB[1]) + (AB[2] + AB[3])) * (((CD[0] + CD[1]) + CD[2]) + CD[3]);
    }
    case 1: {
        const int Nm1 = N - 1;
        v4d AB[4], CD[4];
        for (int j = 0; j < 4; j++) { AB[j] = vset1(A[j]) * B[j]; CD[j] = vset1(Cm[j]) * D[j]; }
        for (int i = 4; i < Nm1; i += 4)
            for (int j = 0; j < 4; j++) {
                AB[j] = vset1(A[i + j]) * B[i + j] + AB[j];
                CD[j] = vset1(Cm[i + j]) * D[i + j] + CD[j];
            }
        return (vset1(A[Nm1]) * B[Nm1] + ((AB[0] + AB[1]) + (AB[2] + AB[3]))) *
               (vset1(Cm[Nm1]) * D[Nm1] + (((CD[0] + CD[1]) + CD[2]) + CD[3]));
    }
    case 2: {
        v4d AB[2], CD[2];
        for (int j = 0; j < 2; j++) { AB[j] = vset1(A[j]) * B[j]; CD[j] = vset1(Cm[j]) * D[j]; }
        for (int i = 2; i < N; i += 2)
            for (int j = 0; j < 2; j++) {
                AB[j] = vset1(A[i + j]) * B[i + j] + AB[j];
                CD[j] = vset1(Cm[i + j]) * D[i + j] + CD[j];
            }
        return (AB[0] + AB[1]) * (CD[0] + CD[1]);
    }
    default: {
        v4d AB[2], CD[2];
        for (int j = 0; j < 2; j++) { AB[j] = vset1(A[j]) * B[j]; CD[j] = vset1(Cm[j]) * D[j]; }
        for (int i = 2; i < N - 1; i += 2)
            for (int j = 0; j < 2; j++) {
                AB[j] = vset1(A[i + j]) * B[i + j] + AB[j];
                CD[j] = vset1(Cm[i + j]) * D[i + j] + CD[j];
            }
        return (vset1(A[N - 1]) * B[N - 1] + (AB[0] + AB[1])) *
               (vset1(Cm[N - 1]) * D[N - 1] + (CD[0] + CD[1]));
    }
    }
}

/* productVecMat with max (458-567): X[i] = sum_j A[j] M[i][j]; Xmax = max|X| */
INLINE void prod_vec_mat(const v4d* A, const double* M, v4d* X, v4d* Xmax, const int N) {
    if (N == 4) {
        for (int i = 0; i < 4; i++) {
            const double* m = M + 4 * i;
            X[i] = (A[1] * vset1(m[1]) + A[0] * vset1(m[0])) + (A[3] * vset1(m[3]) + A[2] * vset1(m[2]));
            if (Xmax) *Xmax = vmax(*Xmax, vabs(X[i]));
        }
        return;
    }
    for (int i = 0; i < N; i++, M += N) {
        v4d r;
        if (N < 4) {
            r = A[0] * vset1(M[0]);
            for (int j = 1; j < N; j++) r = A[j] * vset1(M[j]) + r;
        } else if (N % 4 == 0 || N % 4 == 1) {
            const int lim = (N % 4 == 0) ? N : N - 1;
            v4d Vv[4];
            for (int j = 0; j < 4; j++) Vv[j] = A[j] * vset1(M[j]);
            for (int x = 4; x < lim; x += 4)
                for (int j = 0; j < 4; j++) Vv[j] = A[x + j] * vset1(M[x + j]) + Vv[j];
            r = (Vv[0] + Vv[1]) + (Vv[2] + Vv[3]);
            if (N % 4 == 1) r = A[N - 1] * vset1(M[N - 1]) + r;
        } else {
            const int lim = (N % 4 == 2) ? N : N - 1;
            v4d Vv[2];
            for (int j = 0; j < 2; j++) Vv[j] = A[j] * vset1(M[j]);
            for (int x = 2; x < lim; x += 2)
                for (int j = 0; j < 2; j++) Vv[j] = A[x + j] * vset1(M[x + j]) + Vv[j];
            r = Vv[0] + Vv[1];
            if (N % 4 == 3) r = A[N - 1] * vset1(M[N - 1]) + r;
        }
        X[i] = r;
        if (Xmax) *Xmax = vmax(*Xmax, vabs(r));
    }
}

/* dotProduct3Vec (737-796): sum A[i]*B[i]*C[i], A scalar row */
INLINE v4d dot3(const double* A, const v4d* B, const v4d* Cv, const int N) {
    switch (N % 4) {
    case 0:
    case 1: {
        const int lim = (N % 4 == 0) ? N : N - 1;
        v4d Vv[4];
        for (int j = 0; j < 4; j++) Vv[j] = vset1(A[j]) * B[j] * Cv[j];
        for (int i = 4; i < lim; i += 4)
            for (int j = 0; j < 4; j++) Vv[j] = (vset1(A[i + j]) * B[i + j]) * Cv[i + j] + Vv[j];
        v4d X = (Vv[0] + Vv[1]) + (Vv[2] + Vv[3]);
        if (N % 4 == 1) X = (vset1(A[N - 1]) * B[N - 1]) * Cv[N - 1] + X;
        return X;
    }
    default: {
        const int lim = (N % 4 == 2) ? N : N - 1;
        v4d Vv[2];
        for (int j = 0; j < 2; j++) Vv[j] = vset1(A[j]) * B[j] * Cv[j];
        for (int i = 2; i < lim; i += 2)
            for (int j = 0; j < 2; j++) Vv[j] = (vset1(A[i + j]) * B[i + j]) * Cv[i + j] + Vv[j];
        v4d X = Vv[0] + Vv[1];
        if (N % 4 == 3) X = (vset1(A[N - 1]) * B[N - 1]) * Cv[N - 1] + X;
        return X;
    }
    }
}

/* dotProductTriple (666-725) over N = block entries; pairwise accumulators */
INLINE void dot_triple(const double* A, const double* B, const double* Cm, const v4d* D,
                       v4d* X, v4d* Y, v4d* Z, const int N, const int nstates) {
    v4d AD[2], BD[2], CD[2];
    for (int j = 0; j < 2; j++) {
        AD[j] = vset1(A[j]) * D[j];
        BD[j] = vset1(B[j]) * D[j];
        CD[j] = vset1(Cm[j]) * D[j];
    }
    const int lim = ((nstates & 1) == 0) ? N : N - 1;
    for (int i = 2; i < lim; i += 2)
        for (int j = 0; j < 2; j++) {
            AD[j] = vset1(A[i + j]) * D[i + j] + AD[j];
            BD[j] = vset1(B[i + j]) * D[i + j] + BD[j];
            CD[j] = vset1(Cm[i + j]) * D[i + j] + CD[j];
        }
    if ((nstates & 1) == 0) {
        *X = AD[0] + AD[1];
        *Y = BD[0] + BD[1];
        *Z = CD[0] + CD[1];
    } else {
        *X = vset1(A[N - 1]) * D[N - 1] + (AD[0] + AD[1]);
        *Y = vset1(B[N - 1]) * D[N - 1] + (BD[0] + BD[1]);
        *Z = vset1(Cm[N - 1]) * D[N - 1] + (CD[0] + CD[1]);
    }
}

/* ---- tree object --------------------------------------------------------- */

orc_tree* orc_create(int ntaxa, int nstates, int ncat, int npatterns, int nslots,
                     int state_unknown, int safe_numeric, int nthreads) {
    orc_tree* t = (orc_tree*)calloc(1, sizeof(orc_tree));
    t->ntaxa = ntaxa;
    t->n = nstates;
    t->C = ncat;
    t->P = npatterns;
    t->Ppad = (npatterns + VS - 1) / VS * VS;
    t->nslots = nslots;
    t->SU = state_unknown;
    t->safe = safe_numeric;
    t->block = (size_t)nstates * ncat;
    const size_t Pp = (size_t)t->Ppad;
    t->states = (uint8_t*)amalloc((size_t)ntaxa * Pp);
    memset(t->states, state_unknown, (size_t)ntaxa * Pp);
    t->freq = (double*)amalloc(Pp * 8);
    t->invar = (double*)amalloc(Pp * 8);
    const size_t nn = (size_t)nstates * nstates;
    t->eval = (double*)amalloc(nstates * 8);
    t->evec = (double*)amalloc(nn * 8);
    t->inv_evec = (double*)amalloc(nn * 8);
    t->tip = (double*)amalloc((size_t)(state_unknown + 1) * nstates * 8);
    t->rates = (double*)amalloc(ncat * 8);
    t->props = (double*)amalloc(ncat * 8);
    t->partial = (double*)amalloc((size_t)nslots * Pp * t->block * 8);
    t->scale = (uint16_t*)amalloc((size_t)nslots * Pp * (safe_numeric ? ncat : 1) * 2);
    t->theta = (double*)amalloc(Pp * t->block * 8);
    t->buffer_scale = (double*)amalloc(Pp * 8);
    t->pattern_lh = (double*)amalloc(Pp * 8);
    t->tipbuf = NULL;
    orc_set_threads(t, nthreads);
    return t;
}

void orc_set_threads(orc_tree* t, int nthreads) {
    /* PhyloTree::setNumThreads (tree/phylotreesse.cpp:42-52) */
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 1 && nthreads > t->P / 8) nthreads = t->P / 8 > 1 ? t->P / 8 : 1;
    t->nthreads = nthreads;
    t->npackets = nthreads == 1 ? 1 : nthreads * 2;
    free(t->tipbuf);
    t->tipbuf = (double*)amalloc((size_t)t->npackets * 3 * t->block * VS * 8);
}

void orc_destroy(orc_tree* t) {
    if (!t) return;
    free(t->states); free(t->freq); free(t->invar); free(t->eval); free(t->evec);
    free(t->inv_evec); free(t->tip); free(t->rates); free(t->props); free(t->partial);
    free(t->scale); free(t->theta); free(t->buffer_scale); free(t->pattern_lh); free(t->info);
    free(t->tipbuf);
    free(t->pars_set);
    free(t->pars_changes);
    free(t->pf_tip);
    free(t->pf_node);
    free(t->pf_score);
    free(t->nr_ceval); free(t->nr_cevec); free(t->nr_cinv); free(t->nr_Q); free(t->nr_pi);
    free(t->nr_mix); free(t->nr_flags);
    free(t->sk_cost); free(t->sk_tip); free(t->sk_partial);
    free(t);
}

void orc_set_patterns(orc_tree* t, const uint8_t* states, const double* freq, const double* invar) {
    t->pf_nblocks = 0;   /* bit planes rebuilt on the next parsimony_fast */
    for (int s = 0; s < t->ntaxa; s++)
        memcpy(t->states + (size_t)s * t->Ppad, states + (size_t)s * t->P, t->P);
    memcpy(t->freq, freq, (size_t)t->P * 8);
    memcpy(t->invar, invar, (size_t)t->P * 8);
    for (int p = t->P; p < t->Ppad; p++) { t->freq[p] = 0.0; t->invar[p] = 0.0; }
    orc_set_asc(t, t->nasc);
}

/* Lewis's ascertainment bias correction (ASC_VARIANT; the unobserved constant
 * patterns of Alignment::getUnobservedConstPatterns, alignment/alignment.cpp:
 * 4140-4157, placed after the observed ones as the kernels place them,
 * tree/phylokernelnew.h:1314-1316) */
void orc_set_asc(orc_tree* t, int nasc) {
    t->nasc = nasc;
    t->nsite = 0.0;
    for (int p = 0; p < t->P - nasc; p++) t->nsite += t->freq[p];
}

/* lh of an unobserved constant pattern: rescaled patterns count once at the
 * scaling threshold (tree/phylokernelnew.h:2940-2947, 2506-2515, 3345-3351) */
INLINE double asc_lh(double lh, double scale_log) { return scale_log != 0.0 ? lh * SCALING_THRESHOLD : lh; }

void orc_set_model(orc_tree* t, const double* eval, const double* evec, const double* inv_evec,
                   const double* tip_eigen, const double* rates, const double* props) {
    const size_t nn = (size_t)t->n * t->n;
    memcpy(t->eval, eval, t->n * 8);
    memcpy(t->evec, evec, nn * 8);
    memcpy(t->inv_evec, inv_evec, nn * 8);
    memcpy(t->tip, tip_eigen, (size_t)(t->SU + 1) * t->n * 8);
    memcpy(t->rates, rates, t->C * 8);
    memcpy(t->props, props, t->C * 8);
}

/* computeBounds (tree/phylokernelnew.h:1112-1143) */
static void compute_bounds(int threads, int packets, size_t elements, size_t* limits) {
    elements = (elements + VS - 1) / VS * VS;
    size_t start = 0;
    int k = 0;
    for (int wave = packets / threads; wave >= 1; --wave) {
        size_t this_wave = elements - start;
        if (1 < wave) this_wave = (this_wave * 3) / 4;
        this_wave = (this_wave + VS - 1) / VS * VS;
        const size_t stop = start + this_wave;
        for (int togo = threads; togo >= 1; --togo) {
            limits[k++] = start;
            size_t bs = (stop - start) / togo;
            bs = (bs + VS - 1) / VS * VS;
            start += bs;
        }
    }
    limits[k] = elements;
    for (int i = 1; i <= packets; i++)
        if (limits[i] > elements) limits[i] = elements;
}

/* per-op precomputed tables (computePartialInfo, tree/phylokernelnew.h:1008-1109):
 * echild[child][c][x][i] = evec[x][i] * exp(eval[i] * r_c * t_child) and, for
 * tip children, leaf[child][s][c][x] = sum_i echild[c][x][i] tip[s][i] with
 * leaf[SU] = 1. */
typedef struct {
    const double* echild[2];
    const double* leaf[2];
} op_info;

static void nr_partial_info(orc_tree* t, const orc_op* op, const uint8_t* flags, double* ech,
                            double* leaf);

static void partial_info(orc_tree* t, const orc_op* op, double* ech, double* leaf) {
    const int n = t->n, C = t->C;
    const size_t block = t->block;
    for (int k = 0; k < 2; k++) {
        double* E = ech + (size_t)k * block * n;
        for (int c = 0; c < C; c++) {
            const double len = t->rates[c] * op->child[k].length;
            for (int x = 0; x < n; x++)
                for (int i = 0; i < n; i++)
                    E[((size_t)c * n + x) * n + i] = t->evec[x * n + i] * exp(t->eval[i] * len);
        }
        if (op->child[k].is_tip) {
            double* L = leaf + (size_t)k * (t->SU + 1) * block;
            for (int s = 0; s < t->SU; s++)
                for (int c = 0; c < C; c++)
                    for (int x = 0; x < n; x++) {
                        const double* e = E + ((size_t)c * n + x) * n;
                        double v = e[0] * t->tip[s * n + 0];
                        for (int i = 1; i < n; i++) v += e[i] * t->tip[s * n + i];
                        L[(size_t)s * block + c * n + x] = v;
                    }
            for (size_t x = 0; x < block; x++) L[(size_t)t->SU * block + x] = 1.0;
        }
    }
}

static void ensure_info(orc_tree* t, int nops) {
    const size_t per = 2 * t->block * t->n + 2 * (size_t)(t->SU + 1) * t->block;
    const size_t need = (size_t)(nops > 0 ? nops : 1) * per;
    if (need > t->info_cap) {
        free(t->info);
        t->info = (double*)amalloc(need * 8);
        t->info_cap = need;
    }
}

static double* op_echild(orc_tree* t, int i) {
    const size_t per = 2 * t->block * t->n + 2 * (size_t)(t->SU + 1) * t->block;
    return t->info + (size_t)i * per;
}
static double* op_leaf(orc_tree* t, int i) { return op_echild(t, i) + 2 * t->block * t->n; }

INLINE double* slot_ptr(orc_tree* t, int slot) {
    return t->partial + (size_t)slot * t->Ppad * t->block;
}
INLINE uint16_t* scale_ptr(orc_tree* t, int slot) {
    return t->scale + (size_t)slot * t->Ppad * (t->safe ? t->C : 1);
}

/* load interleaved tip vectors of 4 patterns from a [SU+1][block] table */
INLINE void load_tip(orc_tree* t, const double* table, int taxon, size_t ptn, double* vec) {
    const size_t block = t->block;
    const uint8_t* row = t->states + (size_t)taxon * t->Ppad;
    for (int x = 0; x < VS; x++) {
        const double* src = table + (size_t)row[ptn + x] * block;
        for (size_t i = 0; i < block; i++) vec[i * VS + x] = src[i];
    }
}

/* scaling of one 4-pattern vector block (scaleLikelihood semantics,
 * tree/phylokernelnew.h:1993-2030): rescale a lane when its max |partial| is
 * below 2^-256 and its pattern has no +I mass */
INLINE void scale_lanes(double* dad, uint16_t* sc, const double* invar, v4d lh_max,
                        size_t count, size_t stride_sc) {
    for (int x = 0; x < VS; x++)
        if (lh_max[x] < SCALING_THRESHOLD && invar[x] == 0.0) {
            for (size_t i = 0; i < count; i++)
                dad[i * VS + x] = ldexp(dad[i * VS + x], SCALING_THRESHOLD_EXP);
            sc[x * stride_sc] += 1;
        }
}

/* computePartialLikelihoodSIMD for binary nodes (tree/phylokernelnew.h:1632-2031) */
INLINE void partial_body(orc_tree* t, const orc_op* op, const op_info* inf, size_t lo, size_t hi,
                         double* tipbuf, const int n) {
    const int C = t->C, safe = t->safe;
    const size_t block = (size_t)n * C;
    const int nsc = safe ? C : 1;
    int left = 0, right = 1;
    /* a tip child, when there is one, goes left (1513-1524) */
    if (!op->child[0].is_tip && op->child[1].is_tip) { left = 1; right = 0; }
    const orc_child* L = &op->child[left];
    const orc_child* R = &op->child[right];
    double* dad_base = slot_ptr(t, op->dad_slot);
    uint16_t* sdad = scale_ptr(t, op->dad_slot);
    v4d tmp[64];
    if (L->is_tip && R->is_tip) {
        /* tip-tip: scale counters zero, no rescale (1632-1745) */
        memset(sdad + lo * nsc, 0, (hi - lo) * nsc * sizeof(uint16_t));
        double* vl = tipbuf;
        double* vr = tipbuf + block * VS;
        for (size_t ptn = lo; ptn < hi; ptn += VS) {
            load_tip(t, inf->leaf[left], L->id, ptn, vl);
            load_tip(t, inf->leaf[right], R->id, ptn, vr);
            v4d* dad = (v4d*)(dad_base + ptn * block);
            const v4d* a = (const v4d*)vl;
            const v4d* b = (const v4d*)vr;
            for (int c = 0; c < C; c++) {
                for (int x = 0; x < n; x++) tmp[x] = a[c * n + x] * b[c * n + x];
                prod_vec_mat(tmp, t->inv_evec, dad + c * n, NULL, n);
            }
        }
        return;
    }
    if (L->is_tip) {
        /* tip-internal (1748-1905) */
        const double* rbase = slot_ptr(t, R->id);
        const uint16_t* sright = scale_ptr(t, R->id);
        memcpy(sdad + lo * nsc, sright + lo * nsc, (hi - lo) * nsc * sizeof(uint16_t));
        double* vl = tipbuf;
        const double* ER = inf->echild[right];
        for (size_t ptn = lo; ptn < hi; ptn += VS) {
            load_tip(t, inf->leaf[left], L->id, ptn, vl);
            v4d* dad = (v4d*)(dad_base + ptn * block);
            const v4d* pr = (const v4d*)(rbase + ptn * block);
            const v4d* a = (const v4d*)vl;
            v4d lh_max = vset1(0.0);
            for (int c = 0; c < C; c++) {
                if (safe) lh_max = vset1(0.0);
                const double* e = ER + (size_t)c * n * n;
                for (int x = 0; x < n; x++) tmp[x] = a[c * n + x] * dot_vec_s(e + x * n, pr + c * n, n);
                prod_vec_mat(tmp, t->inv_evec, dad + c * n, &lh_max, n);
                if (safe)
                    scale_lanes((double*)(dad + c * n), sdad + ptn * C + c, t->invar + ptn, lh_max, n, C);
            }
            if (!safe) scale_lanes((double*)dad, sdad + ptn, t->invar + ptn, lh_max, block, 1);
        }
        return;
    }
    /* internal-internal (1907-2031) */
    const double* lbase = slot_ptr(t, L->id);
    const double* rbase = slot_ptr(t, R->id);
    const uint16_t* sl = scale_ptr(t, L->id);
    const uint16_t* sr = scale_ptr(t, R->id);
    const double* EL = inf->echild[left];
    const double* ER = inf->echild[right];
    for (size_t ptn = lo; ptn < hi; ptn += VS) {
        v4d* dad = (v4d*)(dad_base + ptn * block);
        const v4d* pl = (const v4d*)(lbase + ptn * block);
        const v4d* pr = (const v4d*)(rbase + ptn * block);
        v4d lh_max = vset1(0.0);
        if (!safe)
            for (int x = 0; x < VS; x++) sdad[ptn + x] = sl[ptn + x] + sr[ptn + x];
        for (int c = 0; c < C; c++) {
            if (safe) {
                lh_max = vset1(0.0);
                for (int x = 0; x < VS; x++)
                    sdad[(ptn + x) * C + c] = sl[(ptn + x) * C + c] + sr[(ptn + x) * C + c];
            }
            const double* el = EL + (size_t)c * n * n;
            const double* er = ER + (size_t)c * n * n;
            for (int x = 0; x < n; x++)
                tmp[x] = dot_dual(el + x * n, pl + c * n, er + x * n, pr + c * n, n);
            prod_vec_mat(tmp, t->inv_evec, dad + c * n, &lh_max, n);
            if (safe) scale_lanes((double*)(dad + c * n), sdad + ptn * C + c, t->invar + ptn, lh_max, n, C);
        }
        if (!safe) scale_lanes((double*)dad, sdad + ptn, t->invar + ptn, lh_max, block, 1);
    }
}

static void partial_dispatch(orc_tree* t, const orc_op* op, const op_info* inf, size_t lo, size_t hi,
                             double* tipbuf) {
    if (t->n == 4) partial_body(t, op, inf, lo, hi, tipbuf, 4);
    else if (t->n == 20) partial_body(t, op, inf, lo, hi, tipbuf, 20);
    else partial_body(t, op, inf, lo, hi, tipbuf, t->n);
}

static void prepare_ops(orc_tree* t, const orc_op* ops, int nops, op_info* infos) {
    ensure_info(t, nops);
#pragma omp parallel for schedule(static) num_threads(t->nthreads) if (nops >= 3)
    for (int i = 0; i < nops; i++) {
        double* ech = op_echild(t, i);
        double* leaf = op_leaf(t, i);
        if (t->nonrev) nr_partial_info(t, &ops[i], t->nr_flags + 2 * i, ech, leaf);
        else partial_info(t, &ops[i], ech, leaf);
        for (int k = 0; k < 2; k++) {
            infos[i].echild[k] = ech + (size_t)k * t->block * t->n;
            infos[i].leaf[k] = leaf + (size_t)k * (t->SU + 1) * t->block;
        }
    }
}

static void run_ops_packet(orc_tree* t, const orc_op* ops, int nops, const op_info* infos,
                           size_t lo, size_t hi, int packet) {
    double* tipbuf = t->tipbuf + (size_t)packet * 3 * t->block * VS;
    for (int i = 0; i < nops; i++) partial_dispatch(t, &ops[i], &infos[i], lo, hi, tipbuf);
}


/* ---- non-reversible and mixture models in the state basis ---------------
 * The kernel set of tree/phylokernelnonrev.h: partials hold state-basis
 * likelihoods, each child's table is its transition matrix P(r_c t) of the
 * category's mixture component (computePartialInfo's !useRevKernel branch,
 * tree/phylokernelnew.h:934-986: transposed for a child whose direction is
 * toward the root, the state frequencies for the root leaf), and the branch
 * likelihood and its derivatives apply P, QP and QQP between the two ends
 * (tree/phylokernelnonrev.h:586-1056, 1062-1528). The partial update is the
 * binary-node body above with the identity as the output transform: the
 * products (E_l l) * (E_r r) of the reference's nonrev partial kernel
 * (tree/phylokernelnonrev.h:36-580). Child and end flags ride in is_tip:
 * bit 0 a tip, bit 1 toward the root, bit 2 the root leaf. */

/* ModelMarkov::computeTransMatrixNonrev, eigen route (model/modelmarkov.cpp:
 * 480-492): P(t) = Re(V diag(exp(lambda t)) V^-1) */
static void nr_trans(const orc_tree* t, int m, double len, double* P) {
    const int n = t->n;
    const double* ev = t->nr_ceval + (size_t)m * 2 * n;
    const double* V = t->nr_cevec + (size_t)m * 2 * n * n;
    const double* W = t->nr_cinv + (size_t)m * 2 * n * n;
    double er[64], ei[64];
    for (int i = 0; i < n; i++) {
        const double a = exp(ev[2 * i] * len), b = ev[2 * i + 1] * len;
        er[i] = a * cos(b);
        ei[i] = a * sin(b);
    }
    for (int x = 0; x < n; x++)
        for (int y = 0; y < n; y++) {
            double s = 0.0;
            for (int i = 0; i < n; i++) {
                const double vr = V[2 * (x * n + i)], vi = V[2 * (x * n + i) + 1];
                const double dr = vr * er[i] - vi * ei[i], di = vr * ei[i] + vi * er[i];
                s += dr * W[2 * (i * n + y)] - di * W[2 * (i * n + y) + 1];
            }
            P[x * n + y] = s;
        }
}

/* computeTransDerv, non-reversible (model/modelmarkov.cpp:735-766):
 * P, Q P, Q Q P */
static void nr_trans_derv(const orc_tree* t, int m, double len, double* P, double* D1, double* D2) {
    const int n = t->n;
    const double* Q = t->nr_Q + (size_t)m * n * n;
    nr_trans(t, m, len, P);
    for (int x = 0; x < n; x++)
        for (int y = 0; y < n; y++) {
            double s = 0.0;
            for (int k = 0; k < n; k++) s += Q[x * n + k] * P[k * n + y];
            D1[x * n + y] = s;
        }
    for (int x = 0; x < n; x++)
        for (int y = 0; y < n; y++) {
            double s = 0.0;
            for (int k = 0; k < n; k++) s += Q[x * n + k] * D1[k * n + y];
            D2[x * n + y] = s;
        }
}

int orc_set_model_nonrev(orc_tree* t, int nmix, int fused, const double* ceval, const double* cevec,
                         const double* cinv_evec, const double* rate_matrix, const double* state_freq,
                         const double* mix_weights, const double* site_rates, const double* site_props,
                         int nrates, const double* tip, int rooted) {
    const int n = t->n, C = t->C;
    if (nmix < 1 || nrates < 1 || n > 64 || C != (fused ? nrates : nrates * nmix)) return -1;
    if (fused && nmix != nrates) return -1;
    const size_t nn = (size_t)n * n;
    free(t->nr_ceval); free(t->nr_cevec); free(t->nr_cinv); free(t->nr_Q); free(t->nr_pi);
    free(t->nr_mix);
    t->nr_ceval = (double*)amalloc((size_t)nmix * 2 * n * 8);
    t->nr_cevec = (double*)amalloc((size_t)nmix * 2 * nn * 8);
    t->nr_cinv = (double*)amalloc((size_t)nmix * 2 * nn * 8);
    t->nr_Q = (double*)amalloc((size_t)nmix * nn * 8);
    t->nr_pi = (double*)amalloc((size_t)nmix * n * 8);
    t->nr_mix = (int*)amalloc((size_t)C * sizeof(int));
    memcpy(t->nr_ceval, ceval, (size_t)nmix * 2 * n * 8);
    memcpy(t->nr_cevec, cevec, (size_t)nmix * 2 * nn * 8);
    memcpy(t->nr_cinv, cinv_evec, (size_t)nmix * 2 * nn * 8);
    memcpy(t->nr_Q, rate_matrix, (size_t)nmix * nn * 8);
    memcpy(t->nr_pi, state_freq, (size_t)nmix * n * 8);
    /* ncat_mix categories: rate category c % ncat of mixture component
     * c / denom, weight getProp * getMixtureWeight (tree/phylokernelnonrev.h:
     * 1107-1114) */
    const int denom = fused ? 1 : nrates;
    for (int c = 0; c < C; c++) {
        const int m = c / denom;
        t->nr_mix[c] = m;
        t->rates[c] = site_rates[c % nrates];
        t->props[c] = site_props[c % nrates] * mix_weights[m];
    }
    memset(t->inv_evec, 0, nn * 8);
    for (int i = 0; i < n; i++) t->inv_evec[i * n + i] = 1.0;
    memcpy(t->tip, tip, (size_t)(t->SU + 1) * n * 8);
    t->nmix = nmix;
    t->rooted = rooted;
    t->nonrev = 1;
    t->theta_ready = 0;
    return 0;
}

static void nr_partial_info(orc_tree* t, const orc_op* op, const uint8_t* flags, double* ech,
                            double* leaf) {
    const int n = t->n, C = t->C;
    const size_t block = t->block;
    double* P = (double*)malloc((size_t)n * n * 8);
    for (int k = 0; k < 2; k++) {
        double* E = ech + (size_t)k * block * n;
        for (int c = 0; c < C; c++) {
            nr_trans(t, t->nr_mix[c], t->rates[c] * op->child[k].length, P);
            double* e = E + (size_t)c * n * n;
            for (int x = 0; x < n; x++)
                for (int i = 0; i < n; i++) e[x * n + i] = (flags[k] & 2) ? P[i * n + x] : P[x * n + i];
        }
        if (!op->child[k].is_tip) continue;
        double* L = leaf + (size_t)k * (t->SU + 1) * block;
        if (flags[k] & 4) {
            /* the root leaf: the state frequencies (tree/phylokernelnew.h:961-966) */
            for (int s = 0; s <= t->SU; s++)
                for (int c = 0; c < C; c++)
                    memcpy(L + (size_t)s * block + c * n, t->nr_pi + (size_t)t->nr_mix[c] * n, n * 8);
            continue;
        }
        for (int s = 0; s < t->SU; s++)
            for (int c = 0; c < C; c++)
                for (int x = 0; x < n; x++) {
                    const double* e = E + ((size_t)c * n + x) * n;
                    double v = 0.0;
                    for (int i = 0; i < n; i++) v += e[i] * t->tip[s * n + i];
                    L[(size_t)s * block + c * n + x] = v;
                }
        for (size_t x = 0; x < block; x++) L[(size_t)t->SU * block + x] = 1.0;
    }
    free(P);
}

/* the ops with their flags moved to nr_flags (a copy, freed by the caller) */
static orc_op* nr_strip(orc_tree* t, const orc_op* ops, int nops) {
    if (nops <= 0) return NULL;
    if (2 * nops > t->nr_flags_cap) {
        free(t->nr_flags);
        t->nr_flags = (uint8_t*)amalloc((size_t)2 * nops);
        t->nr_flags_cap = 2 * nops;
    }
    orc_op* c = (orc_op*)malloc(sizeof(orc_op) * nops);
    memcpy(c, ops, sizeof(orc_op) * nops);
    for (int i = 0; i < nops; i++)
        for (int k = 0; k < 2; k++) {
            t->nr_flags[2 * i + k] = (uint8_t)(c[i].child[k].is_tip & 6);
            c[i].child[k].is_tip &= 1;
        }
    return c;
}

static void nr_run_ops(orc_tree* t, const orc_op* ops, int nops) {
    if (nops <= 0) return;
    orc_op* own = nr_strip(t, ops, nops);
    op_info* infos = (op_info*)malloc(sizeof(op_info) * nops);
    prepare_ops(t, own, nops, infos);
    size_t limits[1024];
    compute_bounds(t->nthreads, t->npackets, t->Ppad, limits);
#pragma omp parallel for schedule(dynamic, 1) num_threads(t->nthreads)
    for (int pk = 0; pk < t->npackets; pk++) run_ops_packet(t, own, nops, infos, limits[pk], limits[pk + 1], pk);
    free(infos);
    free(own);
}

INLINE double nr_elem(const double* slot, size_t block, size_t ptn, size_t i) {
    return slot[(ptn / VS) * block * VS + i * VS + ptn % VS];
}

/* the safe kernel's per-category factor for a scale count above the minimum */
INLINE double nr_adj(int d) { return d == 0 ? 1.0 : (d == 1 ? SCALING_THRESHOLD : 0.0); }

/* branch lnL (derv = 0; pattern lnL kept) or df, ddf (derv = 1) between the
 * buffered ends at `length` */
static double nr_eval(orc_tree* t, double length, int derv, double* df_out, double* ddf_out, int* status) {
    const int n = t->n, C = t->C, safe = t->safe;
    const size_t block = t->block, nn = (size_t)n * n;
    const orc_end node = t->nr_node, dad = t->nr_dad;
    const int dad_tip = dad.is_tip & 1, dad_root = (dad.is_tip & 4) != 0;
    /* trans_mat (and derivatives) per category: * prop, * rate, rows * pi
     * when unrooted (tree/phylokernelnonrev.h:630-666, 1107-1131) */
    double* T = (double*)amalloc(3 * (size_t)C * nn * 8);
    double* D1 = (double*)malloc(nn * 8);
    double* D2 = (double*)malloc(nn * 8);
    for (int c = 0; c < C; c++) {
        const int m = t->nr_mix[c];
        const double r = t->rates[c], w = t->props[c];
        double* T0 = T + (size_t)c * nn;
        double* T1 = T + ((size_t)C + c) * nn;
        double* T2 = T + ((size_t)2 * C + c) * nn;
        nr_trans_derv(t, m, r * length, T0, D1, D2);
        for (size_t i = 0; i < nn; i++) {
            T0[i] *= w;
            T1[i] = D1[i] * (w * r);
            T2[i] = D2[i] * (w * r * r);
        }
        if (!t->rooted)
            for (int x = 0; x < n; x++)
                for (int y = 0; y < n; y++) {
                    const double f = t->nr_pi[(size_t)m * n + x];
                    T0[x * n + y] *= f;
                    T1[x * n + y] *= f;
                    T2[x * n + y] *= f;
                }
    }
    free(D1);
    free(D2);
    /* a tip dad: lut[k][s][c][x] = sum_y Tk[c][x][y] tip[s][y]; the root leaf:
     * the frequencies times prop, no derivative */
    double* lut = NULL;
    if (dad_tip) {
        const size_t ls = (size_t)(t->SU + 1) * block;
        lut = (double*)amalloc(3 * ls * 8);
        for (int s = 0; s <= t->SU; s++)
            for (int c = 0; c < C; c++)
                for (int x = 0; x < n; x++) {
                    if (dad_root) {
                        lut[(size_t)s * block + c * n + x] = t->nr_pi[(size_t)t->nr_mix[c] * n + x] * t->props[c];
                        continue;
                    }
                    for (int k = 0; k < 3; k++) {
                        const double* Tk = T + ((size_t)k * C + c) * nn + (size_t)x * n;
                        double v = 0.0;
                        for (int y = 0; y < n; y++) v += Tk[y] * t->tip[s * n + y];
                        lut[k * ls + (size_t)s * block + c * n + x] = v;
                    }
                }
    }
    const size_t ls = (size_t)(t->SU + 1) * block;
    const double* pn = slot_ptr(t, node.id);
    const uint16_t* sn = scale_ptr(t, node.id);
    const double* pd = dad_tip ? NULL : slot_ptr(t, dad.id);
    const uint16_t* sd = dad_tip ? NULL : scale_ptr(t, dad.id);
    const uint8_t* drow = dad_tip ? t->states + (size_t)dad.id * t->Ppad : NULL;
    const size_t nobs = (size_t)(t->P - t->nasc);
    double all = 0.0, all_df = 0.0, all_ddf = 0.0, all_pc = 0.0, all_dfc = 0.0, all_ddfc = 0.0;
#pragma omp parallel for num_threads(t->nthreads) \
    reduction(+ : all, all_df, all_ddf, all_pc, all_dfc, all_ddfc)
    for (size_t p = 0; p < (size_t)t->P; p++) {
        double adjn[16], adjd[16], lh_cat[16];
        int minsc = 0;
        if (safe) {
            if (!derv) {
                /* the branch lnL: per-category sums against their minimum (1422-1442) */
                int sum[16];
                for (int c = 0; c < C; c++) {
                    sum[c] = sn[p * C + c] + (sd ? sd[p * C + c] : 0);
                    if (c == 0 || sum[c] < minsc) minsc = sum[c];
                }
                for (int c = 0; c < C; c++) { adjn[c] = nr_adj(sum[c] - minsc); adjd[c] = 1.0; }
            } else {
                /* derivatives: each end against its own minimum (893-947) */
                int mn = 0, md = 0;
                for (int c = 0; c < C; c++) if (c == 0 || sn[p * C + c] < mn) mn = sn[p * C + c];
                if (sd) for (int c = 0; c < C; c++) if (c == 0 || sd[p * C + c] < md) md = sd[p * C + c];
                for (int c = 0; c < C; c++) {
                    adjn[c] = nr_adj(sn[p * C + c] - mn);
                    adjd[c] = sd ? nr_adj(sd[p * C + c] - md) : 1.0;
                }
                minsc = mn + md;
            }
        } else {
            minsc = sn[p] + (sd ? sd[p] : 0);
            for (int c = 0; c < C; c++) adjn[c] = adjd[c] = 1.0;
        }
        double lh = 0.0, d1 = 0.0, d2 = 0.0;
        for (int c = 0; c < C; c++) {
            double a0 = 0.0, a1 = 0.0, a2 = 0.0;
            for (int x = 0; x < n; x++) {
                const double nv = nr_elem(pn, block, p, (size_t)c * n + x) * (derv ? adjn[c] : 1.0);
                double v0, v1 = 0.0, v2 = 0.0;
                if (dad_tip) {
                    const size_t o = (size_t)drow[p] * block + c * n + x;
                    v0 = lut[o];
                    if (derv && !dad_root) { v1 = lut[ls + o]; v2 = lut[2 * ls + o]; }
                } else {
                    const double* T0 = T + (size_t)c * nn + (size_t)x * n;
                    const double* T1 = T0 + (size_t)C * nn;
                    const double* T2 = T1 + (size_t)C * nn;
                    v0 = 0.0;
                    for (int y = 0; y < n; y++) {
                        const double dv = nr_elem(pd, block, p, (size_t)c * n + y) * (derv ? adjd[c] : 1.0);
                        v0 += T0[y] * dv;
                        if (derv) { v1 += T1[y] * dv; v2 += T2[y] * dv; }
                    }
                }
                a0 += nv * v0;
                a1 += nv * v1;
                a2 += nv * v2;
            }
            lh_cat[c] = a0;
            d1 += a1;
            d2 += a2;
        }
        for (int c = 0; c < C; c++) lh += lh_cat[c] * (derv ? 1.0 : adjn[c]);
        lh += t->invar[p];
        if (p >= nobs) {
            /* an unobserved pattern (+ASC): its probability, rescaled once when
             * its counter is not zero (tree/phylokernelnonrev.h:1000-1013,
             * 1455-1469) */
            const double f = minsc != 0 ? SCALING_THRESHOLD : 1.0;
            all_pc += lh * f;
            if (derv) {
                all_dfc += d1 * f;
                all_ddfc += d2 * f;
            }
            continue;
        }
        if (!derv) {
            const double lp = log(lh) + minsc * LOG_SCALING_THRESHOLD;
            t->pattern_lh[p] = lp;
            all += lp * t->freq[p];
        } else {
            const double inv = 1.0 / lh, df_frac = d1 * inv, ddf_frac = d2 * inv;
            const double tmp1 = df_frac * t->freq[p], tmp2 = ddf_frac * t->freq[p];
            all_df += tmp1;
            all_ddf += tmp2 - tmp1 * df_frac;
        }
    }
    free(T);
    free(lut);
    int st = 0;
    if (derv) {
        if (t->nasc) {
            /* ASC_Lewis (tree/phylokernelnonrev.h:1040-1047) */
            const double pv = 1.0 - all_pc;
            const double df_frac = all_dfc / pv, ddf_frac = all_ddfc / pv;
            all_df += t->nsite * df_frac;
            all_ddf += t->nsite * (ddf_frac + df_frac * df_frac);
        }
        *df_out = all_df;
        *ddf_out = all_ddf;
        if (!isfinite(all_df)) {
            *df_out = *ddf_out = 0.0;
            st = 4;
        }
    } else if (!isfinite(all)) {
        /* per-pattern fix-up (tree/phylokernelnonrev.h:1486-1497) */
        st = 4;
        all = 0.0;
        for (size_t p = 0; p < nobs; p++) {
            if (!isfinite(t->pattern_lh[p])) t->pattern_lh[p] = LOG_SCALING_THRESHOLD * 4;
            all += t->pattern_lh[p] * t->freq[p];
        }
    }
    if (!derv && t->nasc) {
        /* ASC_Lewis (tree/phylokernelnonrev.h:1505-1515) */
        const double corr = log(1.0 - all_pc);
        for (size_t p = 0; p < nobs; p++) t->pattern_lh[p] -= corr;
        all -= t->nsite * corr;
    }
    if (status) *status = st;
    return all;
}

static orc_end nr_end(orc_end e) {
    orc_end r = e;
    r.is_tip = e.is_tip & 1 ? (e.is_tip & 4) | 1 : 0;
    return r;
}

void orc_update_partials(orc_tree* t, const orc_op* ops, int nops) {
    if (t->nonrev) {
        nr_run_ops(t, ops, nops);
        return;
    }
    if (nops <= 0) return;
    op_info* infos = (op_info*)malloc(sizeof(op_info) * nops);
    prepare_ops(t, ops, nops, infos);
    size_t limits[1024];
    compute_bounds(t->nthreads, t->npackets, t->Ppad, limits);
#pragma omp parallel for schedule(dynamic, 1) num_threads(t->nthreads)
    for (int pk = 0; pk < t->npackets; pk++) run_ops_packet(t, ops, nops, infos, limits[pk], limits[pk + 1], pk);
    free(infos);
}

/* computeLikelihoodBranchSIMD (tree/phylokernelnew.h:2641-3223), save_log_value = true */
INLINE double branch_body(orc_tree* t, const orc_op* ops, int nops, const op_info* infos,
                          orc_end node, orc_end dad, const double* val, const double* tiplut,
                          const int n, double* prob_const) {
    const int C = t->C, safe = t->safe;
    const size_t block = (size_t)n * C;
    const size_t nobs = (size_t)(t->P - t->nasc);
    size_t limits[1024];
    compute_bounds(t->nthreads, t->npackets, t->Ppad, limits);
    double all = 0.0, all_const = 0.0;
#pragma omp parallel for schedule(dynamic, 1) num_threads(t->nthreads) reduction(+ : all, all_const)
    for (int pk = 0; pk < t->npackets; pk++) {
        const size_t lo = limits[pk], hi = limits[pk + 1];
        run_ops_packet(t, ops, nops, infos, lo, hi, pk);
        double* vtip = t->tipbuf + (size_t)pk * 3 * t->block * VS + 2 * t->block * VS;
        const double* pn = slot_ptr(t, node.id);
        const uint16_t* sn = scale_ptr(t, node.id);
        v4d tree_lh = vset1(0.0);
        for (size_t ptn = lo; ptn < hi; ptn += VS) {
            const v4d* vnode = (const v4d*)(pn + ptn * block);
            v4d lh_cat[16];
            v4d lh = vset1(0.0);
            if (dad.is_tip) {
                load_tip(t, tiplut, dad.id, ptn, vtip);
                const v4d* vt = (const v4d*)vtip;
                for (int c = 0; c < C; c++) {
                    lh_cat[c] = dot_vec_v(vt + c * n, vnode + c * n, n);
                    if (!safe) lh += lh_cat[c];
                }
            } else {
                const v4d* vdad = (const v4d*)(slot_ptr(t, dad.id) + ptn * block);
                for (int c = 0; c < C; c++) {
                    lh_cat[c] = dot3(val + c * n, vdad + c * n, vnode + c * n, n);
                    if (!safe) lh += lh_cat[c];
                }
            }
            v4d minsc;
            if (safe) {
                const uint16_t* sd = dad.is_tip ? NULL : scale_ptr(t, dad.id);
                for (int x = 0; x < VS; x++) {
                    uint16_t sum[16], mn = 0;
                    for (int c = 0; c < C; c++) {
                        sum[c] = (uint16_t)(sn[(ptn + x) * C + c] + (sd ? sd[(ptn + x) * C + c] : 0));
                        if (c == 0 || sum[c] < mn) mn = sum[c];
                    }
                    minsc[x] = mn;
                    for (int c = 0; c < C; c++) {
                        if (sum[c] == mn + 1) lh_cat[c][x] *= SCALING_THRESHOLD;
                        else if (sum[c] > mn + 1) lh_cat[c][x] = 0.0;
                    }
                }
                /* sumVec<VectorClass, true> grouping (43-71) */
                const int cruft = C & 3;
                if (cruft < C) {
                    v4d acc[4] = {vset1(0), vset1(0), vset1(0), vset1(0)};
                    for (int c = 0; c < C - cruft; c += 4)
                        for (int j = 0; j < 4; j++) acc[j] += lh_cat[c + j];
                    acc[0] += acc[2];
                    acc[1] += acc[3];
                    acc[0] += acc[1];
                    lh += acc[0];
                }
                for (int c = C - cruft; c < C; c++) lh += lh_cat[c];
            } else {
                const uint16_t* sd = dad.is_tip ? NULL : scale_ptr(t, dad.id);
                for (int x = 0; x < VS; x++) minsc[x] = (double)(sn[ptn + x] + (sd ? sd[ptn + x] : 0));
            }
            minsc *= vset1(LOG_SCALING_THRESHOLD);
            lh = vabs(lh) + *(const v4d*)(t->invar + ptn);
            for (int x = 0; x < VS; x++) {
                if (ptn + x < nobs) {
                    const double lp = log(lh[x]) + minsc[x];
                    t->pattern_lh[ptn + x] = lp;
                    tree_lh[x] = lp * t->freq[ptn + x] + tree_lh[x];
                } else if (ptn + x < (size_t)t->P) {
                    all_const += asc_lh(lh[x], minsc[x]);
                }
            }
        }
        all += hadd(tree_lh);
    }
    *prob_const = all_const;
    return all;
}

double orc_likelihood_branch(orc_tree* t, const orc_op* ops, int nops, orc_end node, orc_end dad,
                             double length, double* pattern_lh, int* status) {
    if (t->nonrev) {
        nr_run_ops(t, ops, nops);
        t->nr_node = nr_end(node);
        t->nr_dad = nr_end(dad);
        const double lnl = nr_eval(t, length, 0, NULL, NULL, status);
        if (pattern_lh) memcpy(pattern_lh, t->pattern_lh, (size_t)(t->P - t->nasc) * 8);
        return lnl;
    }
    const int n = t->n, C = t->C;
    const size_t block = t->block;
    op_info* infos = (op_info*)malloc(sizeof(op_info) * (nops > 0 ? nops : 1));
    prepare_ops(t, ops, nops, infos);
    double* val = (double*)amalloc(block * 8);
    for (int c = 0; c < C; c++) {
        const double len = t->rates[c] * length;
        for (int i = 0; i < n; i++) val[c * n + i] = exp(t->eval[i] * len) * t->props[c];
    }
    double* tiplut = NULL;
    if (dad.is_tip) {
        tiplut = (double*)amalloc((size_t)(t->SU + 1) * block * 8);
        for (int s = 0; s <= t->SU; s++)
            for (int c = 0; c < C; c++)
                for (int i = 0; i < n; i++)
                    tiplut[(size_t)s * block + c * n + i] = val[c * n + i] * t->tip[s * n + i];
    }
    double lnl, prob_const = 0.0;
    if (n == 4) lnl = branch_body(t, ops, nops, infos, node, dad, val, tiplut, 4, &prob_const);
    else if (n == 20) lnl = branch_body(t, ops, nops, infos, node, dad, val, tiplut, 20, &prob_const);
    else lnl = branch_body(t, ops, nops, infos, node, dad, val, tiplut, n, &prob_const);
    const int nobs = t->P - t->nasc;
    int st = 0;
    if (!isfinite(lnl)) {
        /* per-pattern fix-up (tree/phylokernelnew.h:3115-3132) */
        st = 4;
        lnl = 0.0;
        for (int p = 0; p < nobs; p++) {
            if (!isfinite(t->pattern_lh[p])) t->pattern_lh[p] = LOG_SCALING_THRESHOLD * 4;
            lnl += t->pattern_lh[p] * t->freq[p];
        }
    }
    if (t->nasc) {
        /* ASC_Lewis (tree/phylokernelnew.h:3198-3221) */
        const double corr = log(1.0 - prob_const);
        for (int p = 0; p < nobs; p++) t->pattern_lh[p] -= corr;
        lnl -= t->nsite * corr;
    }
    if (status) *status = st;
    if (pattern_lh) memcpy(pattern_lh, t->pattern_lh, (size_t)nobs * 8);
    free(val);
    free(tiplut);
    free(infos);
    return lnl;
}

/* computeLikelihoodBufferSIMD (tree/phylokernelnew.h:2048-2231) */
void orc_likelihood_buffer(orc_tree* t, const orc_op* ops, int nops, orc_end node, orc_end dad) {
    if (t->nonrev) {
        /* no theta buffer in the nonrev kernel set: the ends, evaluated on demand */
        nr_run_ops(t, ops, nops);
        t->nr_node = nr_end(node);
        t->nr_dad = nr_end(dad);
        t->theta_ready = 1;
        return;
    }
    const int n = t->n, C = t->C, safe = t->safe;
    const size_t block = t->block;
    op_info* infos = (op_info*)malloc(sizeof(op_info) * (nops > 0 ? nops : 1));
    prepare_ops(t, ops, nops, infos);
    /* tip table: tip eigen vector repeated per category */
    double* tipcat = NULL;
    if (dad.is_tip) {
        tipcat = (double*)amalloc((size_t)(t->SU + 1) * block * 8);
        for (int s = 0; s <= t->SU; s++)
            for (int c = 0; c < C; c++)
                for (int i = 0; i < n; i++) tipcat[(size_t)s * block + c * n + i] = t->tip[s * n + i];
    }
    size_t limits[1024];
    compute_bounds(t->nthreads, t->npackets, t->Ppad, limits);
#pragma omp parallel for schedule(dynamic, 1) num_threads(t->nthreads)
    for (int pk = 0; pk < t->npackets; pk++) {
        const size_t lo = limits[pk], hi = limits[pk + 1];
        run_ops_packet(t, ops, nops, infos, lo, hi, pk);
        double* vtip = t->tipbuf + (size_t)pk * 3 * t->block * VS + 2 * t->block * VS;
        const double* pn = slot_ptr(t, node.id);
        const uint16_t* sn = scale_ptr(t, node.id);
        const uint16_t* sd = dad.is_tip ? NULL : scale_ptr(t, dad.id);
        for (size_t ptn = lo; ptn < hi; ptn += VS) {
            v4d* th = (v4d*)(t->theta + ptn * block);
            const v4d* vn = (const v4d*)(pn + ptn * block);
            if (dad.is_tip) {
                load_tip(t, tipcat, dad.id, ptn, vtip);
                const v4d* vt = (const v4d*)vtip;
                for (size_t i = 0; i < block; i++) th[i] = vt[i] * vn[i];
            } else {
                const v4d* vd = (const v4d*)(slot_ptr(t, dad.id) + ptn * block);
                for (size_t i = 0; i < block; i++) th[i] = vn[i] * vd[i];
            }
            for (int x = 0; x < VS; x++) {
                if (safe) {
                    uint16_t sum[16], mn = 0;
                    for (int c = 0; c < C; c++) {
                        sum[c] = (uint16_t)(sn[(ptn + x) * C + c] + (sd ? sd[(ptn + x) * C + c] : 0));
                        if (c == 0 || sum[c] < mn) mn = sum[c];
                    }
                    t->buffer_scale[ptn + x] = mn;
                    double* thx = t->theta + ptn * block + x;
                    for (int c = 0; c < C; c++) {
                        if (sum[c] == mn + 1) {
                            for (int i = 0; i < n; i++) thx[(c * n + i) * VS] *= SCALING_THRESHOLD;
                        } else if (sum[c] > mn + 1) {
                            for (int i = 0; i < n; i++) thx[(c * n + i) * VS] = 0.0;
                        }
                    }
                } else {
                    t->buffer_scale[ptn + x] = (double)(sn[ptn + x] + (sd ? sd[ptn + x] : 0));
                }
                t->buffer_scale[ptn + x] *= LOG_SCALING_THRESHOLD;
            }
        }
    }
    t->theta_ready = 1;
    free(tipcat);
    free(infos);
}

/* computeLikelihoodDervSIMD, normal branch-length model (tree/phylokernelnew.h:2233-2633) */
int orc_likelihood_derv(orc_tree* t, double length, double* df, double* ddf) {
    if (t->nonrev) {
        int st = 0;
        t->last_x = length;
        nr_eval(t, length, 1, df, ddf, &st);
        return st;
    }
    const int n = t->n, C = t->C;
    const size_t block = t->block;
    double* val0 = (double*)amalloc(block * 8);
    double* val1 = (double*)amalloc(block * 8);
    double* val2 = (double*)amalloc(block * 8);
    for (int c = 0; c < C; c++) {
        const double rate = t->rates[c], prop = t->props[c];
        for (int i = 0; i < n; i++) {
            const double cof = t->eval[i] * rate;
            const double v = exp(cof * length) * prop;
            const double v1 = cof * v;
            val0[c * n + i] = v;
            val1[c * n + i] = v1;
            val2[c * n + i] = cof * v1;
        }
    }
    size_t limits[1024];
    compute_bounds(t->nthreads, t->npackets, t->Ppad, limits);
    const size_t nobs = (size_t)(t->P - t->nasc);
    double all_df = 0.0, all_ddf = 0.0, all_pc = 0.0, all_dfc = 0.0, all_ddfc = 0.0;
#pragma omp parallel for schedule(dynamic, 1) num_threads(t->nthreads) \
    reduction(+ : all_df, all_ddf, all_pc, all_dfc, all_ddfc)
    for (int pk = 0; pk < t->npackets; pk++) {
        v4d my_df = vset1(0.0), my_ddf = vset1(0.0);
        for (size_t ptn = limits[pk]; ptn < limits[pk + 1]; ptn += VS) {
            const v4d* th = (const v4d*)(t->theta + ptn * block);
            v4d lh, d1, d2;
            dot_triple(val0, val1, val2, th, &lh, &d1, &d2, (int)block, n);
            lh = vabs(lh) + *(const v4d*)(t->invar + ptn);
            /* padding patterns carry freq 0 and contribute nothing */
            const v4d inv = vset1(1.0) / lh;
            const v4d df_frac = d1 * inv;
            const v4d ddf_frac = d2 * inv;
            v4d freq = *(const v4d*)(t->freq + ptn);
            for (int x = 0; x < VS; x++)
                if (ptn + x >= nobs && ptn + x < (size_t)t->P) {
                    /* an unobserved constant pattern: its probability and derivatives
                     * (tree/phylokernelnew.h:2497-2530) */
                    freq[x] = 0.0;
                    const double bs = t->buffer_scale[ptn + x];
                    all_pc += asc_lh(lh[x], bs);
                    all_dfc += asc_lh(d1[x], bs);
                    all_ddfc += asc_lh(d2[x], bs);
                }
            const v4d tmp1 = df_frac * freq;
            const v4d tmp2 = ddf_frac * freq;
            my_df += tmp1;
            my_ddf += tmp2 - tmp1 * df_frac;
        }
        all_df += hadd(my_df);
        all_ddf += hadd(my_ddf);
    }
    free(val0); free(val1); free(val2);
    if (t->nasc) {
        /* ASC_Lewis (tree/phylokernelnew.h:2620-2627) */
        const double pv = 1.0 - all_pc;
        const double df_frac = all_dfc / pv, ddf_frac = all_ddfc / pv;
        all_df += t->nsite * df_frac;
        all_ddf += t->nsite * (ddf_frac + df_frac * df_frac);
    }
    *df = all_df;
    *ddf = all_ddf;
    if (!isfinite(*df)) {
        *df = *ddf = 0.0;
        return 4;
    }
    return 0;
}

/* computeLikelihoodFromBufferSIMD (tree/phylokernelnew.h:3232-3428) */
double orc_likelihood_from_buffer(orc_tree* t, double length, double* pattern_lh) {
    if (t->nonrev) {
        const double lnl = nr_eval(t, length, 0, NULL, NULL, NULL);
        if (pattern_lh) memcpy(pattern_lh, t->pattern_lh, (size_t)(t->P - t->nasc) * 8);
        return lnl;
    }
    const int n = t->n, C = t->C;
    const size_t block = t->block;
    double* val0 = (double*)amalloc(block * 8);
    for (int c = 0; c < C; c++) {
        const double len = t->rates[c] * length;
        for (int i = 0; i < n; i++) val0[c * n + i] = exp(t->eval[i] * len) * t->props[c];
    }
    const size_t nobs = (size_t)(t->P - t->nasc);
    double all = 0.0, all_const = 0.0;
#pragma omp parallel for num_threads(t->nthreads) reduction(+ : all, all_const)
    for (size_t ptn = 0; ptn < (size_t)t->Ppad; ptn += VS) {
        const v4d* th = (const v4d*)(t->theta + ptn * block);
        v4d lh = dot_vec_s(val0, th, (int)block);
        lh = vabs(lh) + *(const v4d*)(t->invar + ptn);
        v4d acc = vset1(0.0);
        for (int x = 0; x < VS; x++) {
            if (ptn + x < nobs) {
                const double lp = log(fabs(lh[x])) + t->buffer_scale[ptn + x];
                t->pattern_lh[ptn + x] = lp;
                acc[x] = lp * t->freq[ptn + x];
            } else if (ptn + x < (size_t)t->P) {
                all_const += asc_lh(lh[x], t->buffer_scale[ptn + x]);
            }
        }
        all += hadd(acc);
    }
    free(val0);
    if (!isfinite(all)) {
        all = 0.0;
        for (size_t p = 0; p < nobs; p++) {
            if (!isfinite(t->pattern_lh[p])) t->pattern_lh[p] = LOG_SCALING_THRESHOLD * 4;
            all += t->pattern_lh[p] * t->freq[p];
        }
    }
    if (t->nasc) {
        /* ASC_Lewis (tree/phylokernelnew.h:3405-3424) */
        const double corr = log(1.0 - all_const);
        for (size_t p = 0; p < nobs; p++) t->pattern_lh[p] -= corr;
        all -= t->nsite * corr;
    }
    if (pattern_lh) memcpy(pattern_lh, t->pattern_lh, nobs * 8);
    return all;
}

/* Optimization::minimizeNewton (utils/optimization.cpp:422-499) */
double orc_minimize_newton(double x1, double xguess, double x2, double xacc, double* d2l,
                           int max_nr, orc_fderv fn, void* user, int* n_eval, int* status) {
    double df, dx, dxold, f, temp, xh, xl, rts, rts_old;
    int evals = 0;
    *status = 0;
    rts = xguess;
    if (rts < x1) rts = x1;
    if (rts > x2) rts = x2;
    fn(user, rts, &f, &df);
    evals++;
    *d2l = df;
    if (!isfinite(f) || !isfinite(df)) { *status = 4; if (n_eval) *n_eval = evals; return rts; }
    if (df >= 0.0 && fabs(f) < xacc) { if (n_eval) *n_eval = evals; return rts; }
    if (f < 0.0) { xl = rts; xh = x2; } else { xh = rts; xl = x1; }
    dx = dxold = fabs(xh - xl);
    for (int j = 1; j <= max_nr; j++) {
        rts_old = rts;
        if ((df <= 0.0) || (((rts - xh) * df - f) * ((rts - xl) * df - f) >= 0.0)) {
            dxold = dx;
            dx = 0.5 * (xh - xl);
            rts = xl + dx;
            *d2l = df;
            if (xl == rts) { if (n_eval) *n_eval = evals; return rts; }
        } else {
            dxold = dx;
            dx = f / df;
            temp = rts;
            rts -= dx;
            *d2l = df;
            if (temp == rts) { if (n_eval) *n_eval = evals; return rts; }
        }
        if (fabs(dx) < xacc || (j == max_nr)) { if (n_eval) *n_eval = evals; return rts_old; }
        fn(user, rts, &f, &df);
        evals++;
        if (!isfinite(f) || !isfinite(df)) { *status = 4; if (n_eval) *n_eval = evals; return rts; }
        if (df > 0.0 && fabs(f) < xacc) {
            *d2l = df;
            if (n_eval) *n_eval = evals;
            return rts;
        }
        if (f < 0.0) xl = rts;
        else if (f > 0.0) xh = rts;
    }
    *status = 7;
    *d2l = 0.0;
    if (n_eval) *n_eval = evals;
    return 0.0;
}

/* PhyloTree::computeFuncDerv (tree/phylotree.cpp:2529-2536): f = -dlnL/dt */
static void tree_fderv(void* user, double x, double* f, double* df) {
    orc_tree* t = (orc_tree*)user;
    double d1, d2;
    t->last_x = x;   /* the length computeFuncDerv leaves in the tree */
    if (orc_likelihood_derv(t, x, &d1, &d2) != 0 && !t->safe) {
        /* the normal kernel stops with outError (tree/phylokernelnew.h:2577-2579);
         * the safe one continues with df = ddf = 0 (2629-2632) */
        *f = *df = NAN;
        return;
    }
    *f = -d1;
    *df = -d2;
}

/* PhyloTree::optimizeOneBranch, Newton branch (tree/phylotree.cpp:2594-2644) */
double orc_optimize_branch(orc_tree* t, const orc_op* ops, int nops, orc_end node, orc_end dad,
                           double length, double min_len, double max_len, int max_nr,
                           double* lnl, int* n_derv, int* status) {
    orc_likelihood_buffer(t, ops, nops, node, dad);
    double d2l;
    int st = 0, evals = 0;
    double optx = orc_minimize_newton(min_len, length, max_len, min_len, &d2l, max_nr, tree_fderv, t,
                                      &evals, &st);
    if (st == 0 && optx > max_len * 0.95) {
        /* computeLikelihoodFromBuffer at the length the last derivative
         * evaluation left in the tree, then at the old length (2616-2625) */
        const double opt_lh = orc_likelihood_from_buffer(t, t->last_x, NULL);
        const double orig_lh = orc_likelihood_from_buffer(t, length, NULL);
        if (orig_lh > opt_lh) optx = length;
    }
    if (lnl) *lnl = orc_likelihood_from_buffer(t, optx, NULL);
    if (n_derv) *n_derv = evals;
    if (status) *status = st;
    return optx;
}

void orc_read_partial(orc_tree* t, int slot, double* partial, int32_t* scale) {
    const size_t block = t->block;
    const double* src = slot_ptr(t, slot);
    const uint16_t* sc = scale_ptr(t, slot);
    for (int p = 0; p < t->P; p++) {
        const size_t base = (size_t)(p / VS) * block * VS + (p % VS);
        if (partial)
            for (size_t i = 0; i < block; i++) partial[(size_t)p * block + i] = src[base + i * VS];
        if (scale) {
            if (t->safe)
                for (int c = 0; c < t->C; c++) scale[(size_t)p * t->C + c] = sc[(size_t)p * t->C + c];
            else
                scale[p] = sc[p];
        }
    }
}

/* ---- Fitch parsimony (tree/phylotreepars.cpp:32-300) ---------------------- */

/* the Fitch set of an observed tip state, as computePartialParsimonyFast fills
 * its bit planes (tree/phylotreepars.cpp:60-160) */
static uint64_t pars_tip_set(const orc_tree* t, int s) {
    const int n = t->n;
    const uint64_t all = n >= 64 ? ~(uint64_t)0 : (((uint64_t)1 << n) - 1);
    static const int ambi_aa[] = {2, 3, 5, 6, 9, 10};
    if (s < n) return (uint64_t)1 << s;
    if (s == t->SU) return all;
    if (n == 4) return (uint64_t)(s - 3) & all;
    if (n == 20) return ((uint64_t)1 << ambi_aa[2 * (s - 20)]) | ((uint64_t)1 << ambi_aa[2 * (s - 20) + 1]);
    return all;
}

static void pars_end(const orc_tree* t, int is_tip, int id, size_t p, uint64_t* set, int* ch) {
    if (is_tip) {
        *set = pars_tip_set(t, t->states[(size_t)id * t->Ppad + p]);
        *ch = 0;
    } else {
        *set = t->pars_set[(size_t)id * t->Ppad + p];
        *ch = t->pars_changes[(size_t)id * t->Ppad + p];
    }
}

/* computePartialParsimonyFast over the ops, then computeParsimonyBranchFast at
 * (node, dad): the number of changes of every pattern, weighted by its
 * frequency and summed */
double orc_parsimony_branch(orc_tree* t, const orc_op* ops, int nops, orc_end node, orc_end dad,
                            int32_t* ptn_score, double* branch_subst) {
    const size_t Pp = (size_t)t->Ppad;
    if (!t->pars_set) {
        t->pars_set = (uint64_t*)amalloc((size_t)t->nslots * Pp * 8);
        t->pars_changes = (uint16_t*)amalloc((size_t)t->nslots * Pp * 2);
    }
    double total = 0.0, on_branch = 0.0;
    for (size_t p = 0; p < (size_t)t->P; p++) {
        for (int o = 0; o < nops; o++) {
            const orc_op* op = &ops[o];
            uint64_t a, b;
            int ca, cb;
            pars_end(t, op->child[0].is_tip, op->child[0].id, p, &a, &ca);
            pars_end(t, op->child[1].is_tip, op->child[1].id, p, &b, &cb);
            const uint64_t both = a & b;
            t->pars_set[(size_t)op->dad_slot * Pp + p] = both ? both : (a | b);
            t->pars_changes[(size_t)op->dad_slot * Pp + p] = (uint16_t)(ca + cb + (both ? 0 : 1));
        }
        uint64_t a, b;
        int ca, cb;
        pars_end(t, node.is_tip, node.id, p, &a, &ca);
        pars_end(t, dad.is_tip, dad.id, p, &b, &cb);
        const int score = ca + cb + ((a & b) ? 0 : 1);
        if (ptn_score) ptn_score[p] = score;
        total += score * t->freq[p];
        if (!(a & b)) on_branch += t->freq[p];
    }
    if (branch_subst) *branch_subst = on_branch;
    return total;
}

/* The reference's bit-parallel form (computePartialParsimonyFast and
 * computeParsimonyBranchFast, tree/phylotreepars.cpp:32-300, the 256-bit SIMD
 * variant tree/phylokernel.h:1465-2094): every pattern with changes possible
 * (its tips' sets share no state; orderPatternByNumChars drops the invariant
 * ones, alignment/alignment.cpp:1178-1227) is expanded into `frequency` sites,
 * one bit per site in nstates planes; a node's planes are X & Y, or X | Y on
 * the sites where that is empty, counted by popcount; each node keeps its
 * subtree score. The timed CPU baseline of the device kernel; frequencies
 * must be integers. Returns -1 when they are not. */

static int pf_build(orc_tree* t) {
    const int n = t->n;
    int64_t nsites = 0;
    uint64_t* common = (uint64_t*)amalloc(sizeof(uint64_t) * t->P);
    for (int p = 0; p < t->P; p++) {
        if (t->freq[p] != floor(t->freq[p]) || t->freq[p] < 0) {
            free(common);
            return -1;
        }
        uint64_t c = ~(uint64_t)0;
        for (int s = 0; s < t->ntaxa; s++) c &= pars_tip_set(t, t->states[(size_t)s * t->Ppad + p]);
        common[p] = c;
        if (!c) nsites += (int64_t)t->freq[p];
    }
    const int nb = (int)((nsites + 255) / 256);
    t->pf_nblocks = nb < 1 ? 1 : nb;
    const size_t per = (size_t)t->pf_nblocks * n * 8;
    free(t->pf_tip);
    free(t->pf_node);
    free(t->pf_score);
    t->pf_tip = (uint32_t*)amalloc(sizeof(uint32_t) * per * t->ntaxa);
    t->pf_node = (uint32_t*)amalloc(sizeof(uint32_t) * per * t->nslots);
    t->pf_score = (int64_t*)amalloc(sizeof(int64_t) * t->nslots);
    for (int s = 0; s < t->ntaxa; s++) {
        uint32_t* x = t->pf_tip + per * s;
        int64_t site = 0;
        for (int p = 0; p < t->P; p++) {
            if (common[p]) continue;
            const uint64_t set = pars_tip_set(t, t->states[(size_t)s * t->Ppad + p]);
            for (int64_t j = 0; j < (int64_t)t->freq[p]; j++, site++) {
                const size_t b = (size_t)(site / 256), lane = (size_t)(site % 256) / 32;
                const uint32_t bit = 1u << (site % 32);
                for (int k = 0; k < n; k++)
                    if ((set >> k) & 1) x[(b * n + k) * 8 + lane] |= bit;
            }
        }
        /* trailing sites of the last block admit every state (no change) */
        for (; site < (int64_t)t->pf_nblocks * 256; site++) {
            const size_t b = (size_t)(site / 256), lane = (size_t)(site % 256) / 32;
            for (int k = 0; k < n; k++) x[(b * n + k) * 8 + lane] |= 1u << (site % 32);
        }
    }
    free(common);
    return 0;
}

static const uint32_t* pf_end(const orc_tree* t, int is_tip, int id, int64_t* score) {
    const size_t per = (size_t)t->pf_nblocks * t->n * 8;
    *score = is_tip ? 0 : t->pf_score[id];
    return is_tip ? t->pf_tip + per * id : t->pf_node + per * id;
}

double orc_parsimony_fast(orc_tree* t, const orc_op* ops, int nops, orc_end node, orc_end dad,
                          double* branch_subst) {
    if (!t->pf_nblocks && pf_build(t) != 0) return -1.0;
    const int n = t->n, nb = t->pf_nblocks, nth = t->nthreads > 0 ? t->nthreads : 1;
    const size_t per = (size_t)nb * n * 8;
    for (int o = 0; o < nops; o++) {
        int64_t sx, sy, score = 0;
        const uint32_t* x = pf_end(t, ops[o].child[0].is_tip, ops[o].child[0].id, &sx);
        const uint32_t* y = pf_end(t, ops[o].child[1].is_tip, ops[o].child[1].id, &sy);
        uint32_t* z = t->pf_node + per * ops[o].dad_slot;
#pragma omp parallel for num_threads(nth) reduction(+ : score) if (nb > 25)
        for (int b = 0; b < nb; b++) {
            const size_t off = (size_t)b * n * 8;
            uint32_t w[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            for (int k = 0; k < n; k++)
                for (int l = 0; l < 8; l++) {
                    z[off + k * 8 + l] = x[off + k * 8 + l] & y[off + k * 8 + l];
                    w[l] |= z[off + k * 8 + l];
                }
            for (int l = 0; l < 8; l++) {
                w[l] = ~w[l];
                score += __builtin_popcount(w[l]);
            }
            for (int k = 0; k < n; k++)
                for (int l = 0; l < 8; l++) z[off + k * 8 + l] |= w[l] & (x[off + k * 8 + l] | y[off + k * 8 + l]);
        }
        t->pf_score[ops[o].dad_slot] = score + sx + sy;
    }
    int64_t sx, sy, on = 0;
    const uint32_t* x = pf_end(t, node.is_tip, node.id, &sx);
    const uint32_t* y = pf_end(t, dad.is_tip, dad.id, &sy);
#pragma omp parallel for num_threads(nth) reduction(+ : on) if (nb > 25)
    for (int b = 0; b < nb; b++) {
        const size_t off = (size_t)b * n * 8;
        uint32_t w[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int k = 0; k < n; k++)
            for (int l = 0; l < 8; l++) w[l] |= x[off + k * 8 + l] & y[off + k * 8 + l];
        for (int l = 0; l < 8; l++) on += __builtin_popcount(~w[l]);
    }
    if (branch_subst) *branch_subst = (double)on;
    return (double)(sx + sy + on);
}

/* ---- Sankoff parsimony (tree/phylotreepars.cpp:464-892) ------------------ */

/* initCostMatrix / loadCostMatrixFile, then computeTipPartialParsimony's
 * tip costs (tree/phylotreepars.cpp:561-633): a state's row of the cost
 * matrix, an ambiguity set 0 inside and the least cost into it outside
 * (DNA bit mask + 3, protein B/Z/J), STATE_UNKNOWN 0 */
void orc_set_cost_matrix(orc_tree* t, const uint32_t* cost, int seq_type) {
    const int n = t->n;
    free(t->sk_cost);
    free(t->sk_tip);
    t->sk_cost = (uint32_t*)amalloc((size_t)n * n * 4);
    t->sk_tip = (uint32_t*)amalloc((size_t)(t->SU + 1) * n * 4);
    memcpy(t->sk_cost, cost, (size_t)n * n * 4);
    memcpy(t->sk_tip, cost, (size_t)n * n * 4);
    static const int ambi_aa[3] = {4 + 8, 32 + 64, 512 + 1024};
    for (int s = n; s <= t->SU; s++) {
        int mask = 0;
        if (s < t->SU) mask = seq_type == 0 ? s - n + 1 : (seq_type == 1 ? ambi_aa[s - 20] : 0);
        for (int i = 0; i < n; i++) {
            uint32_t v = 0;
            if (mask && !(mask & (1 << i))) {
                v = UINT32_MAX;
                for (int j = 0; j < n; j++)
                    if ((mask & (1 << j)) && cost[i * n + j] < v) v = cost[i * n + j];
            }
            t->sk_tip[(size_t)s * n + i] = v;
        }
    }
    if (!t->sk_partial) t->sk_partial = (uint32_t*)amalloc((size_t)t->nslots * t->Ppad * n * 4);
}

/* min_j child[j] + cost[i][j] */
static inline uint32_t sk_min(const uint32_t* cost_row, const uint32_t* child, int n) {
    uint32_t m = child[0] + cost_row[0];
    for (int j = 1; j < n; j++) {
        const uint32_t v = child[j] + cost_row[j];
        if (v < m) m = v;
    }
    return m;
}

/* computePartialParsimonySankoff for binary nodes (636-789) over the ops,
 * then computeParsimonyBranchSankoff (791-885) at (node, dad) */
double orc_parsimony_sankoff(orc_tree* t, const orc_op* ops, int nops, orc_end node, orc_end dad,
                             uint32_t* ptn_score, double* branch_subst) {
    const int n = t->n;
    const uint32_t* C = t->sk_cost;
    const size_t P = (size_t)t->P, Ppad = (size_t)t->Ppad;
    for (int o = 0; o < nops; o++) {
        const orc_op* op = &ops[o];
        int left = 0, right = 1;
        if (!op->child[0].is_tip && op->child[1].is_tip) { left = 1; right = 0; }
        const orc_child* L = &op->child[left];
        const orc_child* R = &op->child[right];
        uint32_t* dad_p = t->sk_partial + (size_t)op->dad_slot * Ppad * n;
        for (size_t p = 0; p < P; p++) {
            uint32_t* out = dad_p + p * n;
            const uint32_t* lt = L->is_tip ? t->sk_tip + (size_t)t->states[(size_t)L->id * Ppad + p] * n
                                           : t->sk_partial + ((size_t)L->id * Ppad + p) * n;
            const uint32_t* rt = R->is_tip ? t->sk_tip + (size_t)t->states[(size_t)R->id * Ppad + p] * n
                                           : t->sk_partial + ((size_t)R->id * Ppad + p) * n;
            for (int i = 0; i < n; i++) {
                const uint32_t a = L->is_tip ? lt[i] : sk_min(C + (size_t)i * n, lt, n);
                const uint32_t b = R->is_tip ? rt[i] : sk_min(C + (size_t)i * n, rt, n);
                out[i] = a + b;
            }
        }
    }
    double tree = 0.0, branch = 0.0;
    const uint32_t* pn = t->sk_partial + (size_t)node.id * Ppad * n;
    for (size_t p = 0; p < P; p++) {
        const uint32_t* nb = pn + p * n;
        uint32_t best, br;
        if (dad.is_tip) {
            const uint32_t* tp = t->sk_tip + (size_t)t->states[(size_t)dad.id * Ppad + p] * n;
            best = tp[0] + nb[0];
            br = tp[0];
            for (int i = 1; i < n; i++)
                if (tp[i] + nb[i] < best) { best = tp[i] + nb[i]; br = tp[i]; }
        } else {
            const uint32_t* db = t->sk_partial + ((size_t)dad.id * Ppad + p) * n;
            best = UINT32_MAX;
            br = UINT32_MAX;
            for (int i = 0; i < n; i++) {
                const uint32_t* row = C + (size_t)i * n;
                uint32_t m = db[0] + row[0], bs = row[0];
                for (int j = 1; j < n; j++)
                    if (db[j] + row[j] < m) { m = db[j] + row[j]; bs = row[j]; }
                m += nb[i];
                if (m < best) { best = m; br = bs; }
            }
        }
        if (ptn_score) ptn_score[p] = best;
        tree += (double)best * t->freq[p];
        branch += (double)br * t->freq[p];
    }
    if (branch_subst) *branch_subst = branch;
    return tree;
}
