// Substitution-model inputs (FP64, host). See model.h for the contract of
// each function and the reference code it stands in for.
#include "model.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>

namespace plk {

const char kGeneticCodeACGT[65] =
    "KNKNTTTTRSRSIIMIQHQHPPPPRRRRLLLLEDEDAAAAGGGGVVVV*Y*YSSSS*CWCLFLF";

namespace {
struct CodonTables {
    int to_sense[64];
    int to_codon[61];
    CodonTables() {
        int k = 0;
        for (int c = 0; c < 64; c++) {
            if (kGeneticCodeACGT[c] == '*') {
                to_sense[c] = -1;
            } else {
                to_sense[c] = k;
                to_codon[k] = c;
                k++;
            }
        }
    }
};
const CodonTables& codon_tables() {
    static CodonTables t;
    return t;
}
}  // namespace

const int* codon_to_sense() { return codon_tables().to_sense; }
const int* sense_to_codon() { return codon_tables().to_codon; }

int seq_num_states(SeqType t) {
    switch (t) {
    case SEQ_DNA: return 4;
    case SEQ_PROTEIN: return 20;
    case SEQ_CODON: return 61;
    }
    return -1;
}

int seq_state_unknown(SeqType t) {
    switch (t) {
    case SEQ_DNA: return 18;
    case SEQ_PROTEIN: return 23;
    case SEQ_CODON: return 61;
    }
    return -1;
}

void build_reversible_q(int n, const double* exch_upper, const double* pi, double* Q) {
    int k = 0;
    for (int i = 0; i < n; i++) Q[i * n + i] = 0.0;
    for (int i = 0; i < n; i++)
        for (int j = i + 1; j < n; j++, k++) {
            Q[i * n + j] = exch_upper[k] * pi[j];
            Q[j * n + i] = exch_upper[k] * pi[i];
        }
    double mean_rate = 0.0;
    for (int i = 0; i < n; i++) {
        double row = 0.0;
        for (int j = 0; j < n; j++)
            if (j != i) row += Q[i * n + j];
        Q[i * n + i] = -row;
        mean_rate += pi[i] * row;
    }
    const double scale = 1.0 / mean_rate;
    for (int i = 0; i < n * n; i++) Q[i] *= scale;
}

// Cyclic Jacobi rotations on a symmetric matrix A (destroyed). On return
// w holds the eigenvalues and V (row-major, column i = eigenvector i).
static int jacobi_eigen(int n, std::vector<double>& A, double* w, std::vector<double>& V) {
    V.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; i++) V[i * n + i] = 1.0;
    int sweep;
    for (sweep = 0; sweep < 100; sweep++) {
        double off = 0.0, diag = 0.0;
        for (int p = 0; p < n; p++) {
            diag += A[p * n + p] * A[p * n + p];
            for (int q = p + 1; q < n; q++) off += A[p * n + q] * A[p * n + q];
        }
        if (off <= 1e-34 * std::max(diag, 1e-300)) break;
        for (int p = 0; p < n - 1; p++)
            for (int q = p + 1; q < n; q++) {
                const double apq = A[p * n + q];
                if (apq == 0.0) continue;
                const double app = A[p * n + p], aqq = A[q * n + q];
                const double theta = (aqq - app) / (2.0 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) /
                                 (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; k++) {
                    const double akp = A[k * n + p], akq = A[k * n + q];
                    A[k * n + p] = c * akp - s * akq;
                    A[k * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; k++) {
                    const double apk = A[p * n + k], aqk = A[q * n + k];
                    A[p * n + k] = c * apk - s * aqk;
                    A[q * n + k] = s * apk + c * aqk;
                }
                A[p * n + q] = A[q * n + p] = 0.0;
                for (int k = 0; k < n; k++) {
                    const double vkp = V[k * n + p], vkq = V[k * n + q];
                    V[k * n + p] = c * vkp - s * vkq;
                    V[k * n + q] = s * vkp + c * vkq;
                }
            }
    }
    if (sweep == 100) return -1;
    for (int i = 0; i < n; i++) w[i] = A[i * n + i];
    return sweep;
}

int decompose_reversible(int n, const double* Q, const double* pi, double* eval,
                         double* evec, double* inv_evec) {
    std::vector<double> S((size_t)n * n), V;
    std::vector<double> sq(n);
    for (int i = 0; i < n; i++) sq[i] = std::sqrt(pi[i]);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) S[i * n + j] = sq[i] * Q[i * n + j] / sq[j];
    // exact symmetrisation of the rounding noise
    for (int i = 0; i < n; i++)
        for (int j = i + 1; j < n; j++) {
            const double m = 0.5 * (S[i * n + j] + S[j * n + i]);
            S[i * n + j] = S[j * n + i] = m;
        }
    std::vector<double> w(n);
    int sweeps = jacobi_eigen(n, S, w.data(), V);
    if (sweeps < 0) return -1;
    std::vector<int> order(n);
    for (int i = 0; i < n; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return w[a] < w[b]; });
    for (int i = 0; i < n; i++) {
        const int src = order[i];
        eval[i] = w[src];
        for (int x = 0; x < n; x++) {
            evec[x * n + i] = V[x * n + src] / sq[x];
            inv_evec[i * n + x] = V[x * n + src] * sq[x];
        }
    }
    return sweeps;
}

void gy94_exchangeabilities(double kappa, double omega, double* exch_upper) {
    const int* s2c = sense_to_codon();
    int k = 0;
    for (int i = 0; i < 61; i++)
        for (int j = i + 1; j < 61; j++, k++) {
            const int a = s2c[i], b = s2c[j];
            int ndiff = 0, pos = -1;
            for (int p = 0; p < 3; p++) {
                const int sh = 2 * (2 - p);
                if (((a >> sh) & 3) != ((b >> sh) & 3)) { ndiff++; pos = sh; }
            }
            double r = 0.0;
            if (ndiff == 1) {
                const int na = (a >> pos) & 3, nb = (b >> pos) & 3;
                // A=0,C=1,G=2,T=3: transitions are A<->G and C<->T
                const bool transition = (na ^ nb) == 2;
                r = 1.0;
                if (transition) r *= kappa;
                if (kGeneticCodeACGT[a] != kGeneticCodeACGT[b]) r *= omega;
            }
            exch_upper[k] = r;
        }
}

// ---------------------------------------------------------------------------
// Incomplete gamma: power series for x < a + 1, Lentz continued fraction
// otherwise (the standard pairing, e.g. Press et al., Numerical Recipes 6.2).
static double gamma_p_series(double a, double x) {
    double ap = a, sum = 1.0 / a, del = sum;
    for (int n = 0; n < 10000; n++) {
        ap += 1.0;
        del *= x / ap;
        sum += del;
        if (std::fabs(del) < std::fabs(sum) * 1e-17) break;
    }
    return sum * std::exp(-x + a * std::log(x) - std::lgamma(a));
}

static double gamma_q_contfrac(double a, double x) {
    const double tiny = 1e-300;
    double b = x + 1.0 - a, c = 1.0 / tiny, d = 1.0 / b, h = d;
    for (int i = 1; i < 10000; i++) {
        const double an = -i * (i - a);
        b += 2.0;
        d = an * d + b;
        if (std::fabs(d) < tiny) d = tiny;
        c = b + an / c;
        if (std::fabs(c) < tiny) c = tiny;
        d = 1.0 / d;
        const double del = d * c;
        h *= del;
        if (std::fabs(del - 1.0) < 1e-17) break;
    }
    return std::exp(-x + a * std::log(x) - std::lgamma(a)) * h;
}

double gamma_p(double a, double x) {
    if (x <= 0.0) return 0.0;
    if (x < a + 1.0) return gamma_p_series(a, x);
    return 1.0 - gamma_q_contfrac(a, x);
}

double gamma_p_inv(double a, double p) {
    if (p <= 0.0) return 0.0;
    if (p >= 1.0) return INFINITY;
    // bracket, then safeguarded Newton on P(a,x) - p
    double lo = 0.0, hi = std::max(1.0, a);
    while (gamma_p(a, hi) < p) { lo = hi; hi *= 2.0; }
    double x = 0.5 * (lo + hi);
    for (int it = 0; it < 200; it++) {
        const double f = gamma_p(a, x) - p;
        if (f < 0) lo = x; else hi = x;
        const double dens = std::exp(-x + (a - 1.0) * std::log(x) - std::lgamma(a));
        double xn = x - f / dens;
        if (!(xn > lo && xn < hi)) xn = 0.5 * (lo + hi);
        if (std::fabs(xn - x) <= 1e-16 * std::max(x, 1e-300)) { x = xn; break; }
        x = xn;
        if (hi - lo <= 1e-16 * hi) break;
    }
    return x;
}

void gamma_mean_rates(int ncat, double alpha, double p_inv, double* rates, double* props) {
    if (ncat == 1) {
        rates[0] = 1.0 / (1.0 - p_inv);
        props[0] = 1.0 - p_inv;
        return;
    }
    // cut points of Gamma(alpha, 1/alpha) at i/ncat, then category means via
    // P(alpha + 1, alpha * cut) (Yang 1994, eq. 10)
    std::vector<double> cdf1(ncat - 1);
    for (int i = 0; i < ncat - 1; i++) {
        const double cut = gamma_p_inv(alpha, (i + 1.0) / ncat) / alpha;
        cdf1[i] = gamma_p(alpha + 1.0, cut * alpha);
    }
    rates[0] = cdf1[0] * ncat;
    rates[ncat - 1] = (1.0 - cdf1[ncat - 2]) * ncat;
    for (int i = 1; i < ncat - 1; i++) rates[i] = (cdf1[i] - cdf1[i - 1]) * ncat;
    double sum = 0.0;
    for (int i = 0; i < ncat; i++) sum += rates[i];
    const double target = ncat / (1.0 - p_inv);
    for (int i = 0; i < ncat; i++) {
        rates[i] *= target / sum;
        props[i] = (1.0 - p_inv) / ncat;
    }
}

void tip_indicator(SeqType t, int state, double* ind) {
    const int n = seq_num_states(t);
    const int su = seq_state_unknown(t);
    for (int x = 0; x < n; x++) ind[x] = 0.0;
    if (state < n) {
        ind[state] = 1.0;
        return;
    }
    if (state >= su) {
        for (int x = 0; x < n; x++) ind[x] = 1.0;
        return;
    }
    if (t == SEQ_DNA) {
        // ambiguity code: bitmask over A,C,G,T stored as bitmask + 3
        const int mask = state - n + 1;
        for (int x = 0; x < n; x++)
            if (mask & (1 << x)) ind[x] = 1.0;
    } else if (t == SEQ_PROTEIN) {
        // B = N|D, Z = Q|E, J = I|L
        static const int pairs[3][2] = {{2, 3}, {5, 6}, {9, 10}};
        const int k = state - n;
        ind[pairs[k][0]] = 1.0;
        ind[pairs[k][1]] = 1.0;
    } else {
        for (int x = 0; x < n; x++) ind[x] = 1.0;
    }
}

void tip_eigen_vectors(SeqType t, const double* inv_evec, double* tip) {
    const int n = seq_num_states(t);
    const int su = seq_state_unknown(t);
    std::vector<double> ind(n);
    for (int s = 0; s <= su; s++) {
        tip_indicator(t, s, ind.data());
        for (int i = 0; i < n; i++) {
            double acc = 0.0;
            for (int x = 0; x < n; x++) acc += inv_evec[i * n + x] * ind[x];
            tip[s * n + i] = acc;
        }
    }
}

}  // namespace plk
