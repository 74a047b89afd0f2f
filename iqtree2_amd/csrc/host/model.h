// Host-side substitution-model inputs for the likelihood engine: reversible
// rate matrices, their eigen-decomposition, discrete-Gamma category rates and
// the eigen-basis tip vectors. Everything here is FP64.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace plk {

enum SeqType { SEQ_DNA = 0, SEQ_PROTEIN = 1, SEQ_CODON = 2 };

// Number of model states and the alignment's STATE_UNKNOWN per sequence type
// (alignment/alignment.cpp:736-747: DNA 18, protein 23, otherwise num_states).
int seq_num_states(SeqType t);
int seq_state_unknown(SeqType t);

// Standard genetic code indexed by 16*n1 + 4*n2 + n3 with A,C,G,T = 0..3.
extern const char kGeneticCodeACGT[65];
// Sense-codon table: codon64 -> sense index (0..60) or -1 for stop codons,
// and the inverse sense index -> codon64.
const int* codon_to_sense();
const int* sense_to_codon();

// Reversible rate matrix Q (row-major n x n) from exchangeabilities given as
// the upper triangle in row-major order ((0,1),(0,2),...,(n-2,n-1)) and state
// frequencies pi: Q_ij = r_ij * pi_j, rows sum to zero, mean rate -sum_i pi_i
// Q_ii normalised to 1 (model/modelmarkov.cpp:1499-1512).
void build_reversible_q(int n, const double* exch_upper, const double* pi, double* Q);

// Eigen-decomposition of a reversible Q through the symmetric matrix
// Pi^{1/2} Q Pi^{-1/2} (model/modelmarkov.cpp:1514-1556). Eigenvalues are
// returned in ascending order; evec = Pi^{-1/2} V (row x, column i) and
// inv_evec = V^T Pi^{1/2}, so Q = evec * diag(eval) * inv_evec.
// Returns the number of Jacobi sweeps used, or -1 on failure.
int decompose_reversible(int n, const double* Q, const double* pi, double* eval,
                         double* evec, double* inv_evec);

// GY94-style codon exchangeabilities (upper triangle, 61 sense codons):
// codons differing at one position get 1, times kappa for a transition and
// omega for a non-synonymous change (model/modelcodon.cpp:845-870).
void gy94_exchangeabilities(double kappa, double omega, double* exch_upper);

// Discrete-Gamma mean category rates (Yang 1994; model/rategamma.cpp:158-171),
// rescaled so sum(rates) = ncat / (1 - p_inv) as RateGammaInvar does
// (model/rategammainvar.cpp:25-42), and category weights (1 - p_inv) / ncat.
void gamma_mean_rates(int ncat, double alpha, double p_inv, double* rates, double* props);

// Regularised lower incomplete gamma P(a, x) and its inverse in x.
double gamma_p(double a, double x);
double gamma_p_inv(double a, double p);

// Eigen-basis tip vectors tip[s][i] = sum_x inv_evec[i][x] * ind_s(x) for
// s = 0..state_unknown, where ind_s is the indicator (or ambiguity set, or
// all-ones for unknown) of observed state s (tree/phylotreesse.cpp:455-466,
// model/modelsubst.cpp:182-207, model/modeldna.cpp:462-475,
// model/modelprotein.cpp:1149-1163).
void tip_indicator(SeqType t, int state, double* ind);
void tip_eigen_vectors(SeqType t, const double* inv_evec, double* tip);

}  // namespace plk
