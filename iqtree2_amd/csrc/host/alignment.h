// Alignment -> site patterns: character-to-state conversion, first-occurrence
// pattern compression, constant-pattern classification and the per-pattern
// +I term. Integer work, bit-exact with the reference semantics.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "model.h"

namespace plk {

constexpr int kStateInvalid = -1;

// Character -> state of one site (alignment/alignment.cpp:1415-1499).
int convert_state(SeqType t, char c);
// Codon triplet -> sense-codon state, STATE_UNKNOWN for any gap/ambiguity,
// kStateInvalid for unknown characters, -2 for a stop codon
// (alignment/alignment.cpp:1849-1879).
int convert_codon(char c1, char c2, char c3);

struct PatternSet {
    SeqType type = SEQ_DNA;
    int nseq = 0;
    int nsite = 0;       // alignment columns (codons for SEQ_CODON)
    int npattern = 0;
    int n_gaps_only = 0;
    std::vector<uint8_t> states;    // taxon-major [nseq][npattern]
    std::vector<int> freq;          // [npattern]
    std::vector<int> const_char;    // [npattern]
    std::vector<int> site_pattern;  // [nsite]
};

// Build patterns from nseq rows of nchar characters each (row-major, no
// separators). Columns are deduplicated in first-occurrence order
// (alignment/alignment.cpp:1725-1929, 1092-1124) and const_char follows
// Alignment::computeConst (alignment/alignment.cpp:925-1022).
// Returns 0, or a negative error (invalid character / stop codon).
int build_patterns(SeqType t, int nseq, int nchar, const char* chars, PatternSet& out);

// The unobserved patterns of Lewis's correction for parsimony-informative
// sites (ASC_INFORMATIVE; Alignment::getUnobservedConstPatterns,
// alignment/alignment.cpp:4181-4196, with generateSubsets and
// generateUninfPatterns 4106-4138): for every majority state r, every subset
// S of the other states in generateSubsets' order, and every placement of S's
// states into |S| distinct sequences in generateUninfPatterns' order, the
// pattern holding S's states there and r elsewhere (|S| <= nseq - 2, or
// nseq - 1 with r = 0). Rows of nseq states are appended to out; returns the
// count, or -1 when it would exceed max_patterns.
long uninformative_patterns(int nstates, int nseq, long max_patterns, std::vector<uint8_t>& out);

// const_char of one column of states (alignment/alignment.cpp:925-1022).
int compute_const_char(SeqType t, int nseq, const uint8_t* column);

// ptn_invar[p] for p < npattern and p_inv on padding up to padded_len
// (tree/phylotreesse.cpp:549-617).
void compute_ptn_invar(SeqType t, const PatternSet& ps, const double* pi, double p_inv,
                       int padded_len, double* invar);

}  // namespace plk
