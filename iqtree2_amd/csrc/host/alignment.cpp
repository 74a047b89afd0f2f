#include "alignment.h"

#include <cctype>
#include <cstring>
#include <unordered_map>

namespace plk {

static const char kProteinSymbols[] = "ARNDCQEGHILKMFPSTWYVX";

int convert_state(SeqType t, char c) {
    c = (char)std::toupper((unsigned char)c);
    const int su = seq_state_unknown(t);
    if (c == '?' || c == '-' || c == '.' || c == '~') return su;
    if (t == SEQ_DNA || t == SEQ_CODON) {
        switch (c) {
        case 'A': return 0;
        case 'C': return 1;
        case 'G': return 2;
        case 'T': case 'U': return 3;
        // ambiguity codes: bitmask over A=1,C=2,G=4,T=8, stored as mask + 3
        case 'R': return 1 + 4 + 3;
        case 'Y': return 2 + 8 + 3;
        case 'W': return 1 + 8 + 3;
        case 'S': return 2 + 4 + 3;
        case 'M': return 1 + 2 + 3;
        case 'K': return 4 + 8 + 3;
        case 'B': return 2 + 4 + 8 + 3;
        case 'H': return 1 + 2 + 8 + 3;
        case 'D': return 1 + 4 + 8 + 3;
        case 'V': return 1 + 2 + 4 + 3;
        case 'O': case 'N': case 'X': return t == SEQ_DNA ? su : 18;
        default: return kStateInvalid;
        }
    }
    // protein
    if (c == 'B') return 20;
    if (c == 'Z') return 21;
    if (c == 'J') return 22;
    if (c == '*' || c == 'U' || c == 'O') return su;
    const char* loc = std::strchr(kProteinSymbols, c);
    if (!loc || c == '\0') return kStateInvalid;
    const int s = (int)(loc - kProteinSymbols);
    return s < 20 ? s : su;
}

int convert_codon(char c1, char c2, char c3) {
    // each nucleotide through the DNA map; unknown nucleotides map to 18
    const int a = convert_state(SEQ_CODON, c1);
    const int b = convert_state(SEQ_CODON, c2);
    const int c = convert_state(SEQ_CODON, c3);
    if (a >= 0 && a < 4 && b >= 0 && b < 4 && c >= 0 && c < 4) {
        const int codon = a * 16 + b * 4 + c;
        const int sense = codon_to_sense()[codon];
        return sense < 0 ? -2 : sense;
    }
    if (a == kStateInvalid || b == kStateInvalid || c == kStateInvalid) return kStateInvalid;
    return seq_state_unknown(SEQ_CODON);
}

// set of model states compatible with an observed state
// (alignment/alignment.cpp Alignment::getAppearance)
static void appearance(SeqType t, int state, std::vector<char>& app) {
    const int n = seq_num_states(t), su = seq_state_unknown(t);
    if (state == su) {
        std::fill(app.begin(), app.end(), 1);
        return;
    }
    std::fill(app.begin(), app.end(), 0);
    if (state < n) {
        app[state] = 1;
        return;
    }
    if (t == SEQ_DNA) {
        const int mask = state - (n - 1);
        for (int i = 0; i < n; i++)
            if (mask & (1 << i)) app[i] = 1;
    } else if (t == SEQ_PROTEIN) {
        if (state >= 23) return;
        static const int ambi[3] = {4 + 8, 32 + 64, 512 + 1024};
        const int k = state - 20;
        for (int i = 0; i < 11; i++)
            if (ambi[k] & (1 << i)) app[i] = 1;
    }
}

int compute_const_char(SeqType t, int nseq, const uint8_t* column) {
    const int n = seq_num_states(t), su = seq_state_unknown(t);
    std::vector<char> all(n, 1), app(n);
    for (int s = 0; s < nseq; s++) {
        appearance(t, column[s], app);
        for (int j = 0; j < n; j++) all[j] &= app[j];
    }
    int count = 0;
    for (int j = 0; j < n; j++) count += all[j];
    if (count == 0) return su + 1;
    if (count == n) return su;
    if (count == 1) {
        for (int j = 0; j < n; j++)
            if (all[j]) return j;
    }
    if (t == SEQ_DNA) {
        int cc = n - 1;
        for (int j = 0; j < n; j++)
            if (all[j]) cc += (1 << j);
        return cc;
    }
    if (t == SEQ_PROTEIN) {
        if (all[2] && all[3]) return n;
        if (all[5] && all[6]) return n + 1;
        if (all[9] && all[10]) return n + 2;
    }
    return su + 1;
}

namespace {
struct ColumnHash {
    size_t operator()(const std::string& s) const { return std::hash<std::string>()(s); }
};
}  // namespace

int build_patterns(SeqType t, int nseq, int nchar, const char* chars, PatternSet& out) {
    const int step = (t == SEQ_CODON) ? 3 : 1;
    if (nchar % step != 0) return -3;
    const int nsite = nchar / step;
    const int su = seq_state_unknown(t);
    out = PatternSet();
    out.type = t;
    out.nseq = nseq;
    out.nsite = nsite;
    out.site_pattern.assign(nsite, -1);

    std::unordered_map<std::string, int, ColumnHash> index;
    index.reserve((size_t)nsite * 2);
    std::vector<std::string> columns;
    std::string col(nseq, '\0');
    int errors = 0;
    for (int site = 0; site < nsite; site++) {
        for (int s = 0; s < nseq; s++) {
            const char* row = chars + (size_t)s * nchar;
            int st;
            if (t == SEQ_CODON)
                st = convert_codon(row[3 * site], row[3 * site + 1], row[3 * site + 2]);
            else
                st = convert_state(t, row[site]);
            if (st < 0) {
                errors++;
                st = su;
            }
            col[s] = (char)st;
        }
        bool gaps_only = true;
        for (int s = 0; s < nseq; s++)
            if ((int)(uint8_t)col[s] != su) { gaps_only = false; break; }
        if (gaps_only) out.n_gaps_only++;
        auto it = index.find(col);
        if (it == index.end()) {
            const int id = (int)columns.size();
            index.emplace(col, id);
            columns.push_back(col);
            out.freq.push_back(1);
            out.site_pattern[site] = id;
        } else {
            out.freq[it->second] += 1;
            out.site_pattern[site] = it->second;
        }
    }
    if (errors) return -1;
    const int np = (int)columns.size();
    out.npattern = np;
    out.states.assign((size_t)nseq * np, 0);
    out.const_char.assign(np, 0);
    for (int p = 0; p < np; p++) {
        for (int s = 0; s < nseq; s++) out.states[(size_t)s * np + p] = (uint8_t)columns[p][s];
        out.const_char[p] =
            compute_const_char(t, nseq, reinterpret_cast<const uint8_t*>(columns[p].data()));
    }
    return 0;
}

void compute_ptn_invar(SeqType t, const PatternSet& ps, const double* pi, double p_inv,
                       int padded_len, double* invar) {
    const int n = seq_num_states(t), su = seq_state_unknown(t);
    for (int p = 0; p < padded_len; p++) invar[p] = 0.0;
    if (p_inv == 0.0) return;
    for (int p = 0; p < ps.npattern; p++) {
        const int cc = ps.const_char[p];
        if (cc > su) continue;
        if (cc == su) {
            invar[p] = p_inv;
        } else if (cc < n) {
            invar[p] = p_inv * pi[cc];
        } else if (t == SEQ_DNA) {
            double acc = 0.0;
            const int mask = cc - n + 1;
            for (int x = 0; x < n; x++)
                if (mask & (1 << x)) acc += pi[x];
            invar[p] = acc * p_inv;
        } else if (t == SEQ_PROTEIN) {
            static const int ambi[3] = {4 + 8, 32 + 64, 512 + 1024};
            double acc = 0.0;
            const int k = cc - n;
            for (int x = 0; x < 11; x++)
                if (ambi[k] & (1 << x)) acc += pi[x];
            invar[p] = acc * p_inv;
        }
    }
    for (int p = ps.npattern; p < padded_len; p++) invar[p] = p_inv;
}

namespace {

// generateUninfPatterns: the states of S placed, in order, into distinct
// sequences chosen in increasing index at every depth
bool place(int nseq, uint8_t repeat, const std::vector<uint8_t>& sub, std::vector<int>& pos, long max_patterns,
           long& count, std::vector<uint8_t>& out) {
    if (pos.size() == sub.size()) {
        if (++count > max_patterns) return false;
        const size_t base = out.size();
        out.resize(base + nseq, repeat);
        for (size_t i = 0; i < pos.size(); i++) out[base + pos[i]] = sub[i];
        return true;
    }
    for (int seq = 0; seq < nseq; seq++) {
        bool used = false;
        for (int q : pos) used = used || q == seq;
        if (used) continue;
        pos.push_back(seq);
        const bool ok = place(nseq, repeat, sub, pos, max_patterns, count, out);
        pos.pop_back();
        if (!ok) return false;
    }
    return true;
}

}  // namespace

long uninformative_patterns(int nstates, int nseq, long max_patterns, std::vector<uint8_t>& out) {
    if (nstates < 2 || nstates > 31 || nseq < 2) return -1;
    long count = 0;
    std::vector<int> pos;
    for (int r = 0; r < nstates; r++) {
        std::vector<uint8_t> rest;
        for (int s = 0; s < nstates; s++)
            if (s != r) rest.push_back((uint8_t)s);
        const uint64_t total = (uint64_t)1 << rest.size();
        for (uint64_t bin = 0; bin < total; bin++) {
            std::vector<uint8_t> sub;
            for (size_t i = 0; i < rest.size(); i++)
                if (bin & ((uint64_t)1 << i)) sub.push_back(rest[i]);
            const int k = (int)sub.size();
            if (!(k < nseq - 1 || (k == nseq - 1 && r == 0))) continue;
            if (!place(nseq, (uint8_t)r, sub, pos, max_patterns, count, out)) return -1;
        }
    }
    return count;
}

}  // namespace plk
