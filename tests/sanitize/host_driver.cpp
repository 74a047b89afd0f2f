// Runs the device-free host C++ (pattern compression, models, eigensystems,
// discrete Gamma, synthetic workloads, the uninformative-pattern set, the
// Brent and Newton state machines) under AddressSanitizer and
// UndefinedBehaviorSanitizer; built and run by tests/test_sanitizers.py.
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "alignment.h"
#include "brent.h"
#include "model.h"
#include "newton.h"
#include "synth.h"

#define CHECK(c)                                                      \
    do {                                                              \
        if (!(c)) {                                                   \
            std::fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
            return 1;                                                 \
        }                                                             \
    } while (0)

int main() {
    using namespace plk;
    for (int type = 0; type < 3; type++)
        for (int gaps = 0; gaps < 2; gaps++)
            for (int ncat : {1, 4, 8}) {
                WorkloadSpec s;
                s.type = type;
                s.ntaxa = type == 2 ? 8 : 24;
                s.nsite = 400;
                s.ncat = ncat;
                s.p_inv = 0.2;
                s.gap_frac = gaps ? 0.1 : 0.0;
                s.seed = 11 + type * 7 + ncat;
                Workload w;
                CHECK(make_workload(s, w) == 0);
                long total = 0;
                for (int f : w.patterns.freq) total += f;
                CHECK(total == w.patterns.nsite);
                CHECK((int)w.invar.size() == w.patterns.npattern);
            }
    // pattern compression with ambiguity codes, gaps and unknowns
    const std::string rows = "ACGTRYKMSWBDHVN-?ACGT"
                             "ACGTACGTACGTACGTACGTA"
                             "ACGTRYKMSWBDHVN-?ACGA";
    PatternSet ps;
    CHECK(build_patterns(SEQ_DNA, 3, 21, rows.data(), ps) == 0);
    CHECK(ps.npattern > 0 && (int)ps.site_pattern.size() == 21);
    PatternSet bad;
    CHECK(build_patterns(SEQ_DNA, 1, 3, "AC!", bad) < 0);
    // reversible models and their eigensystems, every state count
    for (int n : {4, 20, 61}) {
        std::vector<double> exch(n * (n - 1) / 2), pi(n), Q(n * n), ev(n), U(n * n), Ui(n * n);
        for (size_t i = 0; i < exch.size(); i++) exch[i] = 0.5 + (i % 7) * 0.3;
        double s = 0.0;
        for (int i = 0; i < n; i++) s += (pi[i] = 1.0 + (i % 5));
        for (double& p : pi) p /= s;
        build_reversible_q(n, exch.data(), pi.data(), Q.data());
        CHECK(decompose_reversible(n, Q.data(), pi.data(), ev.data(), U.data(), Ui.data()) >= 0);
        std::vector<double> tip((seq_state_unknown(n == 4 ? SEQ_DNA : n == 20 ? SEQ_PROTEIN : SEQ_CODON) + 1) * n);
        tip_eigen_vectors(n == 4 ? SEQ_DNA : n == 20 ? SEQ_PROTEIN : SEQ_CODON, Ui.data(), tip.data());
    }
    std::vector<double> gexch(61 * 60 / 2);
    gy94_exchangeabilities(2.0, 0.3, gexch.data());
    for (int ncat : {1, 2, 4, 8})
        for (double a : {0.02, 0.5, 3.0, 1000.0}) {
            std::vector<double> r(ncat), p(ncat);
            gamma_mean_rates(ncat, a, 0.1, r.data(), p.data());
            CHECK(std::isfinite(r[ncat - 1]));
        }
    CHECK(std::fabs(gamma_p(2.5, gamma_p_inv(2.5, 0.3)) - 0.3) < 1e-10);
    // the uninformative-pattern set and its limit
    std::vector<uint8_t> un;
    CHECK(uninformative_patterns(4, 5, 1 << 20, un) > 0);
    CHECK(uninformative_patterns(20, 3, 1 << 20, un) > 0);
    CHECK(uninformative_patterns(20, 12, 1000, un) == -1);
    // Brent on a parabola, Newton on a quadratic's derivative
    BrentRun br;
    br.start(0.02, 0.5, 1000.0, 1e-3);
    while (!br.done()) br.feed((br.x() - 1.7) * (br.x() - 1.7));
    CHECK(std::fabs(br.result() - 1.7) < 1e-2);
    NewtonRun nr;
    nr.start(0.1, 1e-6, 10.0, 1e-6, 100);
    while (!nr.done()) nr.feed(2.0 * (nr.x() - 0.3), 2.0);
    CHECK(std::fabs(nr.result() - 0.3) < 1e-5);
    std::printf("host ok\n");
    return 0;
}
