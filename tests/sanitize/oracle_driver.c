/* Runs the oracle (oracle/oracle.c) on a workload written by
 * tests/test_sanitizers.py under AddressSanitizer and
 * UndefinedBehaviorSanitizer, normal and safe kernels, the reversible and the
 * state-basis paths, Newton and both parsimony forms; prints the results for
 * the test to compare with the ordinary build. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static void* slurp(FILE* f, size_t n) {
    void* p = malloc(n ? n : 1);
    if (n && fread(p, 1, n, f) != n) { fprintf(stderr, "short read\n"); exit(2); }
    return p;
}

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    FILE* f = fopen(argv[1], "rb");
    if (!f) return 2;
    int hdr[7];
    if (fread(hdr, sizeof(int), 7, f) != 7) return 2;
    const int ntaxa = hdr[0], n = hdr[1], C = hdr[2], P = hdr[3], nslots = hdr[4], SU = hdr[5], nops = hdr[6];
    uint8_t* states = slurp(f, (size_t)ntaxa * P);
    double* freq = slurp(f, 8 * (size_t)P);
    double* invar = slurp(f, 8 * (size_t)P);
    double* eval = slurp(f, 8 * (size_t)n);
    double* evec = slurp(f, 8 * (size_t)n * n);
    double* inv = slurp(f, 8 * (size_t)n * n);
    double* tip = slurp(f, 8 * (size_t)(SU + 1) * n);
    double* rates = slurp(f, 8 * (size_t)C);
    double* props = slurp(f, 8 * (size_t)C);
    orc_op* ops = slurp(f, sizeof(orc_op) * nops);
    orc_end ends[2];
    if (fread(ends, sizeof(orc_end), 2, f) != 2) return 2;
    double length;
    if (fread(&length, 8, 1, f) != 1) return 2;
    double* ind = slurp(f, 8 * (size_t)(SU + 1) * n);
    double* Q = slurp(f, 8 * (size_t)n * n);
    double* pi = slurp(f, 8 * (size_t)n);
    fclose(f);
    for (int safe = 0; safe < 2; safe++) {
        orc_tree* t = orc_create(ntaxa, n, C, P, nslots, SU, safe, 2);
        orc_set_patterns(t, states, freq, invar);
        orc_set_model(t, eval, evec, inv, tip, rates, props);
        double* plh = malloc(8 * (size_t)P);
        int st = 0;
        const double lnl = orc_likelihood_branch(t, ops, nops, ends[0], ends[1], length, plh, &st);
        orc_likelihood_buffer(t, ops, nops, ends[0], ends[1]);
        double df, ddf;
        orc_likelihood_derv(t, length, &df, &ddf);
        const double lfb = orc_likelihood_from_buffer(t, 1.5 * length, NULL);
        double olnl;
        int nd, ost;
        const double x = orc_optimize_branch(t, ops, nops, ends[0], ends[1], length, 1e-6, 10.0, 100, &olnl, &nd, &ost);
        double sub = 0.0;
        int32_t* ps = malloc(4 * (size_t)P);
        const double pars = orc_parsimony_branch(t, ops, nops, ends[0], ends[1], ps, &sub);
        const double parsf = orc_parsimony_fast(t, ops, nops, ends[0], ends[1], NULL);
        /* the state-basis path with the same model as one real component */
        double* ce = calloc(2 * (size_t)n, 8);
        double* cv = calloc(2 * (size_t)n * n, 8);
        double* ci = calloc(2 * (size_t)n * n, 8);
        for (int i = 0; i < n; i++) ce[2 * i] = eval[i];
        for (int i = 0; i < n * n; i++) { cv[2 * i] = evec[i]; ci[2 * i] = inv[i]; }
        const double one = 1.0;
        const int sm = orc_set_model_nonrev(t, 1, 0, ce, cv, ci, Q, pi, &one, rates, props, C, ind, 0);
        const double nlnl = orc_likelihood_branch(t, ops, nops, ends[0], ends[1], length, NULL, NULL);
        orc_likelihood_buffer(t, ops, nops, ends[0], ends[1]);
        double ndf, nddf;
        orc_likelihood_derv(t, length, &ndf, &nddf);
        printf("%d %.17g %.17g %.17g %.17g %.17g %d %.17g %.17g %d %.17g %.17g %.17g\n", safe, lnl, df, ddf, lfb, x, nd,
               pars, parsf, sm, nlnl, ndf, nddf);
        free(plh); free(ps); free(ce); free(cv); free(ci);
        orc_destroy(t);
    }
    free(states); free(freq); free(invar); free(eval); free(evec); free(inv); free(tip); free(rates);
    free(props); free(ops); free(ind); free(Q); free(pi);
    return 0;
}
