// Seeded pseudo-random generator for the synthetic workloads.
//
// splitmix64 expands the user seed into the 256-bit state of xoshiro256**
// (Blackman & Vigna, "Scrambled linear pseudorandom number generators", 2018).
// All variates are derived from 53-bit uniforms, so a given seed yields the
// same stream on every x86-64 host.
#pragma once
#include <cmath>
#include <cstdint>

namespace plk {

class Rng {
public:
    explicit Rng(uint64_t seed) {
        uint64_t z = seed;
        for (int i = 0; i < 4; i++) s_[i] = splitmix(z);
    }

    uint64_t next() {
        const uint64_t result = rotl(s_[1] * 5, 7) * 9;
        const uint64_t t = s_[1] << 17;
        s_[2] ^= s_[0];
        s_[3] ^= s_[1];
        s_[1] ^= s_[2];
        s_[0] ^= s_[3];
        s_[2] ^= t;
        s_[3] = rotl(s_[3], 45);
        return result;
    }

    // uniform in [0, 1)
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
    double uniform(double lo, double hi) { return lo + (hi - lo) * uniform(); }
    // uniform integer in [0, n)
    uint64_t below(uint64_t n) { return (uint64_t)(uniform() * (double)n); }

    double exponential(double mean) { return -mean * std::log1p(-uniform()); }

    double normal() {
        // Marsaglia polar method
        if (have_spare_) { have_spare_ = false; return spare_; }
        double u, v, s;
        do {
            u = 2.0 * uniform() - 1.0;
            v = 2.0 * uniform() - 1.0;
            s = u * u + v * v;
        } while (s >= 1.0 || s == 0.0);
        double m = std::sqrt(-2.0 * std::log(s) / s);
        spare_ = v * m;
        have_spare_ = true;
        return u * m;
    }

    // Gamma(shape, scale): Marsaglia & Tsang (2000); shape < 1 via the
    // U^(1/shape) boost.
    double gamma(double shape, double scale) {
        if (shape < 1.0) {
            double g = gamma(shape + 1.0, 1.0);
            double u = uniform();
            while (u == 0.0) u = uniform();
            return scale * g * std::pow(u, 1.0 / shape);
        }
        const double d = shape - 1.0 / 3.0, c = 1.0 / std::sqrt(9.0 * d);
        for (;;) {
            double x, v;
            do {
                x = normal();
                v = 1.0 + c * x;
            } while (v <= 0.0);
            v = v * v * v;
            double u = uniform();
            if (u < 1.0 - 0.0331 * x * x * x * x) return scale * d * v;
            if (u > 0.0 && std::log(u) < 0.5 * x * x + d * (1.0 - v + std::log(v)))
                return scale * d * v;
        }
    }

    void dirichlet(int k, const double* alpha, double* out) {
        double sum = 0.0;
        for (int i = 0; i < k; i++) { out[i] = gamma(alpha[i], 1.0); sum += out[i]; }
        for (int i = 0; i < k; i++) out[i] /= sum;
    }

    double lognormal(double mu, double sigma) { return std::exp(mu + sigma * normal()); }

    // draw an index from a cumulative distribution cdf[0..n-1] (cdf[n-1] ~ 1)
    int categorical_cdf(int n, const double* cdf) {
        double u = uniform() * cdf[n - 1];
        int lo = 0, hi = n - 1;
        while (lo < hi) {
            int mid = (lo + hi) >> 1;
            if (u < cdf[mid]) hi = mid; else lo = mid + 1;
        }
        return lo;
    }

private:
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    static uint64_t splitmix(uint64_t& x) {
        uint64_t z = (x += 0x9e3779b97f4a7c15ULL);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
        return z ^ (z >> 31);
    }
    uint64_t s_[4];
    bool have_spare_ = false;
    double spare_ = 0.0;
};

}  // namespace plk
