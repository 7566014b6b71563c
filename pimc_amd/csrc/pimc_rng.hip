// pimc_rng.hip -- host-side MTRand (include/MersenneTwister.h) and the seeded
// worldline generator used to build the measured workloads.
//
// The stream is MT19937 with init_genrand seeding; rand() = u/4294967295,
// randExc() = u/4294967296, randNorm = Marsaglia polar keeping the first
// variate (:290-363).  Slice 0 is Prism::randPosition (src/container.cpp:155-160);
// kind 0 adds randNorm(prev, sqrt(2 lambda tau)) per axis and putInBC, kind 1
// closes a Gaussian bridge from slice 0 back to slice 0 with the staging
// construction (src/move.cpp:270-291).
#include <cmath>
#include <cstdint>
#include <vector>

#include "../../include/pimc_gpu.h"

#pragma STDC FP_CONTRACT OFF

namespace {

struct Mt {
    static constexpr int N = 624, M = 397;
    uint32_t s[N];
    int left = 0, next = 0;
    explicit Mt(uint32_t seed) {
        s[0] = seed;
        for (int i = 1; i < N; ++i) s[i] = 1812433253u * (s[i - 1] ^ (s[i - 1] >> 30)) + (uint32_t)i;
        reload();
    }
    static uint32_t twist(uint32_t m, uint32_t a, uint32_t b) {
        return m ^ (((a & 0x80000000u) | (b & 0x7fffffffu)) >> 1) ^ ((b & 1u) ? 0x9908b0dfu : 0u);
    }
    void reload() {
        int i = 0;
        for (; i < N - M; ++i) s[i] = twist(s[i + M], s[i], s[i + 1]);
        for (; i < N - 1; ++i) s[i] = twist(s[i + M - N], s[i], s[i + 1]);
        s[N - 1] = twist(s[M - 1], s[N - 1], s[0]);
        left = N;
        next = 0;
    }
    uint32_t u32() {
        if (left == 0) reload();
        --left;
        uint32_t y = s[next++];
        y ^= (y >> 11);
        y ^= (y << 7) & 0x9d2c5680u;
        y ^= (y << 15) & 0xefc60000u;
        return y ^ (y >> 18);
    }
    double rand() { return double(u32()) * (1.0 / 4294967295.0); }
    double randExc() { return double(u32()) * (1.0 / 4294967296.0); }
    double randNorm(double mean, double sd) {
        double x, y, r;
        do {
            x = 2.0 * rand() - 1.0;
            y = 2.0 * rand() - 1.0;
            r = x * x + y * y;
        } while (r >= 1.0 || r == 0.0);
        double f = std::sqrt(-2.0 * std::log(r) / r);
        return mean + x * f * sd;
    }
};

inline double put_in_bc(double r, double L) { return r - L * std::floor(r * (1.0 / L) + 0.5); }

}  // namespace

extern "C" int pimc_gen_worldlines(uint32_t seed, double side, int M, int N, double lambda, double tau, int kind,
                                   double* beads) {
    if (!beads || M < 1 || N < 1 || !(side > 0) || (kind != 0 && kind != 1)) return PIMC_E_ARG;
    Mt rng(seed);
    double* b = beads;   // [M][N][3]
    for (int p = 0; p < N; ++p)
        for (int i = 0; i < 3; ++i) b[3 * (size_t)p + i] = side * (-0.5 + rng.randExc());
    if (kind == 0) {
        const double sd = std::sqrt(2.0 * lambda * tau);
        for (int m = 1; m < M; ++m)
            for (int p = 0; p < N; ++p) {
                const double* q = b + 3 * ((size_t)(m - 1) * N + p);
                double* r = b + 3 * ((size_t)m * N + p);
                for (int i = 0; i < 3; ++i) r[i] = rng.randNorm(q[i], sd);
                for (int i = 0; i < 3; ++i) r[i] = put_in_bc(r[i], side);
            }
    } else {
        const double s2lt = std::sqrt(2.0) * std::sqrt(lambda * tau);
        const int stage = M;
        for (int p = 0; p < N; ++p) {
            const double end[3] = {b[3 * (size_t)p], b[3 * (size_t)p + 1], b[3 * (size_t)p + 2]};
            for (int k = 0; k < M - 1; ++k) {
                const double* nb = b + 3 * ((size_t)k * N + p);
                const double f1 = 1.0 * (stage - k - 1);
                const double f2 = 1.0 / (1.0 * (stage - k));
                const double sq = s2lt * std::sqrt(f1 * f2);
                double r[3];
                for (int i = 0; i < 3; ++i) r[i] = put_in_bc(end[i] - nb[i], side);
                for (int i = 0; i < 3; ++i) r[i] *= f2;
                for (int i = 0; i < 3; ++i) r[i] += nb[i];
                for (int i = 0; i < 3; ++i) r[i] = rng.randNorm(r[i], sq);
                double* o = b + 3 * ((size_t)(k + 1) * N + p);
                for (int i = 0; i < 3; ++i) o[i] = put_in_bc(r[i], side);
            }
        }
    }
    return PIMC_OK;
}
