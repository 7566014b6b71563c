// pimc_mtrand.hpp -- host-side MTRand (include/MersenneTwister.h): MT19937 with
// init_genrand seeding (:173-188), tempered 32-bit draws (:290-304), rand() =
// u/4294967295 on [0,1] (:326), randExc() = u/4294967296 on [0,1) (:331),
// randInt(n) by masked rejection on [0,n] (:306-323) and randNorm by the
// Marsaglia polar method keeping the first variate (:349-363).  Every draw the
// moves make goes through this stream, so a chain's decisions follow the
// reference's for the same seed.
#pragma once
#include <cmath>
#include <cstdint>

#pragma STDC FP_CONTRACT OFF

namespace pimc {

struct MtRand {
    static constexpr int N = 624, M = 397;
    uint32_t s[N];
    int left = 0, next = 0;
    explicit MtRand(uint32_t seed) { seed_with(seed); }
    void seed_with(uint32_t seed) {
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
    uint32_t randInt(uint32_t n) {
        uint32_t used = n;
        used |= used >> 1;
        used |= used >> 2;
        used |= used >> 4;
        used |= used >> 8;
        used |= used >> 16;
        uint32_t i;
        do {
            i = u32() & used;
        } while (i > n);
        return i;
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

// Container::putInBC for one periodic axis (include/container.h:50-59).
inline double put_in_bc(double r, double L) { return r - L * std::floor(r * (1.0 / L) + 0.5); }

}  // namespace pimc
