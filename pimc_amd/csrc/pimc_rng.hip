// pimc_rng.hip -- the seeded worldline generator used to build the measured
// workloads, on the MTRand stream of pimc_mtrand.hpp.
//
// Slice 0 is Prism::randPosition (src/container.cpp:155-160);
// kind 0 adds randNorm(prev, sqrt(2 lambda tau)) per axis and putInBC, kind 1
// closes a Gaussian bridge from slice 0 back to slice 0 with the staging
// construction (src/move.cpp:270-291).
#include <cmath>
#include <cstdint>
#include <vector>

#include "../../include/pimc_gpu.h"
#include "pimc_mtrand.hpp"

#pragma STDC FP_CONTRACT OFF

using pimc::put_in_bc;
using Mt = pimc::MtRand;

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
