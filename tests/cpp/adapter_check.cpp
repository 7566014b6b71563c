// Drives include/pimc_gpu.hpp the way pimc's callers would (LocalAction /
// PotentialBase names) and prints the results as JSON for tests/test_gpu_adapter.py.
#include <cmath>
#include <cstdio>
#include <vector>

#include "pimc_gpu.hpp"

int main(int argc, char** argv) {
    const int N = argc > 1 ? std::atoi(argv[1]) : 32;
    const int M = argc > 2 ? std::atoi(argv[2]) : 16;
    const double tau = 0.004, lambda = 6.0596;
    const double L = std::pow(1.0 * N / 0.02198, 1.0 / 3.0);
    std::vector<double> beads(3 * (size_t)M * N);
    if (pimc_gen_worldlines(139853, L, M, N, lambda, tau, 1, beads.data()) != PIMC_OK) return 2;
    try {
        pimc_gpu::Action act(L, M, N, tau, lambda, L, {1.0, 1.0}, 1979, PIMC_V_TABLE);
        act.uploadPath(beads);
        const double total = act.potentialAction();
        act.refreshSlices();
        const auto v3 = act.potential(3);
        const double d3 = act.derivPotentialActionTau(3);
        std::vector<pimc_gpu::beadLocator> seg = {{2, 5}, {3, 5}, {4, 5}, {5, 5}};
        const double list = act.potentialAction(seg);
        const double range = act.potentialAction(pimc_gpu::beadLocator{2, 5}, pimc_gpu::beadLocator{5, 5});
        std::vector<pimc_gpu::dVec> np;
        for (auto& b : seg) {
            const double* p = &beads[3 * ((size_t)b[0] * N + b[1])];
            np.push_back({p[0] + 0.3, p[1] - 0.2, p[2] + 0.1});
        }
        const auto dv = act.deltaAction(seg, np, false);
        act.setShift(2);
        const double shifted = act.potentialAction();
        act.setShift(1);
        const auto e = act.energy({}, 0.5);
        // the cache follows the path: V(3) after bead (3, 5) moves
        act.updateBead(pimc_gpu::beadLocator{3, 5}, np[1]);
        const double v3_after = act.potential(3)[1];
        act.updateBead(pimc_gpu::beadLocator{3, 5}, pimc_gpu::dVec{beads[3 * ((size_t)3 * N + 5)],
                                                                  beads[3 * ((size_t)3 * N + 5) + 1],
                                                                  beads[3 * ((size_t)3 * N + 5) + 2]});
        // GSF factors: the thermodynamic V of EnergyEstimator::accumulate from the
        // adapter's per-slice tau and lambda derivatives (src/estimator.cpp:985-1003)
        act.setActionFactors({2.0 / 3.0, 4.0 / 3.0}, {0.0, 2.0 / 9.0}, 2);
        double t1 = 0.0, t2 = 0.0;
        for (int m = 0; m < M; ++m) {
            t1 += act.derivPotentialActionLambda(m);
            t2 += act.derivPotentialActionTau(m);
        }
        t1 *= lambda / (tau * M);
        t2 /= 1.0 * M;
        const auto eg = act.energy();
        pimc_gpu::Potential pot(1979, std::sqrt(3.0) * L / 2.0, L);
        const double v = pot.V(pimc_gpu::dVec{2.9673, 0.0, 0.0});
        std::printf("{\"N\": %d, \"M\": %d, \"L\": %.17g, \"total\": %.17g, \"v3\": [%.17g, %.17g], \"d3\": %.17g, "
                    "\"list\": %.17g, \"range\": %.17g, \"dv\": [%.17g, %.17g, %.17g], \"shifted\": %.17g, "
                    "\"E\": [%.17g, %.17g, %.17g, %.17g, %.17g, %.17g, %.17g, %.17g, %.17g], \"V_rm\": %.17g, "
                    "\"tailV\": %.17g, \"v3_after\": %.17g, \"gsf_t1\": %.17g, \"gsf_t2\": %.17g, "
                    "\"gsf_E\": [%.17g, %.17g, %.17g, %.17g, %.17g]}\n",
                    N, M, L, total, v3[0], v3[1], d3, list, range, dv[0], dv[1], dv[2], shifted, e[0], e[1], e[2],
                    e[3], e[4], e[5], e[6], e[7], e[8], v, pot.tailV, v3_after, t1, t2, eg[0], eg[1], eg[2], eg[3],
                    eg[4]);
    } catch (const std::exception& ex) {
        std::fprintf(stderr, "adapter error: %s\n", ex.what());
        return 1;
    }
    return 0;
}
