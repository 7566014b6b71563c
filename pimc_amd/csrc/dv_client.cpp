// dv_client.cpp -- a C++ caller of the C-ABI, as the reference's moves would be:
// p50/p99 wall latency of pimc_delta_action (Bisection segments of 8/16/32
// slices: old + new Vnn of the s-1 moved beads) and of pimc_total_action on
// seeded worldlines, printed as one JSON object.
//
//   dv_client N M tau seed requests [gsf]
// "gsf" selects the GSF action (VFactor {2/3, 4/3}, gradVFactor {0, 2/9}): the
// bead actions then carry the three-body gradient correction.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <string>
#include <vector>

#include "../../include/pimc_gpu.h"

using Clk = std::chrono::steady_clock;

static double pct(std::vector<double> v, double q) {
    std::sort(v.begin(), v.end());
    return v[(size_t)(q * (v.size() - 1))];
}

int main(int argc, char** argv) {
    const int N = argc > 1 ? std::atoi(argv[1]) : 256;
    const int M = argc > 2 ? std::atoi(argv[2]) : 167;
    const double tau = argc > 3 ? std::atof(argv[3]) : 0.004;
    const unsigned seed = argc > 4 ? (unsigned)std::atol(argv[4]) : 139853u;
    const int R = argc > 5 ? std::atoi(argv[5]) : 2000;
    const bool gsf = argc > 6 && std::string(argv[6]) == "gsf";
    const double lambda = 6.0596, density = 0.02198;
    const double side = std::pow(N / density, 1.0 / 3.0);
    std::vector<double> beads(3 * (size_t)M * N);
    if (pimc_gen_worldlines(seed, side, M, N, lambda, tau, 0, beads.data()) != PIMC_OK) return 2;
    pimc_desc d{side, M, N, tau, lambda, side, 1979, PIMC_V_MIXED};
    pimc_ctx* c = nullptr;
    if (pimc_ctx_create(&d, &c) != PIMC_OK) {
        std::fprintf(stderr, "pimc_ctx_create: %s\n", pimc_last_error(nullptr));
        return 1;
    }
    if (pimc_upload_path(c, beads.data(), nullptr, nullptr) != PIMC_OK) {
        std::fprintf(stderr, "upload: %s\n", pimc_last_error(c));
        return 1;
    }
    if (gsf) {
        const double vf_gsf[2] = {2.0 / 3.0, 4.0 / 3.0}, gvf[2] = {0.0, 2.0 / 9.0};
        if (pimc_set_action_factors(c, vf_gsf, gvf, 2) != PIMC_OK) return 1;
    }
    const double* vf = nullptr;   // the context's factors
    double a = 0.0;
    std::vector<double> t_full;
    for (int k = 0; k < R + 50; ++k) {
        const auto t0 = Clk::now();
        if (pimc_total_action(c, 1, vf, &a, nullptr) != PIMC_OK) return 1;
        if (k >= 50) t_full.push_back(std::chrono::duration<double, std::micro>(Clk::now() - t0).count());
    }
    std::mt19937_64 rng(seed);
    std::normal_distribution<double> G(0.0, std::sqrt(lambda * tau));
    std::printf("{\"N\": %d, \"M\": %d, \"action\": \"%s\", \"requests\": %d, \"total_action_p50_us\": %.3f, "
                "\"total_action_p99_us\": %.3f",
                N, M, gsf ? "gsf" : "primitive", R, pct(t_full, 0.5), pct(t_full, 0.99));
    for (int seg : {8, 16, 32}) {
        // requests drawn before timing
        std::vector<std::vector<int32_t>> sps(R + 50);
        std::vector<std::vector<double>> nps(R + 50);
        for (int k = 0; k < R + 50; ++k) {
            const int p = (int)(rng() % (unsigned)N), m0 = (int)(rng() % (unsigned)M);
            for (int s = 1; s < seg; ++s) {
                const int m = (m0 + s) % M;
                sps[k].push_back(m);
                sps[k].push_back(p);
                for (int x = 0; x < 3; ++x) {
                    double r = beads[3 * ((size_t)m * N + p) + x] + G(rng);
                    r -= side * std::floor(r / side + 0.5);
                    nps[k].push_back(r);
                }
            }
        }
        std::vector<double> t;
        for (int k = 0; k < R + 50; ++k) {
            double out3[3];
            const auto t0 = Clk::now();
            if (pimc_delta_action(c, sps[k].data(), nullptr, nps[k].data(), seg - 1, 1, vf, 0, out3) != PIMC_OK)
                return 1;
            if (k >= 50) t.push_back(std::chrono::duration<double, std::micro>(Clk::now() - t0).count());
        }
        double mean = 0.0;
        for (double x : t) mean += x;
        mean /= (double)t.size();
        std::printf(", \"dv%d_p50_us\": %.3f, \"dv%d_p99_us\": %.3f, \"dv%d_mean_us\": %.3f", seg, pct(t, 0.5), seg,
                    pct(t, 0.99), seg, mean);
    }
    std::printf("}\n");
    pimc_ctx_destroy(c);
    return 0;
}
