// pimc_gpu.hpp -- header-only C++ adapter over include/pimc_gpu.h with the
// reference's method names and argument meaning, so pimc's own callers can hold
// a GPU action or potential where they hold a LocalAction or AzizPotential.
//
//   pimc_gpu::Action      ~ LocalAction (include/action.h:30-145, src/action.cpp:456-947)
//   pimc_gpu::Potential   ~ AzizPotential as PotentialBase (include/potential.h:40-117)
//   pimc_gpu::Estimators  ~ Static / IntermediateScatteringFunction estimators'
//                           accumulate() (src/estimator.cpp:3705-3737, :3923-3961)
//
// Errors: every failing C-ABI call throws pimc_gpu::Error with pimc_last_error()'s
// text (the reference aborts through GPU_ASSERT, include/common_gpu.h:34; a thrown
// exception lets a caller decide).
#pragma once
#include <array>
#include <stdexcept>
#include <string>
#include <vector>

#include "pimc_gpu.h"

namespace pimc_gpu {

using dVec = std::array<double, 3>;
using beadLocator = std::array<int, 2>;

struct Error : std::runtime_error {
    explicit Error(const std::string& s) : std::runtime_error(s) {}
};

inline void check(int rc, const pimc_ctx* c, const char* what) {
    if (rc != PIMC_OK) throw Error(std::string(what) + ": " + pimc_last_error(c));
}

class Action {
   public:
    // The constants LocalAction reads from constants() and the box from Prism.
    Action(double side, int numTimeSlices, int numAlloc, double tau, double lambda, double rc,
           std::array<double, 2> VFactor = {1.0, 1.0}, int year = 1979, int v_mode = PIMC_V_MIXED)
        : VFactor_(VFactor), M_(numTimeSlices), N_(numAlloc), tau_(tau), lambda_(lambda) {
        pimc_desc d{side, numTimeSlices, numAlloc, tau, lambda, rc, year, v_mode};
        check(pimc_ctx_create(&d, &ctx_), nullptr, "pimc_ctx_create");
    }
    ~Action() { pimc_ctx_destroy(ctx_); }
    Action(const Action&) = delete;
    Action& operator=(const Action&) = delete;

    // Path::Path / loadState + LookupTable::updateGrid: beads [M][N][3].
    void uploadPath(const std::vector<double>& beads, const std::vector<uint8_t>* beadOn = nullptr,
                    const std::vector<int32_t>* numBeadsAtSlice = nullptr) {
        check(pimc_upload_path(ctx_, beads.data(), beadOn ? beadOn->data() : nullptr,
                               numBeadsAtSlice ? numBeadsAtSlice->data() : nullptr),
              ctx_, "upload_path");
    }
    // Path::updateBead (src/path.cpp:197-200)
    void updateBead(const beadLocator& b, const dVec& pos) {
        const int32_t sp[2] = {b[0], b[1]};
        check(pimc_update_beads(ctx_, sp, pos.data(), 1), ctx_, "update_beads");
    }
    // ActionBase::setShift (include/action.h:92)
    void setShift(int shift) { shift_ = shift; }
    int getShift() const { return shift_; }

    // LocalAction::potentialAction() (src/action.cpp:456-472)
    double potentialAction() {
        double a = 0.0;
        check(pimc_total_action(ctx_, shift_, VFactor_.data(), &a, nullptr), ctx_, "total_action");
        return a;
    }
    // LocalAction::potentialAction(const beadLocator&) (src/action.cpp:487-514)
    double potentialAction(const beadLocator& b) { return potentialAction(std::vector<beadLocator>{b}); }
    // LocalAction::potentialAction(const std::vector<beadLocator>&) (src/action.cpp:572-620)
    double potentialAction(const std::vector<beadLocator>& beads) {
        std::vector<int32_t> sp;
        for (const auto& b : beads) {
            sp.push_back(b[0]);
            sp.push_back(b[1]);
        }
        double a = 0.0;
        check(pimc_action_list(ctx_, sp.data(), (int)beads.size(), shift_, VFactor_.data(), &a, nullptr), ctx_,
              "action_list");
        return a;
    }
    // LocalAction::potentialAction(start, end) over next links (src/action.cpp:522-570);
    // nextLinks [M][N][2] or empty for closed periodic worldlines.
    double potentialAction(const beadLocator& start, const beadLocator& end,
                           const std::vector<int32_t>& nextLinks = {}) {
        const int32_t s[2] = {start[0], start[1]}, e[2] = {end[0], end[1]};
        double a = 0.0;
        check(pimc_action_range(ctx_, s, e, nextLinks.empty() ? nullptr : nextLinks.data(), shift_,
                                VFactor_.data(), &a),
              ctx_, "action_range");
        return a;
    }
    // A move's oldAction/newAction pair in one pass (src/move.cpp:1216-1248, :1370-1394):
    // returns {oldAction, newAction, newAction - oldAction}; commit applies Path::updateBead.
    std::array<double, 3> deltaAction(const std::vector<beadLocator>& beads, const std::vector<dVec>& newPos,
                                      bool commit) {
        std::vector<int32_t> sp;
        std::vector<double> np;
        for (size_t k = 0; k < beads.size(); ++k) {
            sp.push_back(beads[k][0]);
            sp.push_back(beads[k][1]);
            np.insert(np.end(), newPos[k].begin(), newPos[k].end());
        }
        std::array<double, 3> out{};
        check(pimc_delta_action(ctx_, sp.data(), nullptr, np.data(), (int)beads.size(), shift_, VFactor_.data(),
                                commit ? 1 : 0, out.data()),
              ctx_, "delta_action");
        return out;
    }
    // LocalAction's factors (src/setup.cpp:1232-1256): VFactor, gradVFactor of the
    // GSF / Li-Broughton correction, estimator period.
    void setActionFactors(std::array<double, 2> VFactor, std::array<double, 2> gradVFactor, int period) {
        check(pimc_set_action_factors(ctx_, VFactor.data(), gradVFactor.data(), period), ctx_, "set_action_factors");
        VFactor_ = VFactor;
        gradVFactor_ = gradVFactor;
    }
    // LocalAction::potential(int) == V(int) = {Vext, Vint} (src/action.cpp:902-947)
    std::array<double, 2> potential(int slice) {
        refreshSlices();
        return {vs_[2 * slice], vs_[2 * slice + 1]};
    }
    // LocalAction::derivPotentialActionTau(int) (src/action.cpp:751-767):
    // VFactor sum(V(slice)) + 3 gradVFactor tau^2 lambda gradVSquared(slice)
    double derivPotentialActionTau(int slice) {
        refreshSlices();
        const int eo = slice % 2;
        double dU = VFactor_[eo] * (vs_[2 * slice] + vs_[2 * slice + 1]);
        if (gradVFactor_[eo] > kEps) dU += 3.0 * gradVFactor_[eo] * tau_ * tau_ * lambda_ * g2_[slice];
        return dU;
    }
    // LocalAction::derivPotentialActionLambda(int) (src/action.cpp:796-802)
    double derivPotentialActionLambda(int slice) {
        refreshSlices();
        const int eo = slice % 2;
        return gradVFactor_[eo] > kEps ? gradVFactor_[eo] * tau_ * tau_ * tau_ * g2_[slice] : 0.0;
    }
    // V(slice) (and gradVSquared(slice) with a finite correction) for every slice
    // in one pass, kept until the context's state changes: an estimator's
    // per-slice calls (src/estimator.cpp:985-987) cost one launch per sample.
    void refreshSlices() {
        uint64_t v = 0;
        check(pimc_path_version(ctx_, &v), ctx_, "path_version");
        if (v == stamp_ && !vs_.empty()) return;
        vs_.resize(2 * (size_t)M_);
        check(pimc_potential_slices(ctx_, vs_.data(), nullptr), ctx_, "potential_slices");
        g2_.assign(M_, 0.0);
        if (gradVFactor_[0] > kEps || gradVFactor_[1] > kEps)
            check(pimc_gradv_squared(ctx_, g2_.data()), ctx_, "gradv_squared");
        stamp_ = v;
    }
    // EnergyEstimator::accumulate columns {K, V, V_ext, V_int, E, E_mu, K/N, V/N, E/N}
    std::array<double, 9> energy(const std::vector<int32_t>& nextLinks = {}, double mu = 0.0) {
        std::array<double, 9> o{};
        check(pimc_energy(ctx_, nextLinks.empty() ? nullptr : nextLinks.data(), mu, o.data()), ctx_, "energy");
        return o;
    }
    // StaticStructureFactorEstimator::accumulate contribution sf/N
    std::vector<double> staticStructureFactor(const std::vector<dVec>& q) {
        std::vector<double> flat, out(q.size());
        for (const auto& v : q) flat.insert(flat.end(), v.begin(), v.end());
        check(pimc_ssf(ctx_, flat.data(), (int)q.size(), out.data()), ctx_, "ssf");
        return out;
    }
    // IntermediateScatteringFunctionEstimator::accumulate contribution isf/N, [nq][M]
    std::vector<double> intermediateScatteringFunction(const std::vector<dVec>& q) {
        std::vector<double> flat, out(q.size() * (size_t)M_);
        for (const auto& v : q) flat.insert(flat.end(), v.begin(), v.end());
        check(pimc_isf(ctx_, flat.data(), (int)q.size(), out.data()), ctx_, "isf");
        return out;
    }
    pimc_ctx* raw() { return ctx_; }

   private:
    static constexpr double kEps = 1.0e-7;   // a gradVFactor counts above EPS (src/action.cpp:760)
    pimc_ctx* ctx_ = nullptr;
    std::array<double, 2> VFactor_;
    std::array<double, 2> gradVFactor_{0.0, 0.0};
    int shift_ = 1;
    int M_, N_;
    double tau_, lambda_;
    std::vector<double> vs_, g2_;
    uint64_t stamp_ = ~uint64_t(0);
};

// AzizPotential as a batched PotentialBase (include/potential.h:86, :985-1018).
class Potential {
   public:
    Potential(int year, double maxSep, double rc, int v_mode = PIMC_V_TABLE) {
        if (pimc_aziz_create(year, maxSep, rc, v_mode, &az_) != PIMC_OK) throw Error("pimc_aziz_create");
        tailV = pimc_aziz_tailV(az_);
    }
    ~Potential() { pimc_aziz_destroy(az_); }
    Potential(const Potential&) = delete;
    Potential& operator=(const Potential&) = delete;
    // PotentialBase::V(const dVec*, double*, int)
    void V(const dVec* sep, double* out, int n) {
        if (n < 0) throw std::runtime_error("negative count");  // src/potential.cpp:128-129
        if (pimc_aziz_V(az_, sep[0].data(), out, n) != PIMC_OK) throw Error("pimc_aziz_V");
    }
    double V(const dVec& sep) {
        double o = 0.0;
        V(&sep, &o, 1);
        return o;
    }
    dVec gradV(const dVec& sep) {
        dVec g{};
        if (pimc_aziz_gradV(az_, sep.data(), g.data(), 1) != PIMC_OK) throw Error("pimc_aziz_gradV");
        return g;
    }
    double grad2V(const dVec& sep) {
        double o = 0.0;
        if (pimc_aziz_grad2V(az_, sep.data(), &o, 1) != PIMC_OK) throw Error("pimc_aziz_grad2V");
        return o;
    }
    double tailV = 0.0;

   private:
    pimc_aziz* az_ = nullptr;
};

}  // namespace pimc_gpu
