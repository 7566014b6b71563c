// pimc_chain.hip -- one Markov chain's diagonal moves over a resident path:
// StagingMove, BisectionMove and CenterOfMassMove::attemptMove (src/move.cpp:
// 1124-1259, 1315-1408, 1484-1661) for closed worldlines (next(m, i) =
// (m+1 mod M, i), every bead on), the PIMC case of those moves.
//
// Proposals are drawn on the host from the chain's MTRand stream in the
// reference's order, against a host copy of the positions that receives every
// Path::updateBead the reference makes; each oldAction/newAction pair is one
// device pass (pimc_delta_action) over the moved beads, and the Metropolis
// test rand() < exp(-(newAction - oldAction)) uses the same stream.  Given the
// same seed and worldlines, a run therefore proposes the same positions and
// takes the same decisions as the reference (up to the ~1e-15 relative rounding
// of the device pair sums against exp's argument).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "pimc_host.hpp"
#include "pimc_mtrand.hpp"

using pimc::MtRand;
using pimc::put_in_bc;

namespace {
constexpr double kEps = 1.0e-7;   // constants()->EPS
}

struct pimc_chain {
    pimc_ctx* ctx = nullptr;
    pimc_chain_desc d{};
    MtRand rng{0u};
    int M = 0, N = 0;
    double side = 0, tau = 0, lambda = 0;
    double sqrtLambdaTau = 0, sqrt2LambdaTau = 0;
    std::vector<double> pos;              // [M][N][3], equal to the device copy
    std::vector<int> wind;                // winding sectors, 3 per sector
    std::vector<double> cumrho0;
    std::string err;
    // scratch
    std::vector<int32_t> sp;
    std::vector<double> np, op;
    // device-resident runs: MTRand state, winding table, per-move outputs
    uint32_t* d_mt = nullptr;
    int* d_wind = nullptr;
    int* d_kinds = nullptr;
    int* d_acc = nullptr;
    double* d_info = nullptr;
    int* d_near = nullptr;
    int cap_moves = 0;
    ~pimc_chain() {
        for (void* p : {(void*)d_mt, (void*)d_wind, (void*)d_kinds, (void*)d_acc, (void*)d_info, (void*)d_near})
            if (p) (void)hipFree(p);
    }
    int check_device(int n_moves, const int* kinds);
    int device_args(int n_moves, const int* kinds, pimc::ChainRunArgs* A, pimc::AzizDev* az);
    int device_results(int n_moves, int* accepted, double* info, int* near);

    double* P(int m, int i) { return pos.data() + 3 * ((size_t)m * N + i); }
    int fail(int code, const std::string& m) {
        err = m;
        return code;
    }
    // Path::updateBead on the host copy (the device copy is updated by the caller).
    void set(int m, int i, const double r[3]) { std::memcpy(P(m, i), r, sizeof(double) * 3); }

    // action.cpp rho0(dVec, int): free-particle propagator over Mlen slices
    double rho0(const double vel[3], int Mlen) const {
        const double norm = std::pow(4.0 * M_PI * lambda * Mlen * tau, -0.5 * 3);
        const double v2 = vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2];
        return (norm * std::exp(-v2 / (4.0 * lambda * Mlen * tau)));
    }
    // MoveBase::sampleWindingSector (move.cpp:359-396)
    void sample_winding(int m_s, int m_e, int i, int stageLength, int w[3]) {
        const double* pe = P(m_e, i);
        const double* ps = P(m_s, i);
        const double vel[3] = {pe[0] - ps[0], pe[1] - ps[1], pe[2] - ps[2]};
        const int nw = (int)cumrho0.size();
        cumrho0[0] = rho0(vel, stageLength);
        double total = cumrho0[0];
        for (int n = 1; n < nw; ++n) {
            double vw[3];
            for (int k = 0; k < 3; ++k) vw[k] = vel[k] + wind[3 * n + k] * side;
            const double c = rho0(vw, stageLength);
            total += c;
            cumrho0[n] = cumrho0[n - 1] + c;
        }
        for (auto& c : cumrho0) c /= total;
        const double u = rng.rand();
        const int idx = (int)(std::lower_bound(cumrho0.begin(), cumrho0.end(), u) - cumrho0.begin());
        for (int k = 0; k < 3; ++k) w[k] = wind[3 * idx + k];
    }
    // MoveBase::newStagingPosition, winding form (move.cpp:306-346)
    void staging_position(int m_prev, int m_end, int i, int stageLength, int k, int w[3], double r[3]) {
        const double f1 = 1.0 * (stageLength - k - 1);
        const double f2 = 1.0 / (1.0 * (stageLength - k));
        const double sq = sqrt2LambdaTau * std::sqrt(f1 * f2);
        const double* npv = P(m_prev, i);
        const double* pe = P(m_end, i);
        for (int a = 0; a < 3; ++a) r[a] = (pe[a] + side * w[a]) - npv[a];
        for (int a = 0; a < 3; ++a) r[a] *= f2;
        for (int a = 0; a < 3; ++a) r[a] += npv[a];
        for (int a = 0; a < 3; ++a) r[a] = rng.randNorm(r[a], sq);
        for (int a = 0; a < 3; ++a) {
            while (r[a] < -0.5 * side) {
                w[a]++;
                r[a] += side;
            }
            while (r[a] >= 0.5 * side) {
                w[a]--;
                r[a] -= side;
            }
        }
        for (int a = 0; a < 3; ++a) r[a] = put_in_bc(r[a], side);
    }
    // MoveBase::newBisectionPosition (move.cpp:467-488)
    void bisection_position(int m, int i, int lshift, double r[3]) {
        const double delta = sqrtLambdaTau * std::sqrt(1.0 * lshift);
        const int mp = (m - lshift % M + M) % M, mn = (m + lshift) % M;
        const double* a = P(mn, i);
        const double* b = P(mp, i);
        for (int k = 0; k < 3; ++k) r[k] = put_in_bc(a[k] - b[k], side);   // Path::getSeparation
        for (int k = 0; k < 3; ++k) r[k] *= 0.5;
        for (int k = 0; k < 3; ++k) r[k] += b[k];
        for (int k = 0; k < 3; ++k) r[k] = rng.randNorm(r[k], delta);
        for (int k = 0; k < 3; ++k) r[k] = put_in_bc(r[k], side);
    }
    // One device pass: {A_old, A_new} of the listed beads (old: device copy, or
    // `old` when given), new positions `np`; commit applies them on the device.
    int delta(int n, const double* old, int shift, bool commit, double out3[3]) {
        // the context's action factors (VFactor and any gradient correction)
        int r = pimc_delta_action(ctx, sp.data(), old, np.data(), n, shift, nullptr, commit ? 1 : 0, out3);
        if (r) return fail(r, std::string("delta action: ") + pimc_last_error(ctx));
        return PIMC_OK;
    }
    int device_set(const std::vector<int32_t>& beads, const std::vector<double>& p) {
        const int n = (int)beads.size() / 2;
        if (n == 0) return PIMC_OK;
        int r = pimc_update_beads(ctx, beads.data(), p.data(), n);
        if (r) return fail(r, std::string("update beads: ") + pimc_last_error(ctx));
        return PIMC_OK;
    }

    int staging(int* acc, double* info);
    int bisection(int* acc, double* info);
    int com(int* acc, double* info);
};

// StagingMove::attemptMove (move.cpp:1315-1408)
int pimc_chain::staging(int* acc, double* info) {
    *acc = 0;
    int stageLength = d.Mbar;
    const int s0 = (int)rng.randInt(M - 1);
    const int s1 = (int)rng.randInt(N - 1);
    if (d.var_updates) stageLength = 2 + (int)rng.randInt(d.Mbar - 2);
    const int e0 = (s0 + stageLength) % M;
    info[0] = 1.0;
    info[2] = stageLength;
    int w[3];
    sample_winding(s0, e0, s1, stageLength, w);
    // actionRange(startBead, prev(endBead)): the start bead and the stageLength-1 moved beads
    const int nb = stageLength;
    sp.assign(2 * nb, s1);
    np.resize(3 * nb);
    op.resize(3 * nb);
    for (int k = 0; k < nb; ++k) sp[2 * k] = (s0 + k) % M;
    std::memcpy(np.data(), P(s0, s1), sizeof(double) * 3);
    std::memcpy(op.data(), P(s0, s1), sizeof(double) * 3);
    for (int k = 0; k < nb - 1; ++k) {
        const int m = (s0 + k + 1) % M, mp = (s0 + k) % M;
        std::memcpy(&op[3 * (k + 1)], P(m, s1), sizeof(double) * 3);
        double r[3];
        staging_position(mp, e0, s1, stageLength, k, w, r);
        set(m, s1, r);
        std::memcpy(&np[3 * (k + 1)], r, sizeof(double) * 3);
    }
    double o[3];
    int rc = delta(nb, nullptr, 1, true, o);
    if (rc) return rc;
    const double oldAction = o[0], newAction = o[1];
    info[1] = newAction - oldAction;
    if (rng.rand() < std::exp(-(newAction - oldAction))) {
        *acc = 1;
        return PIMC_OK;
    }
    std::vector<int32_t> moved(sp.begin() + 2, sp.end());
    std::vector<double> back(op.begin() + 3, op.end());
    for (int k = 1; k < nb; ++k) set(sp[2 * k], s1, &op[3 * k]);
    return device_set(moved, back);
}

// BisectionMove::attemptMove, local-action branch (move.cpp:1484-1661)
int pimc_chain::bisection(int* acc, double* info) {
    *acc = 0;
    int numLevels = d.levels;
    if (d.var_updates) numLevels = 1 + (int)rng.randInt(d.levels - 1);
    const int numActive = (int)std::floor(std::pow(2.0, 1.0 * numLevels) + kEps) - 1;
    const int s0 = (int)rng.randInt(M - 1);
    const int s1 = (int)rng.randInt(N - 1);
    info[0] = 1.0;
    info[2] = numLevels;
    std::vector<char> include(numActive, 1);
    std::vector<double> orig(3 * numActive), newp(3 * numActive);
    double oldDelta = 0.0;
    for (int level = numLevels; level > 0; --level) {
        const int shift = (int)std::floor(std::pow(2.0, 1.0 * (level - 1)) + kEps);
        std::vector<int> inc, prev;   // offsets n: bead (s0 + n + 1) mod M
        for (int k = 1; k * shift < numActive + 1; ++k) {
            const int n = k * shift - 1;
            if (include[n]) {
                std::memcpy(&orig[3 * n], P((s0 + n + 1) % M, s1), sizeof(double) * 3);
                inc.push_back(n);
            } else if (level == 1) {
                prev.push_back(n);
            }
        }
        // proposals of this level (their neighbours at +-shift are fixed here)
        for (int n : inc) {
            double r[3];
            bisection_position((s0 + n + 1) % M, s1, shift, r);
            std::memcpy(&newp[3 * n], r, sizeof(double) * 3);
            set((s0 + n + 1) % M, s1, r);
            include[n] = 0;
        }
        double oldAction = 0.0, newAction = 0.0, o[3];
        const int ni = (int)inc.size();
        sp.resize(2 * ni);
        np.resize(3 * ni);
        for (int t = 0; t < ni; ++t) {
            sp[2 * t] = (s0 + inc[t] + 1) % M;
            sp[2 * t + 1] = s1;
            std::memcpy(&np[3 * t], &newp[3 * inc[t]], sizeof(double) * 3);
        }
        int rc = delta(ni, nullptr, shift, true, o);
        if (rc) return rc;
        oldAction += o[0];
        newAction += o[1];
        if (level == 1 && !prev.empty()) {
            // newAction over the earlier levels' beads, then oldAction with them restored
            const int np_ = (int)prev.size();
            sp.resize(2 * np_);
            np.resize(3 * np_);
            op.resize(3 * np_);
            for (int t = 0; t < np_; ++t) {
                sp[2 * t] = (s0 + prev[t] + 1) % M;
                sp[2 * t + 1] = s1;
                std::memcpy(&np[3 * t], &newp[3 * prev[t]], sizeof(double) * 3);
                std::memcpy(&op[3 * t], &orig[3 * prev[t]], sizeof(double) * 3);
            }
            rc = delta(np_, op.data(), shift, false, o);
            if (rc) return rc;
            newAction += o[1];
            oldAction += o[0];
        }
        const double dlt = (newAction - oldAction);
        info[1] = dlt;
        if (rng.rand() < std::exp(-dlt + oldDelta)) {
            if (level == 1) {
                *acc = 1;
                return PIMC_OK;
            }
        } else {
            // undoMove (move.cpp:1686-1703): every bead moved so far back to its original
            std::vector<int32_t> moved;
            std::vector<double> back;
            for (int n = 0; n < numActive; ++n)
                if (!include[n]) {
                    const int m = (s0 + n + 1) % M;
                    set(m, s1, &orig[3 * n]);
                    moved.push_back(m);
                    moved.push_back(s1);
                    back.insert(back.end(), &orig[3 * n], &orig[3 * n] + 3);
                }
            return device_set(moved, back);
        }
        oldDelta = dlt;
    }
    return PIMC_OK;
}

// CenterOfMassMove::attemptMove for a closed worldline (move.cpp:1124-1259)
int pimc_chain::com(int* acc, double* info) {
    *acc = 0;
    const int p = (int)rng.randInt(N - 1);   // a bead on slice 0; its worldline is slices 0..M-1
    info[0] = 1.0;
    info[2] = M;
    double sh[3];
    for (int k = 0; k < 3; ++k) sh[k] = d.com_delta * (-0.5 + rng.rand());
    sp.resize(2 * M);
    np.resize(3 * M);
    for (int m = 0; m < M; ++m) {
        sp[2 * m] = m;
        sp[2 * m + 1] = p;
        double* q = P(m, p);
        double r[3];
        for (int k = 0; k < 3; ++k) r[k] = put_in_bc(q[k] + sh[k], side);
        set(m, p, r);
        std::memcpy(&np[3 * m], r, sizeof(double) * 3);
    }
    double o[3];
    int rc = delta(M, nullptr, 1, true, o);
    if (rc) return rc;
    const double oldAction = o[0], newAction = o[1];
    info[1] = newAction - oldAction;
    if (rng.rand() < std::exp(-(newAction - oldAction))) {
        *acc = 1;
        return PIMC_OK;
    }
    // the reference shifts back (it does not restore saved positions)
    std::vector<double> back(3 * M);
    for (int m = 0; m < M; ++m) {
        double* q = P(m, p);
        double r[3];
        for (int k = 0; k < 3; ++k) r[k] = put_in_bc(q[k] - sh[k], side);
        set(m, p, r);
        std::memcpy(&back[3 * m], r, sizeof(double) * 3);
    }
    std::vector<int32_t> beads(sp.begin(), sp.end());
    return device_set(beads, back);
}

// Device-resident runs (pimc_dchain.hip): every chain's moves of `kinds` in
// one launch, one workgroup per chain, on the chains' own stream states: each
// host MTRand state goes up before and comes back after, so host moves and
// device runs continue one stream per chain.
int pimc_chain::check_device(int n_moves, const int* kinds) {
    if (n_moves < 0 || (n_moves > 0 && !kinds)) return fail(PIMC_E_ARG, "bad move list");
    for (int k = 0; k < n_moves; ++k)
        if (kinds[k] < 0 || kinds[k] > 2) return fail(PIMC_E_ARG, "move kind must be 0, 1 or 2");
    if (M > pimc::kChainMaxBeads) return fail(PIMC_E_UNSUPPORTED, "device runs hold at most 512 slices");
    if ((int)cumrho0.size() > pimc::kChainMaxWind)
        return fail(PIMC_E_UNSUPPORTED, "device runs sample at most 343 windings");
    if (ctx->correction() && N > pimc::kChainMaxList)
        return fail(PIMC_E_UNSUPPORTED, "device runs with the gradient correction hold at most 8192 beads per slice");
    return PIMC_OK;
}

int pimc_chain::device_args(int n_moves, const int* kinds, pimc::ChainRunArgs* A, pimc::AzizDev* az) {
    auto chk = [&](hipError_t e, const char* what) {
        return e == hipSuccess ? PIMC_OK : fail(PIMC_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
    };
    int r;
    if (!d_mt) {
        if ((r = chk(hipMalloc(&d_mt, sizeof(uint32_t) * (MtRand::N + 2)), "hipMalloc"))) return r;
        if ((r = chk(hipMalloc(&d_wind, sizeof(int) * wind.size()), "hipMalloc"))) return r;
        if ((r = chk(hipMalloc(&d_near, sizeof(int)), "hipMalloc"))) return r;
        if ((r = chk(hipMemcpy(d_wind, wind.data(), sizeof(int) * wind.size(), hipMemcpyHostToDevice), "upload")))
            return r;
    }
    if (cap_moves < n_moves) {
        for (void* p : {(void*)d_kinds, (void*)d_acc, (void*)d_info})
            if (p) (void)hipFree(p);
        d_kinds = d_acc = nullptr;
        d_info = nullptr;
        if ((r = chk(hipMalloc(&d_kinds, sizeof(int) * n_moves), "hipMalloc"))) return r;
        if ((r = chk(hipMalloc(&d_acc, sizeof(int) * n_moves), "hipMalloc"))) return r;
        if ((r = chk(hipMalloc(&d_info, sizeof(double) * 3 * n_moves), "hipMalloc"))) return r;
        cap_moves = n_moves;
    }
    uint32_t mt[MtRand::N + 2];
    std::memcpy(mt, rng.s, sizeof(uint32_t) * MtRand::N);
    mt[MtRand::N] = (uint32_t)rng.left;
    mt[MtRand::N + 1] = (uint32_t)rng.next;
    if ((r = chk(hipMemcpy(d_mt, mt, sizeof(mt), hipMemcpyHostToDevice), "upload"))) return r;
    if ((r = chk(hipMemcpy(d_kinds, kinds, sizeof(int) * n_moves, hipMemcpyHostToDevice), "upload"))) return r;
    *az = pimc::context_aziz(ctx, &r);
    if (r) return fail(r, pimc_last_error(ctx));
    *A = pimc::ChainRunArgs{};
    A->pos = ctx->d_pos;
    A->plane = ctx->plane;
    A->nalloc = N;
    A->num_slices = M;
    A->cell = ctx->d_cell;
    A->label = ctx->d_label;
    A->numlab = ctx->d_numlab;
    A->hash = ctx->d_hash;
    A->box = ctx->box;
    A->side = side;
    A->tau = tau;
    A->lambda = lambda;
    A->vf0 = ctx->vf_default[0];
    A->vf1 = ctx->vf_default[1];
    A->gvf0 = ctx->correction() ? ctx->gvf[0] : 0.0;
    A->gvf1 = ctx->correction() ? ctx->gvf[1] : 0.0;
    A->sqrtLambdaTau = sqrtLambdaTau;
    A->sqrt2LambdaTau = sqrt2LambdaTau;
    A->com_delta = d.com_delta;
    A->Mbar = d.Mbar;
    A->levels = d.levels;
    A->var_updates = d.var_updates;
    A->num_wind = (int)cumrho0.size();
    A->wind = d_wind;
    A->mt = d_mt;
    A->kinds = d_kinds;
    A->n_moves = n_moves;
    A->accepted = d_acc;
    A->info = d_info;
    A->near = d_near;
    return PIMC_OK;
}

int pimc_chain::device_results(int n_moves, int* accepted, double* info, int* near) {
    auto chk = [&](hipError_t e, const char* what) {
        return e == hipSuccess ? PIMC_OK : fail(PIMC_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
    };
    int r;
    uint32_t mt[MtRand::N + 2];
    if ((r = chk(hipMemcpy(mt, d_mt, sizeof(mt), hipMemcpyDeviceToHost), "download"))) return r;
    if (accepted && (r = chk(hipMemcpy(accepted, d_acc, sizeof(int) * n_moves, hipMemcpyDeviceToHost), "download")))
        return r;
    if (info && (r = chk(hipMemcpy(info, d_info, sizeof(double) * 3 * n_moves, hipMemcpyDeviceToHost), "download")))
        return r;
    int nr = 0;
    if ((r = chk(hipMemcpy(&nr, d_near, sizeof(int), hipMemcpyDeviceToHost), "download"))) return r;
    std::memcpy(rng.s, mt, sizeof(uint32_t) * MtRand::N);
    rng.left = (int)mt[MtRand::N];
    rng.next = (int)mt[MtRand::N + 1];
    if (near) *near = nr;
    // the host copy follows the device path
    r = pimc_download_path(ctx, pos.data());
    if (r) return fail(r, pimc_last_error(ctx));
    return PIMC_OK;
}

static int run_chains_on_device(pimc_chain* const* chains, int n, int n_moves, const int* kinds, int* accepted,
                                double* info, int* near) {
    if (!chains || n < 1) return PIMC_E_ARG;
    for (int c = 0; c < n; ++c) {
        if (!chains[c]) return PIMC_E_ARG;
        int r = chains[c]->check_device(n_moves, kinds);
        if (r) return r;
        if (chains[c]->ctx->desc.v_mode != chains[0]->ctx->desc.v_mode)
            return chains[c]->fail(PIMC_E_ARG, "chains of one run share the pair-term mode");
        if (chains[c]->ctx->correction() != chains[0]->ctx->correction())
            return chains[c]->fail(PIMC_E_ARG, "chains of one run all carry the gradient correction or none does");
        for (int e = 0; e < c; ++e)
            if (chains[e]->ctx == chains[c]->ctx) return chains[c]->fail(PIMC_E_ARG, "two chains on one context");
    }
    if (n_moves == 0) {
        if (near)
            for (int c = 0; c < n; ++c) near[c] = 0;
        return PIMC_OK;
    }
    // every context's queued work first: the launch reads and writes their paths
    for (int c = 0; c < n; ++c)
        if (hipStreamSynchronize(chains[c]->ctx->stream) != hipSuccess)
            return chains[c]->fail(PIMC_E_HIP, "stream synchronisation failed");
    std::vector<pimc::ChainRunArgs> args(n);
    std::vector<pimc::AzizDev> azs(n);
    for (int c = 0; c < n; ++c) {
        int r = chains[c]->device_args(n_moves, kinds, &args[c], &azs[c]);
        if (r) return r;
    }
    pimc_chain* lead = chains[0];
    pimc::ChainRunArgs* d_args = nullptr;
    pimc::AzizDev* d_azs = nullptr;
    hipStream_t st = lead->ctx->stream;
    hipError_t e = hipMalloc(&d_args, sizeof(pimc::ChainRunArgs) * n);
    if (e == hipSuccess) e = hipMalloc(&d_azs, sizeof(pimc::AzizDev) * n);
    if (e == hipSuccess) e = hipMemcpy(d_args, args.data(), sizeof(pimc::ChainRunArgs) * n, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_azs, azs.data(), sizeof(pimc::AzizDev) * n, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        lead->ctx->timed_begin(8);
        e = pimc::launch_chain_run(lead->ctx->desc.v_mode, lead->ctx->correction(), st, n, d_args, d_azs);
        lead->ctx->timed_end();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (d_args) (void)hipFree(d_args);
    if (d_azs) (void)hipFree(d_azs);
    if (e != hipSuccess) return lead->fail(PIMC_E_HIP, std::string("chain kernel: ") + hipGetErrorString(e));
    for (int c = 0; c < n; ++c) ++chains[c]->ctx->path_version;   // the launch moved their beads
    for (int c = 0; c < n; ++c) {
        int r = chains[c]->device_results(n_moves, accepted ? accepted + (size_t)c * n_moves : nullptr,
                                          info ? info + (size_t)c * 3 * n_moves : nullptr, near ? near + c : nullptr);
        if (r) return r;
    }
    return PIMC_OK;
}

extern "C" {

int pimc_chain_create(pimc_ctx* ctx, const pimc_chain_desc* d, pimc_chain** out) {
    if (!ctx || !d || !out) return PIMC_E_ARG;
    *out = nullptr;
    if (!ctx->have_path) return PIMC_E_STATE;
    const pimc_desc* ctx_desc = &ctx->desc;
    const int M = ctx_desc->num_slices, N = ctx_desc->num_alloc;
    if (M < 2 || N < 1) return PIMC_E_ARG;
    if (d->Mbar < 2 || d->Mbar >= M || d->levels < 1 || (1 << d->levels) >= M || d->max_wind < 0 ||
        !(d->com_delta >= 0))
        return PIMC_E_ARG;
    // closed worldlines: every bead on, no worm ends
    for (int m = 0; m < M; ++m)
        if (ctx->h_nbs[m] != N) return PIMC_E_STATE;
    for (auto v : ctx->h_on)
        if (!v) return PIMC_E_STATE;
    if (ctx->worm.head[0] >= 0 || ctx->worm.tail[0] >= 0) return PIMC_E_STATE;
    std::unique_ptr<pimc_chain> ch(new pimc_chain());
    ch->ctx = ctx;
    ch->d = *d;
    ch->rng.seed_with(d->seed);
    ch->M = M;
    ch->N = N;
    ch->side = ctx_desc->side;
    ch->tau = ctx_desc->tau;
    ch->lambda = ctx_desc->lambda;
    // MoveBase ctor (move.cpp:58-112)
    ch->sqrtLambdaTau = std::sqrt(ch->lambda * ch->tau);
    ch->sqrt2LambdaTau = std::sqrt(2.0) * ch->sqrtLambdaTau;
    const int nw1 = 2 * d->max_wind + 1;
    const int numWind = (int)std::floor(std::pow(1.0 * nw1, 3.0) + kEps);
    ch->cumrho0.assign(numWind, 0.0);
    ch->wind.assign(3 * numWind, 0);
    for (int n = 0; n < numWind; ++n)
        for (int i = 0; i < 3; ++i) {
            int scale = 1;
            for (int j = i + 1; j < 3; ++j) scale *= nw1;
            int w = (n / scale) % nw1;
            w -= (w > d->max_wind) * nw1;
            ch->wind[3 * n + i] = w;
        }
    // the winding sectors keep this order: the reference's stable_sort
    // comparator compares a sector with itself (move.cpp:120-134)
    ch->pos.resize(3 * (size_t)M * N);
    const int r = pimc_download_path(ctx, ch->pos.data());
    if (r) return r;
    *out = ch.release();
    return PIMC_OK;
}

void pimc_chain_destroy(pimc_chain* ch) { delete ch; }

const char* pimc_chain_last_error(const pimc_chain* ch) { return ch ? ch->err.c_str() : ""; }

int pimc_chain_run_device(pimc_chain* ch, int n_moves, const int* kinds, int* accepted, double* info3,
                          int* near_threshold) {
    if (!ch) return PIMC_E_ARG;
    return run_chains_on_device(&ch, 1, n_moves, kinds, accepted, info3, near_threshold);
}

int pimc_chain_run_device_many(pimc_chain* const* chains, int n_chains, int n_moves, const int* kinds,
                               int* accepted, double* info3, int* near_threshold) {
    return run_chains_on_device(chains, n_chains, n_moves, kinds, accepted, info3, near_threshold);
}

int pimc_chain_move(pimc_chain* ch, int kind, int* accepted, double* info3) {
    if (!ch || !accepted) return PIMC_E_ARG;
    double info[3] = {0.0, 0.0, 0.0};
    int r;
    if (kind == 0)
        r = ch->staging(accepted, info);
    else if (kind == 1)
        r = ch->bisection(accepted, info);
    else if (kind == 2)
        r = ch->com(accepted, info);
    else
        return ch->fail(PIMC_E_ARG, "move kind must be 0 (staging), 1 (bisection) or 2 (center of mass)");
    if (info3) std::memcpy(info3, info, sizeof(info));
    return r;
}

int pimc_chain_positions(const pimc_chain* ch, double* beads) {
    if (!ch || !beads) return PIMC_E_ARG;
    std::memcpy(beads, ch->pos.data(), sizeof(double) * ch->pos.size());
    return PIMC_OK;
}

}  // extern "C"
