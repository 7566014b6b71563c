// pimc_dchain.hip -- device-resident Markov chain: Staging, Bisection and
// CenterOfMass moves (src/move.cpp:1124-1259, 1315-1408, 1484-1661) run inside
// one kernel, many moves per launch, with no host round trip per move.
//
// One workgroup per chain (a launch may run many chains, one workgroup each,
// each with its own arguments).  Lane 0 replays the chain's MTRand stream (state in
// LDS, include/MersenneTwister.h) in the reference's draw order and builds each
// move's proposal; the workgroup then evaluates the move's old and new
// Vnn sums over every partner of the moved beads (the pimc_delta_action
// contract) and lane 0 takes the Metropolis decision; accepted moves are
// written to the resident path and its cell lists, rejected CenterOfMass moves
// are shifted back as the reference shifts them.  The transcendental units
// (log, sqrt, exp, pow) are the device's, so a proposal can differ from the
// host's in the last ulp; a decision whose uniform deviate lies within a
// relative 1e-10 of its acceptance probability is counted (guard band) so a
// caller can tell when such a difference could have flipped it.
#include <hip/hip_runtime.h>

#include "pimc_kernels.hpp"

namespace pimc {

namespace {

constexpr int kMt = 624, kMtM = 397;
constexpr int kPolarBeads = 64;   // beads whose Gaussian deviates are drawn per round

// The MTRand state (include/MersenneTwister.h:76-419) in LDS: the current
// 624-word block and, once refilled, the next one (MTRand::reload of the
// current block, computed ahead of need by the whole workgroup).  Lane 0 draws
// through a cursor that keeps the block index and position in registers.
struct MtLds {
    uint32_t s[2][kMt];
    int cur, next, have_next;
};

__device__ __forceinline__ uint32_t mt_twist(uint32_t m, uint32_t a, uint32_t b) {
    return m ^ (((a & 0x80000000u) | (b & 0x7fffffffu)) >> 1) ^ ((b & 1u) ? 0x9908b0dfu : 0u);
}

// MTRand::reload (include/MersenneTwister.h:271-288) out of place: dst = the
// block that follows src.  Its three ranges depend on earlier ones only, so the
// workgroup computes each range in parallel.
__device__ void mt_refill(MtLds& g) {
    __syncthreads();
    if (g.have_next) return;
    const uint32_t* src = g.s[g.cur];
    uint32_t* dst = g.s[g.cur ^ 1];
    for (int i = threadIdx.x; i < kMt - kMtM; i += (int)blockDim.x) dst[i] = mt_twist(src[i + kMtM], src[i], src[i + 1]);
    __syncthreads();
    for (int i = kMt - kMtM + threadIdx.x; i < 2 * (kMt - kMtM); i += (int)blockDim.x)
        dst[i] = mt_twist(dst[i + kMtM - kMt], src[i], src[i + 1]);
    __syncthreads();
    for (int i = 2 * (kMt - kMtM) + threadIdx.x; i < kMt - 1; i += (int)blockDim.x)
        dst[i] = mt_twist(dst[i + kMtM - kMt], src[i], src[i + 1]);
    __syncthreads();
    if (threadIdx.x == 0) {
        dst[kMt - 1] = mt_twist(dst[kMtM - 1], src[kMt - 1], dst[0]);
        g.have_next = 1;
    }
    __syncthreads();
}

// Lane 0's cursor over the stream.
struct MtCur {
    MtLds* g;
    int cur, next, have_next;
    __device__ explicit MtCur(MtLds& m) : g(&m), cur(m.cur), next(m.next), have_next(m.have_next) {}
    __device__ void close() {
        g->cur = cur;
        g->next = next;
        g->have_next = have_next;
    }
};

__device__ uint32_t mt_u32(MtCur& c) {
    if (c.next == kMt) {
        if (!c.have_next) {
            // a section drew past the prepared block: one reload on this lane
            const uint32_t* src = c.g->s[c.cur];
            uint32_t* dst = c.g->s[c.cur ^ 1];
            int i = 0;
            for (; i < kMt - kMtM; ++i) dst[i] = mt_twist(src[i + kMtM], src[i], src[i + 1]);
            for (; i < kMt - 1; ++i) dst[i] = mt_twist(dst[i + kMtM - kMt], src[i], src[i + 1]);
            dst[kMt - 1] = mt_twist(dst[kMtM - 1], src[kMt - 1], dst[0]);
        }
        c.cur ^= 1;
        c.next = 0;
        c.have_next = 0;
    }
    uint32_t y = c.g->s[c.cur][c.next++];
    y ^= (y >> 11);
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    return y ^ (y >> 18);
}

__device__ uint32_t mt_randint(MtCur& g, uint32_t n) {
    uint32_t used = n;
    used |= used >> 1;
    used |= used >> 2;
    used |= used >> 4;
    used |= used >> 8;
    used |= used >> 16;
    uint32_t i;
    do {
        i = mt_u32(g) & used;
    } while (i > n);
    return i;
}

__device__ double mt_rand(MtCur& g) {
#pragma clang fp contract(off)
    return double(mt_u32(g)) * (1.0 / 4294967295.0);
}

// MTRand::randNorm (include/MersenneTwister.h:349-363) split in two: lane 0 draws
// the accepted polar pairs (x, r) of cnt calls in stream order; then every lane
// forms x * sqrt(-2 log r / r), so a caller's mean + (x f) sd is
// randNorm(mean, sd) to the bit.
__device__ void draw_polar(MtLds& m, int cnt, double* gx, double* gr) {
#pragma clang fp contract(off)
    MtCur g(m);
    for (int c = 0; c < cnt; ++c) {
        double x, y, r;
        do {
            x = 2.0 * mt_rand(g) - 1.0;
            y = 2.0 * mt_rand(g) - 1.0;
            r = x * x + y * y;
        } while (r >= 1.0 || r == 0.0);
        gx[c] = x;
        gr[c] = r;
    }
    g.close();
}

__device__ __forceinline__ double pib(double r, double L) {
#pragma clang fp contract(off)
    return r - L * floor(r * (1.0 / L) + 0.5);
}

struct ChainLds {
    MtLds mt;
    // the move's bead list: entry k is bead (slice[k], ptcl) with weight co[k] =
    // VFactor * shift * tau and positions oldp / newp [at[k]]
    int at[kChainMaxBeads];
    int slice[kChainMaxBeads];
    double co[kChainMaxBeads];
    double oldp[kChainMaxBeads][3];
    double newp[kChainMaxBeads][3];
    // bisection: current positions of the segment (offset o = bead s0 + o)
    double seg[kChainMaxBeads + 2][3];
    unsigned char include[kChainMaxBeads];
    double cumrho[kChainMaxWind];
    int c_old[kChainMaxBeads], c_new[kChainMaxBeads];
    uint32_t inlist[kChainMaxList / 32];   // the correction's interaction list, one bit per bead
    double red4[kChainThreads / 64][4];
    // Marsaglia polar pairs of up to kPolarBeads beads: x (then x f) and r
    double gx[3 * kPolarBeads], gr[3 * kPolarBeads];
    int lvl_cnt;
    double red[kChainThreads / 64][2];   // up to kChainThreads lanes
    int nb, ptcl, accept, done, near;
    double ao, an;
};

__device__ __forceinline__ void load_pos(const ChainRunArgs& A, int m, int i, double r[3]) {
    const double* p = A.pos + (int64_t)m * A.nalloc + i;
    r[0] = p[0];
    r[1] = p[A.plane];
    r[2] = p[2 * A.plane];
}

// The workgroup's pass over the bead list: sum over beads of co * Vnn at the
// old and at the new position (src/action.cpp:1026-1060 with the interaction
// list's cutoff and cells, src/lookuptable.cpp:464-516; every bead on).
template <int MODE>
__device__ void delta_pass(const ChainRunArgs& A, const AzizDev& az, ChainLds& L) {
    const double inv_dr = 1.0 / az.dr;
    const bool use_cells = A.box.ngrid >= 4;
    const int N = A.nalloc, nb = L.nb, me = L.ptcl;
    if (use_cells) {
        for (int k = threadIdx.x; k < nb; k += (int)blockDim.x) {
            const int a = L.at[k];
            L.c_old[k] = cell_of(L.oldp[a][0], L.oldp[a][1], L.oldp[a][2], A.box);
            L.c_new[k] = cell_of(L.newp[a][0], L.newp[a][1], L.newp[a][2], A.box);
        }
        __syncthreads();
    }
    // (bead k, partner j) pairs over every lane: t = k N + j
    double to = 0.0, tn = 0.0;
    const int T = (int)blockDim.x;
    const int qk = T / N, rj = T % N;
    int k = (int)threadIdx.x / N, j = (int)threadIdx.x % N;
    for (int t = threadIdx.x; t < nb * N; t += T) {
        if (j != me) {
            const int m = L.slice[k], a = L.at[k];
            const double* px = A.pos + (int64_t)m * N;
            const double xj = px[j], yj = px[j + A.plane], zj = px[j + 2 * A.plane];
            const double r2o = sep_r2<MODE>(L.oldp[a][0], L.oldp[a][1], L.oldp[a][2], xj, yj, zj, A.box);
            const double r2n = sep_r2<MODE>(L.newp[a][0], L.newp[a][1], L.newp[a][2], xj, yj, zj, A.box);
            bool oko = r2o < A.box.rc2, okn = r2n < A.box.rc2;
            if (use_cells) {
                const int cj = A.cell[m * N + j];
                oko = oko && cells_adjacent(L.c_old[k], cj, A.box.ngrid);
                okn = okn && cells_adjacent(L.c_new[k], cj, A.box.ngrid);
            }
            const double vo = pair_value_from_r2<MODE>(r2o, az, inv_dr);
            const double vn = pair_value_from_r2<MODE>(r2n, az, inv_dr);
            to += oko ? L.co[k] * vo : 0.0;
            tn += okn ? L.co[k] * vn : 0.0;
        }
        j += rj;
        k += qk;
        if (j >= N) {
            j -= N;
            ++k;
        }
    }
    to = wave_sum(to);
    tn = wave_sum(tn);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) {
        L.red[wave][0] = to;
        L.red[wave][1] = tn;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double so = 0.0, sn = 0.0;
        for (int w = 0; w < (int)blockDim.x / 64; ++w) {
            so += L.red[w][0];
            sn += L.red[w][1];
        }
        L.ao = so;
        L.an = sn;
    }
    __syncthreads();
}

// gradVnnSquared(bead1) (src/action.cpp:1282-1337, external gradV = 0) of the
// bead (m, me) placed at p1: the interaction list (partners with r^2 < rc^2
// in neighbouring cells, src/lookuptable.cpp:464-516) as a bit mask, then one
// lane per bead2 with its sum over bead3; returns (on lane 0)
// sum_2 [|F2|^2 + 2 F2.F3] + |F1|^2, F2 = gradV(r2 - r1), F1 = -sum F2.
template <int MODE>
__device__ double gnn_sum(const ChainRunArgs& A, const AzizDev& az, ChainLds& L, int m, int me, const double p1[3],
                          int c1) {
#pragma clang fp contract(off)
    const double inv_dr = 1.0 / az.dr;
    const bool use_cells = A.box.ngrid >= 4;
    const int N = A.nalloc;
    const double* px = A.pos + (int64_t)m * N;
    const double* py = px + A.plane;
    const double* pz = py + A.plane;
    for (int w = threadIdx.x; w < (N + 31) / 32; w += (int)blockDim.x) L.inlist[w] = 0u;
    __syncthreads();
    for (int j = threadIdx.x; j < N; j += (int)blockDim.x) {
        if (j == me) continue;
        const double r2 = sep_r2<MODE>(px[j], py[j], pz[j], p1[0], p1[1], p1[2], A.box);
        const bool in = (r2 < A.box.rc2) && (!use_cells || cells_adjacent(c1, A.cell[m * N + j], A.box.ngrid));
        if (in) atomicOr(&L.inlist[j >> 5], 1u << (j & 31));
    }
    __syncthreads();
    double term = 0.0, f1x = 0.0, f1y = 0.0, f1z = 0.0;
    for (int b2 = threadIdx.x; b2 < N; b2 += (int)blockDim.x) {
        if (!((L.inlist[b2 >> 5] >> (b2 & 31)) & 1u)) continue;
        const double x2 = px[b2], y2 = py[b2], z2 = pz[b2];
        double sx, sy, sz;
        if (MODE == 0) {
            sx = min_image_fast(x2 - p1[0], A.box.L, A.box.invL);
            sy = min_image_fast(y2 - p1[1], A.box.L, A.box.invL);
            sz = min_image_fast(z2 - p1[2], A.box.L, A.box.invL);
        } else {
            sx = min_image_exact(x2 - p1[0], A.box.L, A.box.invL);
            sy = min_image_exact(y2 - p1[1], A.box.L, A.box.invL);
            sz = min_image_exact(z2 - p1[2], A.box.L, A.box.invL);
        }
        const double g2 = grad_factor_from_r2<MODE>(norm2_exact(sx, sy, sz), az, inv_dr);
        const double F2x = g2 * sx, F2y = g2 * sy, F2z = g2 * sz;
        f1x -= F2x;
        f1y -= F2y;
        f1z -= F2z;
        double F3x = 0.0, F3y = 0.0, F3z = 0.0;
        for (int b3 = 0; b3 < N; ++b3) {
            if (b3 == b2 || !((L.inlist[b3 >> 5] >> (b3 & 31)) & 1u)) continue;
            double tx, ty, tz;
            if (MODE == 0) {
                tx = min_image_fast(x2 - px[b3], A.box.L, A.box.invL);
                ty = min_image_fast(y2 - py[b3], A.box.L, A.box.invL);
                tz = min_image_fast(z2 - pz[b3], A.box.L, A.box.invL);
            } else {
                tx = min_image_exact(x2 - px[b3], A.box.L, A.box.invL);
                ty = min_image_exact(y2 - py[b3], A.box.L, A.box.invL);
                tz = min_image_exact(z2 - pz[b3], A.box.L, A.box.invL);
            }
            const double g3 = grad_factor_from_r2<MODE>(norm2_exact(tx, ty, tz), az, inv_dr);
            F3x += g3 * tx;
            F3y += g3 * ty;
            F3z += g3 * tz;
        }
        term += norm2_exact(F2x, F2y, F2z) + 2.0 * ((F2x * F3x + F2y * F3y) + F2z * F3z);
    }
    term = wave_sum(term);
    f1x = wave_sum(f1x);
    f1y = wave_sum(f1y);
    f1z = wave_sum(f1z);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) {
        L.red4[wave][0] = term;
        L.red4[wave][1] = f1x;
        L.red4[wave][2] = f1y;
        L.red4[wave][3] = f1z;
    }
    __syncthreads();
    double G = 0.0;
    if (threadIdx.x == 0) {
        double t = 0.0, fx = 0.0, fy = 0.0, fz = 0.0;
        for (int w = 0; w < (int)blockDim.x / 64; ++w) {
            t += L.red4[w][0];
            fx += L.red4[w][1];
            fy += L.red4[w][2];
            fz += L.red4[w][3];
        }
        G = t + norm2_exact(fx, fy, fz);
    }
    __syncthreads();
    return G;
}

// The shift-1 bead actions' correction gradVFactor tau^3 lambda gradVnnSquared
// (src/action.cpp:497-503) of every listed bead at its old and new position,
// added to L.ao / L.an.
template <int MODE>
__device__ void correction_pass(const ChainRunArgs& A, const AzizDev& az, ChainLds& L) {
    const double t3l = A.tau * A.tau * A.tau * A.lambda;
    for (int k = 0; k < L.nb; ++k) {
        const int m = L.slice[k];
        const double gvf = (m & 1) ? A.gvf1 : A.gvf0;
        if (!(gvf > kCorrEps)) continue;
        const int a = L.at[k];
        const bool use_cells = A.box.ngrid >= 4;
        const int co = use_cells ? cell_of(L.oldp[a][0], L.oldp[a][1], L.oldp[a][2], A.box) : 0;
        const int cn = use_cells ? cell_of(L.newp[a][0], L.newp[a][1], L.newp[a][2], A.box) : 0;
        const double go = gnn_sum<MODE>(A, az, L, m, L.ptcl, L.oldp[a], co);
        const double gn = gnn_sum<MODE>(A, az, L, m, L.ptcl, L.newp[a], cn);
        if (threadIdx.x == 0) {
            L.ao += gvf * t3l * go;
            L.an += gvf * t3l * gn;
        }
    }
    __syncthreads();
}

// Write list entries [k0, nb) to the path, with the cell-list update of
// Path::updateBead; the beads sit on distinct slices, one lane each.
__device__ void commit_list(const ChainRunArgs& A, ChainLds& L, int k0) {
    for (int k = k0 + (int)threadIdx.x; k < L.nb; k += (int)blockDim.x) {
        const int m = L.slice[k], i = L.ptcl;
        const double* r = L.newp[L.at[k]];
        double* q = A.pos + (int64_t)m * A.nalloc + i;
        q[0] = r[0];
        q[A.plane] = r[1];
        q[2 * A.plane] = r[2];
        const int cn = cell_of(r[0], r[1], r[2], A.box);
        cell_update_bead(m, i, cn, A.nalloc, A.num_slices, A.cell, A.label, A.numlab, A.hash);
    }
    __syncthreads();
}

__device__ bool metropolis(ChainLds& L, double arg) {
    MtCur g(L.mt);
    const double u = mt_rand(g);
    g.close();
    const double e = exp(arg);
    if (e <= 1.0 && fabs(u - e) <= 1.0e-10 * e) ++L.near;   // u <= 1: only e <= 1 can be close
    return u < e;
}

// rho0 over Mlen slices (src/action.cpp:119-127)
__device__ double rho0(const ChainRunArgs& A, const double v[3], int Mlen) {
#pragma clang fp contract(off)
    const double norm = pow(4.0 * 3.14159265358979323846 * A.lambda * Mlen * A.tau, -0.5 * 3);
    const double v2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    return norm * exp(-v2 / (4.0 * A.lambda * Mlen * A.tau));
}

// StagingMove::attemptMove (src/move.cpp:1315-1408), lane 0 proposes.
template <int MODE, bool COR>
__device__ void staging(const ChainRunArgs& A, const AzizDev& az, ChainLds& L, int* acc, double* info) {
#pragma clang fp contract(off)
    const int M = A.num_slices;
    __shared__ int s_s0, s_s1, s_len, s_w[3];
    __shared__ double s_vel[3], s_pe[3];
    if (threadIdx.x == 0) {
        MtCur g(L.mt);
        int stageLength = A.Mbar;
        const int s0 = (int)mt_randint(g, M - 1);
        const int s1 = (int)mt_randint(g, A.nalloc - 1);
        if (A.var_updates) stageLength = 2 + (int)mt_randint(g, A.Mbar - 2);
        g.close();
        const int e0 = (s0 + stageLength) % M;
        info[0] = 1.0;
        info[2] = stageLength;
        double ps[3], pe[3];
        load_pos(A, s0, s1, ps);
        load_pos(A, e0, s1, pe);
        for (int k = 0; k < 3; ++k) s_vel[k] = pe[k] - ps[k];
        s_s0 = s0;
        s_s1 = s1;
        s_len = stageLength;
    }
    __syncthreads();
    // MoveBase::sampleWindingSector (src/move.cpp:359-396): the sector weights
    // on the lanes, their running sum on lane 0 in the reference's order
    for (int n = threadIdx.x; n < A.num_wind; n += (int)blockDim.x) {
        double vw[3];
        for (int k = 0; k < 3; ++k) vw[k] = n == 0 ? s_vel[k] : s_vel[k] + A.wind[3 * n + k] * A.side;
        L.cumrho[n] = rho0(A, vw, s_len);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int stageLength = s_len, s0 = s_s0, s1 = s_s1;
        const int e0 = (s0 + stageLength) % M;
        double pe[3];
        load_pos(A, e0, s1, pe);
        double total = L.cumrho[0];
        for (int n = 1; n < A.num_wind; ++n) {
            const double c = L.cumrho[n];
            total += c;
            L.cumrho[n] = L.cumrho[n - 1] + c;
        }
        for (int n = 0; n < A.num_wind; ++n) L.cumrho[n] /= total;
        MtCur g(L.mt);
        const double u = mt_rand(g);
        g.close();
        int idx = 0;
        while (idx < A.num_wind && L.cumrho[idx] < u) ++idx;   // std::lower_bound
        if (idx == A.num_wind) idx = A.num_wind - 1;   // cumrho ends at 1 and u <= 1: not reached
        for (int k = 0; k < 3; ++k) {
            s_w[k] = A.wind[3 * idx + k];
            s_pe[k] = pe[k];
        }
        // the start bead and the stageLength - 1 moved beads
        L.nb = stageLength;
        L.ptcl = s1;
        const double co = A.tau;
        for (int k = 0; k < stageLength; ++k) {
            const int m = (s0 + k) % M;
            L.at[k] = k;
            L.slice[k] = m;
            L.co[k] = ((m & 1) ? A.vf1 : A.vf0) * co;
            load_pos(A, m, s1, L.oldp[k]);
        }
        for (int d = 0; d < 3; ++d) L.newp[0][d] = L.oldp[0][d];
    }
    __syncthreads();
    // MoveBase::newStagingPosition, winding form (src/move.cpp:306-346): each
    // bead starts from the previous bead's new position, so lane 0 composes
    // them in order; the Gaussian deviates come from the split randNorm
    const int nm = s_len - 1;
    for (int kb = 0; kb < nm; kb += kPolarBeads) {
        const int cnt = nm - kb < kPolarBeads ? nm - kb : kPolarBeads;
        mt_refill(L.mt);
        if (threadIdx.x == 0) draw_polar(L.mt, 3 * cnt, L.gx, L.gr);
        __syncthreads();
        for (int c = threadIdx.x; c < 3 * cnt; c += (int)blockDim.x) L.gx[c] = L.gx[c] * sqrt(-2.0 * log(L.gr[c]) / L.gr[c]);
        __syncthreads();
        if (threadIdx.x == 0) {
            const int stageLength = s_len;
            for (int k = kb; k < kb + cnt; ++k) {
                const double f1 = 1.0 * (stageLength - k - 1);
                const double f2 = 1.0 / (1.0 * (stageLength - k));
                const double sq = A.sqrt2LambdaTau * sqrt(f1 * f2);
                const double* npv = L.newp[k];
                double r[3];
                for (int a = 0; a < 3; ++a) r[a] = (s_pe[a] + A.side * s_w[a]) - npv[a];
                for (int a = 0; a < 3; ++a) r[a] *= f2;
                for (int a = 0; a < 3; ++a) r[a] += npv[a];
                for (int a = 0; a < 3; ++a) r[a] = r[a] + L.gx[3 * (k - kb) + a] * sq;
                for (int a = 0; a < 3; ++a) {
                    while (r[a] < -0.5 * A.side) {
                        s_w[a]++;
                        r[a] += A.side;
                    }
                    while (r[a] >= 0.5 * A.side) {
                        s_w[a]--;
                        r[a] -= A.side;
                    }
                }
                for (int a = 0; a < 3; ++a) L.newp[k + 1][a] = pib(r[a], A.side);
            }
        }
        __syncthreads();
    }
    delta_pass<MODE>(A, az, L);
    if (COR) correction_pass<MODE>(A, az, L);
    if (threadIdx.x == 0) {
        const double dlt = L.an - L.ao;
        info[1] = dlt;
        L.accept = metropolis(L, -dlt) ? 1 : 0;
        *acc = L.accept;
    }
    __syncthreads();
    if (L.accept) commit_list(A, L, 1);
}

// BisectionMove::attemptMove, local action (src/move.cpp:1484-1661).
template <int MODE, bool COR>
__device__ void bisection(const ChainRunArgs& A, const AzizDev& az, ChainLds& L, int* acc, double* info) {
#pragma clang fp contract(off)
    const int M = A.num_slices;
    __shared__ int s_levels, s_active, s_s0;
    __shared__ double s_old_delta;
    if (threadIdx.x == 0) {
        MtCur g(L.mt);
        int numLevels = A.levels;
        if (A.var_updates) numLevels = 1 + (int)mt_randint(g, A.levels - 1);
        const int numActive = (1 << numLevels) - 1;
        const int s0 = (int)mt_randint(g, M - 1);
        const int s1 = (int)mt_randint(g, A.nalloc - 1);
        g.close();
        info[0] = 1.0;
        info[2] = numLevels;
        s_levels = numLevels;
        s_active = numActive;
        s_s0 = s0;
        s_old_delta = 0.0;
        L.ptcl = s1;
        L.done = 0;
        for (int o = 0; o <= numActive + 1; ++o) load_pos(A, (s0 + o) % M, s1, L.seg[o]);
        for (int n = 0; n < numActive; ++n) L.include[n] = 1;
    }
    __syncthreads();
    for (int level = s_levels; level > 0; --level) {
        const int shift = 1 << (level - 1);
        if (threadIdx.x == 0) {
            const int s0 = s_s0, numActive = s_active;
            const double co_s = shift * A.tau;
            int nb = 0;
            // this level's beads in order (src/move.cpp:1513-1535)
            for (int k = 1; k * shift < numActive + 1; ++k) {
                const int n = k * shift - 1;
                if (!L.include[n]) continue;
                const int o = n + 1;
                for (int d = 0; d < 3; ++d) L.oldp[n][d] = L.seg[o][d];
                L.include[n] = 0;
                const int m = (s0 + o) % M;
                L.at[nb] = n;
                L.slice[nb] = m;
                L.co[nb] = ((m & 1) ? A.vf1 : A.vf0) * co_s;
                ++nb;
            }
            L.lvl_cnt = nb;
            // at level 1 the earlier levels' beads join (newAction and oldAction over them)
            if (level == 1)
                for (int k = 1; k * shift < numActive + 1; ++k) {
                    const int n = k * shift - 1;
                    if ((n + 1) % 2 == 1) continue;   // this level's beads: already listed
                    const int o = n + 1, m = (s0 + o) % M;
                    L.at[nb] = n;
                    L.slice[nb] = m;
                    L.co[nb] = ((m & 1) ? A.vf1 : A.vf0) * co_s;
                    ++nb;
                }
            L.nb = nb;
        }
        __syncthreads();
        // their proposals (src/move.cpp:467-488): the deviates in stream order,
        // then one lane per bead; a bead's neighbours at +-shift belong to
        // earlier levels, so the beads of a level are independent
        {
            const double delta = A.sqrtLambdaTau * sqrt(1.0 * shift);
            const int nl = L.lvl_cnt;
            for (int kb = 0; kb < nl; kb += kPolarBeads) {
                const int cnt = nl - kb < kPolarBeads ? nl - kb : kPolarBeads;
                mt_refill(L.mt);
                if (threadIdx.x == 0) draw_polar(L.mt, 3 * cnt, L.gx, L.gr);
                __syncthreads();
                for (int c = threadIdx.x; c < 3 * cnt; c += (int)blockDim.x)
                    L.gx[c] = L.gx[c] * sqrt(-2.0 * log(L.gr[c]) / L.gr[c]);
                __syncthreads();
                for (int t = threadIdx.x; t < cnt; t += (int)blockDim.x) {
                    const int n = L.at[kb + t], o = n + 1;
                    double r[3];
                    for (int d = 0; d < 3; ++d) r[d] = pib(L.seg[o + shift][d] - L.seg[o - shift][d], A.side);
                    for (int d = 0; d < 3; ++d) r[d] *= 0.5;
                    for (int d = 0; d < 3; ++d) r[d] += L.seg[o - shift][d];
                    for (int d = 0; d < 3; ++d) r[d] = r[d] + L.gx[3 * t + d] * delta;
                    for (int d = 0; d < 3; ++d) r[d] = pib(r[d], A.side);
                    for (int d = 0; d < 3; ++d) {
                        L.newp[n][d] = r[d];
                        L.seg[o][d] = r[d];
                    }
                }
                __syncthreads();
            }
        }
        delta_pass<MODE>(A, az, L);
        if (COR && shift == 1) correction_pass<MODE>(A, az, L);
        if (threadIdx.x == 0) {
            const double dlt = L.an - L.ao;
            info[1] = dlt;
            if (metropolis(L, -dlt + s_old_delta)) {
                if (level == 1) {
                    *acc = 1;
                    // every bead of the segment takes its proposal
                    for (int n = 0; n < s_active; ++n) {
                        L.at[n] = n;
                        L.slice[n] = (s_s0 + n + 1) % M;
                    }
                    L.nb = s_active;
                    L.accept = 1;
                }
            } else {
                L.accept = 0;
                L.done = 1;   // undoMove: the path still holds every original position
            }
            s_old_delta = dlt;
        }
        __syncthreads();
        if (L.done) break;
    }
    if (L.accept && !L.done) commit_list(A, L, 0);
}

// CenterOfMassMove::attemptMove for a closed worldline (src/move.cpp:1124-1259).
template <int MODE, bool COR>
__device__ void com(const ChainRunArgs& A, const AzizDev& az, ChainLds& L, int* acc, double* info) {
#pragma clang fp contract(off)
    const int M = A.num_slices;
    __shared__ double sh[3];
    if (threadIdx.x == 0) {
        MtCur g(L.mt);
        const int p = (int)mt_randint(g, A.nalloc - 1);
        info[0] = 1.0;
        info[2] = M;
        for (int k = 0; k < 3; ++k) sh[k] = A.com_delta * (-0.5 + mt_rand(g));
        g.close();
        L.ptcl = p;
        L.nb = M;
    }
    __syncthreads();
    for (int m = threadIdx.x; m < M; m += (int)blockDim.x) {
        L.at[m] = m;
        L.slice[m] = m;
        L.co[m] = ((m & 1) ? A.vf1 : A.vf0) * A.tau;
        load_pos(A, m, L.ptcl, L.oldp[m]);
        for (int k = 0; k < 3; ++k) L.newp[m][k] = pib(L.oldp[m][k] + sh[k], A.side);
    }
    __syncthreads();
    delta_pass<MODE>(A, az, L);
    if (COR) correction_pass<MODE>(A, az, L);
    if (threadIdx.x == 0) {
        const double dlt = L.an - L.ao;
        info[1] = dlt;
        L.accept = metropolis(L, -dlt) ? 1 : 0;
        *acc = L.accept;
    }
    __syncthreads();
    if (!L.accept) {
        // the reference shifts the worldline back (it does not restore saved positions)
        for (int m = threadIdx.x; m < M; m += (int)blockDim.x)
            for (int k = 0; k < 3; ++k) L.newp[m][k] = pib(L.newp[m][k] - sh[k], A.side);
        __syncthreads();
    }
    commit_list(A, L, 0);
}

// COR: the action carries the gradient correction; its three-body sums need
// more registers, so that variant runs 512 lanes (two waves per SIMD).
template <int MODE, bool COR>
__global__ __launch_bounds__(COR ? kChainThreadsCor : kChainThreads) void k_chain_run(
    const ChainRunArgs* __restrict__ args, const AzizDev* __restrict__ azs) {
    __shared__ ChainLds L;
    const ChainRunArgs A = args[blockIdx.x];
    const AzizDev az = azs[blockIdx.x];
    for (int t = threadIdx.x; t < kMt; t += (int)blockDim.x) L.mt.s[0][t] = A.mt[t];
    if (threadIdx.x == 0) {
        L.mt.cur = 0;
        L.mt.next = (int)A.mt[kMt + 1];
        L.mt.have_next = 0;
        L.near = 0;
    }
    __syncthreads();
    for (int mv = 0; mv < A.n_moves; ++mv) {
        const int kind = A.kinds[mv];
        int* acc = A.accepted + mv;
        double* info = A.info + 3 * mv;
        mt_refill(L.mt);
        if (threadIdx.x == 0) {
            *acc = 0;
            L.accept = 0;
        }
        if (kind == 0)
            staging<MODE, COR>(A, az, L, acc, info);
        else if (kind == 1)
            bisection<MODE, COR>(A, az, L, acc, info);
        else
            com<MODE, COR>(A, az, L, acc, info);
        __syncthreads();
    }
    for (int t = threadIdx.x; t < kMt; t += (int)blockDim.x) A.mt[t] = L.mt.s[L.mt.cur][t];
    if (threadIdx.x == 0) {
        A.mt[kMt] = (uint32_t)(kMt - L.mt.next);
        A.mt[kMt + 1] = (uint32_t)L.mt.next;
        A.near[0] = L.near;
    }
}

}  // namespace

hipError_t launch_chain_run(int mode, bool cor, hipStream_t st, int n_chains, const ChainRunArgs* args,
                            const AzizDev* azs) {
    (void)hipGetLastError();
    const dim3 g(n_chains), b(cor ? kChainThreadsCor : kChainThreads);
    if (cor) {
        if (mode == 1)
            hipLaunchKernelGGL((k_chain_run<1, true>), g, b, 0, st, args, azs);
        else if (mode == 2)
            hipLaunchKernelGGL((k_chain_run<2, true>), g, b, 0, st, args, azs);
        else
            hipLaunchKernelGGL((k_chain_run<0, true>), g, b, 0, st, args, azs);
    } else {
        if (mode == 1)
            hipLaunchKernelGGL((k_chain_run<1, false>), g, b, 0, st, args, azs);
        else if (mode == 2)
            hipLaunchKernelGGL((k_chain_run<2, false>), g, b, 0, st, args, azs);
        else
            hipLaunchKernelGGL((k_chain_run<0, false>), g, b, 0, st, args, azs);
    }
    return hipGetLastError();
}

}  // namespace pimc
