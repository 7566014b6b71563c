// pimc_pair.hip -- all-bead-pairs kernels for gfx950.
//
//  k_pair_action : LocalAction::V(int) for every slice of one or many walkers
//                  (src/action.cpp:902-947) and potentialAction() (:456-472).
//  k_vnn         : LocalAction::Vnn(bead) (src/action.cpp:1044-1060) for a list of
//                  beads at old and/or new positions: the Delta V of a move.
//  k_sep_hist    : ActionBase::updateSepHist over one slice (src/action.cpp:216-224).
//  k_kinetic     : the link sum of EnergyEstimator (src/estimator.cpp:966-976).
//  cell kernels  : LookupTable::updateGrid / updateBead / updateInteractionList
//                  (src/lookuptable.cpp:314-516), bit-exact integer results.
//  k_aziz_batch  : PotentialBase::V / gradV / grad2V over a batch of separations.
//
// No MFMA: the pair term is a transcendental sum.  Bead positions are FP64 SoA
// planes; a slice is staged in LDS once per workgroup and every lane walks a
// strided range of the slice's flattened pair index space.  Sums are FP64 with
// fixed-order wave/workgroup trees and a fixed-order final pass, so results do
// not depend on how many other slices or walkers share the launch.
#include <algorithm>
#include <type_traits>
#include <hip/hip_runtime.h>

#include "pimc_kernels.hpp"

namespace pimc {

// ---------------------------------------------------------------------------
// All i<j pairs of each slice, row tiles (RowTile in pimc_kernels.hpp).
// ---------------------------------------------------------------------------
// A loop-invariant value held in a vector register, so that an instruction
// with two loop-invariant operands reads at most one scalar register (the
// constant bus of gfx950 takes one) and needs no copy per pair.
template <typename T>
__device__ __forceinline__ T in_vgpr(T x) {
    asm("" : "+v"(x));
    return x;
}

// Mixed-mode bin of one pair, the row bead's part precomputed.  The minimum
// image in box units: u = x_j/L + (1/2 - x_i/L) = (x_j - x_i)/L + 1/2 and the
// separation is L (fract(u) - 1/2), here scaled to bins by S = L/dr, so the bin
// is the integer part of |w| (an FP32 root refined by one Newton step in FP64).
// u carries an FP64 rounding of ~1e-16, |w| an error of ~1e-8 bin at L = 45 A:
// the bin differs from the exact one only when |r|/dr lies that close to an
// integer (the image choice itself can differ only at |x_j - x_i| = L/2, where
// both images are equally far).
struct RowMixed {
    double cx, cy, cz;   // 1/2 - x_i/L per axis
    double invL, S, Sh;  // 1/L, L/dr, -S/2
};

// |w| in bins (>= 0); y_out: the reciprocal-root estimate of |w|
__device__ __forceinline__ double sep_bins_mixed(const RowMixed& q, double xj, double yj, double zj, double* y_out) {
    const double wx = fma(__builtin_amdgcn_fract(fma(xj, q.invL, q.cx)), q.S, q.Sh);
    const double wy = fma(__builtin_amdgcn_fract(fma(yj, q.invL, q.cy)), q.S, q.Sh);
    const double wz = fma(__builtin_amdgcn_fract(fma(zj, q.invL, q.cz)), q.S, q.Sh);
    const double r2 = fma(wz, wz, fma(wy, wy, wx * wx));
    // |w|: the hardware FP64 reciprocal-root estimate (relative error ~2^-23)
    // and one Newton step, |w| = r0 + (r2 - r0^2) y / 2 (~2^-46)
    const double y = __builtin_amdgcn_rsq(r2);
    *y_out = y;
    const double r0 = r2 * y;
    const double res = fma(-r0, r0, r2);
    return fma(res * y, 0.5, r0);
}

__device__ __forceinline__ int bin_mixed(const RowMixed& q, double xj, double yj, double zj) {
    double y;
    return (int)sep_bins_mixed(q, xj, yj, zj, &y);
}

// One lane: row bead i (held in registers) against partners j = (i + d) mod n for
// d = d, d + G, ... < d_end (0 <= i < n, 1 <= d <= n/2), T iterations, the
// count of the tile's first lane group (uniform, so the loop runs on a scalar
// counter; a lane group with one offset fewer masks its last iteration).
// PS > 0: the slice is staged in LDS as x | y | z planes PSL doubles apart, the
// partner walked by its byte offset, so one address register serves the three
// reads; PS == 0: px, py, pz in global memory.  DUP: each LDS plane holds the
// slice followed by its first n/2 + 1 beads again, so j = i + d < 3n/2 needs no
// wrap; otherwise the cursor wraps with an unsigned min.
// mask: the offsets reach d = n/2 of an even n, whose pairs count once (rows
// i < n/2).  The group holding offset n/2 = d_end - 1 has it as its largest
// offset, and its offset count, ceil((dn - g)/G) for g = (dn - 1) mod G, is T:
// so d = n/2 only ever falls in the last iteration, where the lane drops it.
template <int PS>
constexpr bool pair_lds_dup() {
    // at 2048 beads the extra 24 KB per workgroup costs more occupancy than the
    // wrap it saves (N = 2048 x 250: 1.013 ms against 0.952)
    return PS > 0 && PS <= 1024;
}
template <int PS>
constexpr int pair_lds_plane() {
    return pair_lds_dup<PS>() ? PS + PS / 2 + 64 : PS;   // room for n + n/2 + 1 beads, 64-bead aligned
}

// One mixed-mode pair term against the row bead of q.  FAR: a wave whose pairs
// all lie in far bins (k_far <= k < table_length) takes aziz_vfar (no exp2,
// damping test or clamp), a wave whose pairs all lie at or beyond D (k_d <= k <
// table_length) aziz_vmid (no damping test or clamp), any other wave the full
// term.  The tiers differ from the full term by at most an FP64 rounding and a
// ~1e-12 relative 1/k^2, so which one a pair takes depends on its wave's other
// pairs but its value does not beyond that.  av: a with the loop-invariant
// operands in vector registers; tb: the tiers' bin bounds.
// The tier tests run on the bin as an FP64 integer (floor of |w|, equal to the
// truncated int for |w| >= 0), so the far tier needs no integer round trip.
struct TierBounds {
    double far_lo, mid_lo, hi;   // k_far, k_d, table_length
};
template <bool BETA0, bool FAR>
__device__ __forceinline__ double pair_term_mixed(const RowMixed& q, double x, double y, double z, const AzizDev& a,
                                                  const AzizDev& av, int lenm1, const TierBounds& tb) {
    double yr;
    const double w = sep_bins_mixed(q, x, y, z, &yr);
    if (FAR) {
        const double kd = floor(w);
        if (__builtin_amdgcn_ballot_w64(!(kd >= tb.far_lo && kd < tb.hi)) == 0) return aziz_vfar(kd, yr, av);
        if (__builtin_amdgcn_ballot_w64(!(kd >= tb.mid_lo && kd < tb.hi)) == 0) return aziz_vmid<BETA0>(kd, yr, av);
    }
    const int k = (int)w;
    // k_eps <= k < tableLength, else V = 0 (include/potential.h:252-257): the
    // clamped bin keeps the arithmetic finite, and only a wave holding such a
    // pair pays for the select
    const int kk = min(max(k, a.k_eps), lenm1);
    double v = aziz_vmix<BETA0>(kk, av);
    if (__builtin_amdgcn_ballot_w64(k != kk)) {
        asm volatile("");
        v = (k != kk) ? 0.0 : v;
    }
    return v;
}

template <int MODE, int PS, bool FACTORS, bool BETA0, bool NOWRAP, bool FAR>
__device__ __forceinline__ double row_loop(const double* px, const double* py, const double* pz, int n, int i,
                                           int d, int d_end, int T, int G, bool mask, const BoxDev& box,
                                           const AzizDev& a, double inv_dr, const WormDev& worm, int slice) {
    constexpr bool DUP = pair_lds_dup<PS>();
    constexpr int PSL = pair_lds_plane<PS>();
    const double xi = px[i];
    const double yi = PS ? px[PSL + i] : py[i];
    const double zi = PS ? px[2 * PSL + i] : pz[i];
    int j0 = i + d;
    if (!DUP && j0 >= n) j0 -= n;
    constexpr unsigned kW = PS ? 8u : 1u;   // units of the partner cursor: bytes (LDS) or beads
    const unsigned step = in_vgpr((unsigned)(G % n) * kW);
    const unsigned span = in_vgpr((unsigned)n * kW);
    unsigned jc = (unsigned)j0 * kW;
    const int skip_d = (mask && i >= (n >> 1)) ? (n >> 1) : -1;
    RowMixed q;
    if (MODE == 0) {
        q.invL = box.invL;
        q.cx = fma(-xi, box.invL, 0.5);
        q.cy = fma(-yi, box.invL, 0.5);
        q.cz = fma(-zi, box.invL, 0.5);
        q.S = box.L * inv_dr;
        q.Sh = in_vgpr(-0.5 * q.S);
    }
    const int lenm1 = in_vgpr(a.table_length - 1);
    const TierBounds tb{(double)a.k_far, (double)a.k_d, (double)a.table_length};
    AzizDev av = a;
    if (MODE == 0) {
        av.rep_c1 = in_vgpr(a.rep_c1);
        av.eC8k = in_vgpr(a.eC8k);
    }
    const char* lb = reinterpret_cast<const char*>(px);
    // A partner's position (load) and its pair term (eval).  Evaluating two pairs'
    // branch-free parts back to back, fenced so the scheduler interleaves them,
    // measured slower at N = 2048 x 250 (1.009 ms against 0.981) and N = 256 x 167.
    struct Pos {
        double x, y, z;
        unsigned c;
    };
    auto load = [&](unsigned c) -> Pos {
        Pos p;
        p.c = c;
        if (PS) {
            p.x = *reinterpret_cast<const double*>(lb + c);
            p.y = *reinterpret_cast<const double*>(lb + c + 8 * PSL);
            p.z = *reinterpret_cast<const double*>(lb + c + 16 * PSL);
        } else {
            p.x = px[c];
            p.y = py[c];
            p.z = pz[c];
        }
        return p;
    };
    auto eval = [&](const Pos& p) -> double {
        double v;
        if (MODE == 0) {
            v = pair_term_mixed<BETA0, FAR>(q, p.x, p.y, p.z, a, av, lenm1, tb);
        } else {
            const double r2 = sep_r2<MODE>(xi, yi, zi, p.x, p.y, p.z, box);
            v = pair_value_from_r2<MODE>(r2, a, inv_dr);
        }
        if (FACTORS) {
            int j = (int)(p.c / kW);
            if (DUP && j >= n) j -= n;
            const int lo = i < j ? i : j, hi = i < j ? j : i;
            v *= pair_factor(bead_state(worm, slice, lo), bead_state(worm, slice, hi));
        }
        return v;
    };
    auto advance = [&]() {
        const unsigned jn = jc + step;
        jc = (DUP || NOWRAP) ? jn : min(jn, jn - span);   // unsigned: jn - span wraps above jn while jn < span
    };
    double acc = 0.0;
    auto add = [&](double v) { acc += v; };
    // the last offset is present for the lane groups whose d reaches it (a lane
    // past d_end reads bead 0 and drops it)
    const bool ok = (d + (T - 1) * G < d_end) && (d + (T - 1) * G) != skip_d;
    if (T == 1) {
        const double v = eval(load(ok ? jc : 0u));
        return ok ? v : 0.0;
    }
    // each partner is read one iteration ahead of its evaluation
    Pos p = load(jc);
    for (int t = 1; t < T - 1; ++t) {
        advance();
        const Pos pn = load(jc);
        add(eval(p));
        p = pn;
    }
    advance();
    const Pos pl = load(ok ? jc : 0u);
    add(eval(p));
    const double v = eval(pl);
    acc += ok ? v : 0.0;
    return acc;
}

// The loop variants, each fixed per tile so none costs a branch per pair:
// FACTORS (a worm on the slice), BETA0 (beta = 0, mixed mode) and NOWRAP (no
// row + offset of the tile reaches n, so the partner cursor needs no wrap even
// without the duplicated staging).
template <int MODE, int PS, bool FAR = false>
__device__ __forceinline__ double row_tile(const double* px, const double* py, const double* pz, int n, int i, int d,
                                           int d_end, int T, int G, bool mask, const BoxDev& box, const AzizDev& a,
                                           double inv_dr, bool factors, bool nowrap, const WormDev& worm,
                                           int slice) {
#define PIMC_ROW(F_, B_, W_) \
    row_loop<MODE, PS, F_, B_, W_, FAR>(px, py, pz, n, i, d, d_end, T, G, mask, box, a, inv_dr, worm, slice)
    if (factors) return PIMC_ROW(true, false, false);
    const bool beta0 = MODE == 0 && a.rep_c2 == 0.0;
    if (!pair_lds_dup<PS>() && nowrap) return beta0 ? PIMC_ROW(false, true, true) : PIMC_ROW(false, false, true);
    return beta0 ? PIMC_ROW(false, true, false) : PIMC_ROW(false, false, false);
#undef PIMC_ROW
}

// Spatial order of a slice: the index of each bead's cell along a Hilbert curve
// through a 32^3 grid over the box (J. Skilling, "Programming the Hilbert
// curve", AIP Conf. Proc. 707, 381 (2004): axes to transposed index, then the
// bits interleaved, x the most significant of each triple).  Consecutive cells
// along the curve are face neighbours, so runs of consecutive beads form
// compact clusters (on a random 2048-bead slice 66.6% of the window form's
// waves are all far, against 62.6% along a Morton curve).
__device__ __forceinline__ unsigned hilbert_cell(double x, double y, double z, double invL) {
    auto cell = [&](double v) { return (unsigned)min(max((int)floor(fma(v, invL, 0.5) * 32.0), 0), 31); };
    unsigned X[3] = {cell(x), cell(y), cell(z)};
    for (unsigned Q = 16; Q > 1; Q >>= 1) {
        const unsigned P = Q - 1;
        for (int i = 0; i < 3; ++i) {
            if (X[i] & Q) {
                X[0] ^= P;
            } else {
                const unsigned t = (X[0] ^ X[i]) & P;
                X[0] ^= t;
                X[i] ^= t;
            }
        }
    }
    X[1] ^= X[0];
    X[2] ^= X[1];
    unsigned t = 0;
    for (unsigned Q = 16; Q > 1; Q >>= 1)
        if (X[2] & Q) t ^= Q - 1;
    X[0] ^= t;
    X[1] ^= t;
    X[2] ^= t;
    unsigned m = 0;
    for (int b = 4; b >= 0; --b) m = (m << 3) | (((X[0] >> b) & 1u) << 2) | (((X[1] >> b) & 1u) << 1) | ((X[2] >> b) & 1u);
    return m;
}

// Spatial order of each slice (one workgroup per slice): keys (Hilbert cell << 11
// | bead), unique, sorted ascending by a bitonic network over the next power of
// two (padding keys 0xffffffff); perm[pos + t] is the bead of the t-th key, pos
// the slice's offset in the x plane (so one int array the size of the position
// planes holds every slice's order).  The order depends only on the positions.
// n <= kLdsBeads (bead indices take 11 bits).
constexpr int kOrderThreads = 1024;
__global__ __launch_bounds__(kOrderThreads) void k_slice_order(const SliceDesc* __restrict__ slices,
                                                               const double* __restrict__ pos, double invL,
                                                               int* __restrict__ perm) {
    static_assert(kLdsBeads <= 2048, "bead index keys take 11 bits");
    __shared__ unsigned key[kLdsBeads];
    const SliceDesc sd = slices[blockIdx.x];
    const int n = sd.n;
    if (n > kLdsBeads) return;
    const double* px = pos + sd.pos;
    const double* py = px + sd.plane;
    const double* pz = py + sd.plane;
    int P = 1;
    while (P < n) P <<= 1;
    for (int t = threadIdx.x; t < P; t += kOrderThreads)
        key[t] = t < n ? (hilbert_cell(px[t], py[t], pz[t], invL) << 11) | (unsigned)t : 0xffffffffu;
    __syncthreads();
    // lane t compares keys a and a + j of the 128 keys [128 w, 128 w + 128) of its
    // wave w whenever j <= 64, so a pass with j <= 64 followed by another needs
    // only the wave's own LDS order (51 of the 66 passes at 2048 beads)
    static_assert(kLdsBeads / 2 <= kOrderThreads, "one compare-exchange per lane and pass");
    const int t = threadIdx.x;
    for (int k = 2; k <= P; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            if (t < (P >> 1)) {
                const int a = ((t & ~(j - 1)) << 1) | (t & (j - 1));   // pair (a, a + j)
                const unsigned ka = key[a], kb = key[a + j];
                if ((ka > kb) == ((a & k) == 0)) {
                    key[a] = kb;
                    key[a + j] = ka;
                }
            }
            const int j_next = j > 1 ? j >> 1 : k;   // the next pass's distance
            if (j <= 64 && j_next <= 64) {
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            } else {
                __syncthreads();
            }
        }
    }
    __syncthreads();
    int* out = perm + sd.pos;
    for (int t = threadIdx.x; t < n; t += kOrderThreads) out[t] = (int)(key[t] & 0x7ffu);
}

// Persistent: the workgroups of one XCD take consecutive tile ranges (so a
// slice is staged into one XCD's L2), each range a run of tiles grouped by
// slice; a slice is restaged in LDS only when it changes.  One FP64 partial per
// tile, whatever the grid.  PS: LDS plane stride in doubles (slices of up to PS
// beads are staged; 0: no staging).
template <int MODE, int PS>
__global__ __launch_bounds__(kPairThreads) void k_pair_action(const RowTile* __restrict__ tiles,
                                                              const int* __restrict__ ranges,
                                                              const WalkerDesc* __restrict__ walkers,
                                                              const double* __restrict__ pos, AzizDev az,
                                                              double* __restrict__ partials) {
    // x | y | z of one slice, PSL each, then the block sum's wave partials: a
    // static array, so its address is a constant folded into the reads' offsets
    constexpr int PSL = pair_lds_plane<PS>();
    __shared__ double lds[3 * PSL + kPairThreads / 64];
    double* red = lds + 3 * PSL;
    const int r = xcd_contiguous_item(blockIdx.x, gridDim.x);
    const int k_begin = ranges[r], k_end = ranges[r + 1];
    if (k_begin >= k_end) return;
    int staged = -1;
    RowTile tl = tiles[k_begin];
    for (int k = k_begin; k < k_end; ++k) {
        // the next tile's descriptor loads while this tile computes (clamped index:
        // a plain uniform load, no select of the struct)
        const RowTile tl_next = tiles[k + 1 < k_end ? k + 1 : k];
        const WalkerDesc& wk = walkers[tl.walker];
        AzizDev a = az;
        a.table_length = wk.table_length;
        const double inv_dr = 1.0 / a.dr;
        const int n = tl.n;
        const double* px = pos + tl.pos;
        const double* py = px + tl.plane;
        const double* pz = py + tl.plane;
        const bool in_lds = PS > 0 && n <= PS;
        if (in_lds && staged != tl.slice) {
            __syncthreads();   // the previous slice is no longer read
            for (int t = threadIdx.x; t < n; t += kPairThreads) {
                const double x = px[t], y = py[t], z = pz[t];
                lds[t] = x;
                lds[PSL + t] = y;
                lds[2 * PSL + t] = z;
                if (pair_lds_dup<PS>() && t <= (n >> 1)) {
                    lds[n + t] = x;
                    lds[PSL + n + t] = y;
                    lds[2 * PSL + n + t] = z;
                }
            }
            staged = tl.slice;
            __syncthreads();
        }
        const int R = n < kPairThreads ? n : kPairThreads;
        const int G = kPairThreads / R;
        const int grp = threadIdx.x / R;
        const int i = tl.i0 + (threadIdx.x - grp * R);
        double acc = 0.0;
        if (grp < G && i < tl.i_end) {
            const bool factors = worm_touches_slice(wk.worm, tl.m);
            const int d = tl.d0 + grp, d_end = tl.d0 + tl.dn;
            const int T = (tl.dn + G - 1) / G;
            const bool mask = ((n & 1) == 0) && d_end > (n >> 1) && tl.i_end > (n >> 1);
            const bool nowrap = (tl.i_end - 1) + (d_end - 1) < n;
            acc = in_lds ? row_tile<MODE, PS>(lds, nullptr, nullptr, n, i, d, d_end, T, G, mask, wk.box, a, inv_dr,
                                              factors, nowrap, wk.worm, tl.m)
                         : row_tile<MODE, 0>(px, py, pz, n, i, d, d_end, T, G, mask, wk.box, a, inv_dr, factors,
                                             nowrap, wk.worm, tl.m);
        }
        const double s = block_sum<kPairThreads>(acc, red);
        if (threadIdx.x == 0) partials[k] = s;
        tl = tl_next;
    }
}

// ---------------------------------------------------------------------------
// Window form (mixed mode, slices in spatial order): tiles of kWinRows rows x at
// most kWinOffsets offsets on slices of kWinMinBeads beads and more.  A tile's partners
// j = (i + d) mod n, i in [i0, i0 + 32), d in [d0, d0 + dn), are the cyclic run
// of 32 + dn - 1 beads from (i0 + d0) mod n; only that window is staged, 26 KB
// of LDS against 48 KB for a 2048-bead slice, so a CU holds six workgroups
// instead of three.  Each wave takes 8 rows x 8 offset
// groups, so its 64 pairs join two small clusters of beads and are mostly all
// far or all near (the FAR path); the partner cursor walks the window without a
// wrap.  Slices of fewer than 64 beads are staged whole, row_tile's layout.
// ---------------------------------------------------------------------------
constexpr int kWinRows = 32;      // rows per tile: 4 waves x 8
constexpr int kWinMinBeads = 64;  // window tiles from this slice size (32 + n/2 - 1 < n)
constexpr int kWinGroups = 8;     // offset groups per wave (lane bits 3..5)
// offsets per tile, at most (N = 2048 x 250, interleaved runs: 256 / 512 / 1024
// offsets 711 / 687 / 676 us, bead-order staging 801 us)
constexpr int kWinOffsets = 1024;
// window slots per plane: 32 + 1024 - 1 staged, the masked last iteration of a
// lane reads at most 7 slots further; rounded up to a multiple of 64, plus 2 so
// the plane stride (8720 bytes) is no multiple of 512: the three reads of a
// partner are then single reads with immediate offsets from the cursor (a
// stride of 64 x 512 bytes pairs the x and y reads into one read2st64 whose
// offset units cannot hold the next slot's 64 bytes, which costs an add per pair)
constexpr int kWinBeads = 1090;
static_assert(kWinRows + kWinOffsets - 1 + kWinGroups - 1 <= kWinBeads, "window too small");
static_assert(pair_lds_plane<256>() <= kWinBeads, "small slices are staged whole in the window's planes");

// One lane of a window tile: row bead (xi, yi, zi) of global index `row`,
// partners at window slots slot0, slot0 + 8, ... (offsets d, d + 8, ... < d_end,
// T iterations as in row_loop, the same masking of d = n/2), w0: the slice index
// of window slot 0 (FACTORS: partner index (w0 + slot) mod n, bead order).
template <bool FACTORS, bool BETA0>
__device__ __forceinline__ double row_loop_win(const double* lw, double xi, double yi, double zi, int row, int n,
                                               int slot0, int d, int d_end, int T, bool mask, int w0,
                                               const BoxDev& box, const AzizDev& a, double inv_dr,
                                               const WormDev& worm, int slice) {
    RowMixed q;
    q.invL = box.invL;
    q.cx = fma(-xi, box.invL, 0.5);
    q.cy = fma(-yi, box.invL, 0.5);
    q.cz = fma(-zi, box.invL, 0.5);
    q.S = box.L * inv_dr;
    q.Sh = in_vgpr(-0.5 * q.S);
    const int lenm1 = in_vgpr(a.table_length - 1);
    const TierBounds tb{(double)a.k_far, (double)a.k_d, (double)a.table_length};
    AzizDev av = a;
    av.rep_c1 = in_vgpr(a.rep_c1);
    av.eC8k = in_vgpr(a.eC8k);
    const int skip_d = (mask && row >= (n >> 1)) ? (n >> 1) : -1;
    const char* lb = reinterpret_cast<const char*>(lw);
    struct Pos {
        double x, y, z;
        unsigned c;
    };
    auto load = [&](unsigned c) -> Pos {   // c: byte offset of the slot in the x plane
        Pos p;
        p.c = c;
        p.x = *reinterpret_cast<const double*>(lb + c);
        p.y = *reinterpret_cast<const double*>(lb + c + 8 * kWinBeads);
        p.z = *reinterpret_cast<const double*>(lb + c + 16 * kWinBeads);
        return p;
    };
    auto eval = [&](const Pos& p) -> double {
        double v = pair_term_mixed<BETA0, true>(q, p.x, p.y, p.z, a, av, lenm1, tb);
        if (FACTORS) {
            int j = w0 + (int)(p.c >> 3);
            if (j >= n) j -= n;
            const int lo = row < j ? row : j, hi = row < j ? j : row;
            v *= pair_factor(bead_state(worm, slice, lo), bead_state(worm, slice, hi));
        }
        return v;
    };
    constexpr unsigned kStep = 8u * kWinGroups;
    unsigned jc = (unsigned)slot0 * 8u;
    const bool ok = (d + (T - 1) * kWinGroups < d_end) && (d + (T - 1) * kWinGroups) != skip_d;
    if (T == 1) {
        const double v = eval(load(ok ? jc : 0u));
        return ok ? v : 0.0;
    }
    double acc = 0.0;
    Pos p = load(jc);
    // two partners per trip, the second read taking its cursor step as an
    // immediate offset; the sum keeps the partners' order
    int t = 1;
    for (; t + 1 < T - 1; t += 2) {
        const Pos p1 = load(jc + kStep);
        acc += eval(p);
        const Pos p2 = load(jc + 2 * kStep);
        acc += eval(p1);
        p = p2;
        jc += 2 * kStep;
    }
    if (t < T - 1) {
        jc += kStep;
        const Pos pn = load(jc);
        acc += eval(p);
        p = pn;
    }
    jc += kStep;
    const Pos pl = load(ok ? jc : 0u);
    acc += eval(p);
    const double v = eval(pl);
    acc += ok ? v : 0.0;
    return acc;
}

__global__ __launch_bounds__(kPairThreads) void k_pair_window(const RowTile* __restrict__ tiles,
                                                              const int* __restrict__ ranges,
                                                              const WalkerDesc* __restrict__ walkers,
                                                              const double* __restrict__ pos, AzizDev az,
                                                              double* __restrict__ partials,
                                                              const int* __restrict__ perm) {
    constexpr int PSL = pair_lds_plane<256>();   // plane stride of a small slice staged whole
    __shared__ double lds[3 * kWinBeads + kPairThreads / 64];
    double* red = lds + 3 * kWinBeads;
    const int r = xcd_contiguous_item(blockIdx.x, gridDim.x);
    const int k_begin = ranges[r], k_end = ranges[r + 1];
    if (k_begin >= k_end) return;
    int staged = -1;   // the small slice staged whole, if any
    RowTile tl = tiles[k_begin];
    for (int k = k_begin; k < k_end; ++k) {
        const RowTile tl_next = tiles[k + 1 < k_end ? k + 1 : k];
        const WalkerDesc& wk = walkers[tl.walker];
        AzizDev a = az;
        a.table_length = wk.table_length;
        const double inv_dr = 1.0 / a.dr;
        const int n = tl.n;
        const double* px = pos + tl.pos;
        const double* py = px + tl.plane;
        const double* pz = py + tl.plane;
        // spatial order, or bead order where the worm's pair factors need the
        // bead index
        const bool factors = worm_touches_slice(wk.worm, tl.m);
        const int* pm = factors ? nullptr : perm + tl.pos;
        const bool mask_n = (n & 1) == 0;
        const int d_end = tl.d0 + tl.dn;
        double acc = 0.0;
        if (n >= kWinMinBeads) {
            // the window: slots t < 32 + dn - 1 hold beads (w0 + t) mod n (i0 + d0
            // < 3n/2, 32 + dn - 1 < n: one wrap at most)
            int w0 = tl.i0 + tl.d0;
            if (w0 >= n) w0 -= n;
            const int len = (tl.i_end - tl.i0) + tl.dn - 1;
            __syncthreads();   // the previous window is no longer read
            for (int t = threadIdx.x; t < len; t += kPairThreads) {
                int b = w0 + t;
                if (b >= n) b -= n;
                const int src = pm ? min(max(pm[b], 0), n - 1) : b;
                lds[t] = px[src];
                lds[kWinBeads + t] = py[src];
                lds[2 * kWinBeads + t] = pz[src];
            }
            staged = -1;
            __syncthreads();
            const int row = tl.i0 + 8 * (threadIdx.x >> 6) + (threadIdx.x & 7);
            const int c = (threadIdx.x >> 3) & 7;
            if (row < tl.i_end) {
                const int rs = pm ? min(max(pm[row], 0), n - 1) : row;
                const double xi = px[rs], yi = py[rs], zi = pz[rs];
                const int T = (tl.dn + kWinGroups - 1) / kWinGroups;
                const bool mask = mask_n && d_end > (n >> 1) && tl.i_end > (n >> 1);
                const int slot0 = (row - tl.i0) + c;
                const int d = tl.d0 + c;
#define PIMC_WIN(F_, B_) \
    row_loop_win<F_, B_>(lds, xi, yi, zi, row, n, slot0, d, d_end, T, mask, w0, wk.box, a, inv_dr, wk.worm, tl.m)
                if (factors)
                    acc = PIMC_WIN(true, false);
                else
                    acc = a.rep_c2 == 0.0 ? PIMC_WIN(false, true) : PIMC_WIN(false, false);
#undef PIMC_WIN
            }
        } else {
            if (staged != tl.slice) {
                __syncthreads();
                for (int t = threadIdx.x; t < n; t += kPairThreads) {
                    const int src = pm ? min(max(pm[t], 0), n - 1) : t;
                    const double x = px[src], y = py[src], z = pz[src];
                    lds[t] = x;
                    lds[PSL + t] = y;
                    lds[2 * PSL + t] = z;
                    if (t <= (n >> 1)) {
                        lds[n + t] = x;
                        lds[PSL + n + t] = y;
                        lds[2 * PSL + n + t] = z;
                    }
                }
                staged = tl.slice;
                __syncthreads();
            }
            const int G = kPairThreads / n;
            const int grp = threadIdx.x / n;
            const int i = tl.i0 + (threadIdx.x - grp * n);
            if (grp < G && i < tl.i_end) {
                const int d = tl.d0 + grp;
                const int T = (tl.dn + G - 1) / G;
                const bool mask = mask_n && d_end > (n >> 1) && tl.i_end > (n >> 1);
                acc = row_tile<0, 256, true>(lds, nullptr, nullptr, n, i, d, d_end, T, G, mask, wk.box, a, inv_dr,
                                             factors, true, wk.worm, tl.m);
            }
        }
        const double s = block_sum<kPairThreads>(acc, red);
        if (threadIdx.x == 0) partials[k] = s;
        tl = tl_next;
    }
}

// Per-slice sums of the item partials (items are grouped by slice, in order) and
// the per-walker potential action (src/action.cpp:462-469).  One workgroup of
// 1024 lanes: the partials and the slices' item ranges are first pulled into LDS
// with every load in flight; then each lane sums the items of one slice in item
// order, and one wave per walker sums its slices (lanes strided, fixed shuffle
// tree), so the result depends only on the work list, never on timing.
constexpr int kFinThreads = 1024;
constexpr int kFinLds = 5120;      // partials staged in LDS (beyond: read from HBM)
constexpr int kFinSlices = 1024;   // slice ranges / sums staged in LDS (beyond: HBM)
__global__ __launch_bounds__(kFinThreads) void k_pair_finalize(int n_items, const SliceDesc* __restrict__ slices,
                                                              int n_slices, const WalkerDesc* __restrict__ walkers,
                                                              int n_walkers, const double* __restrict__ partials,
                                                              double* __restrict__ v_slice,
                                                              double* __restrict__ action,
                                                              unsigned long long* done, unsigned long long seq,
                                                              const double* __restrict__ g2_part, int g2_blocks,
                                                              double* __restrict__ g2_slice) {
    __shared__ double sp[kFinLds];
    __shared__ double vs[kFinSlices];
    __shared__ double gs[kFinSlices];
    __shared__ int fi[kFinSlices + 1];
    const bool staged = n_items <= kFinLds;
    const bool small = n_slices <= kFinSlices;
    if (staged)
        for (int t = threadIdx.x; t < n_items; t += kFinThreads) sp[t] = partials[t];
    if (small) {
        for (int t = threadIdx.x; t < n_slices; t += kFinThreads) fi[t] = slices[t].first_item;
        if (threadIdx.x == 0) fi[n_slices] = n_items;
    }
    __syncthreads();
    const double* src = staged ? sp : partials;
    for (int sl = threadIdx.x; sl < n_slices; sl += kFinThreads) {
        const int b = small ? fi[sl] : slices[sl].first_item;
        const int e = small ? fi[sl + 1] : ((sl + 1 < n_slices) ? slices[sl + 1].first_item : n_items);
        double v = 0.0;
        for (int k = b; k < e; ++k) v += src[k];
        v_slice[sl] = v;
        if (small) vs[sl] = v;
        if (g2_part) {
            double g = 0.0;
            for (int b2 = 0; b2 < g2_blocks; ++b2) g += g2_part[(int64_t)sl * g2_blocks + b2];
            g2_slice[sl] = g;
            if (small) gs[sl] = g;
        }
    }
    __threadfence_block();   // v_slice is re-read below when the slices do not fit in LDS
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    constexpr int kWaves = kFinThreads / 64;
    for (int w = wave; w < n_walkers; w += kWaves) {
        const WalkerDesc& W = walkers[w];
        const double tau_s = W.shift * W.tau;
        double tot = 0.0;
        for (int m = lane; m < W.num_slices; m += 64) {
            const double vf = (m & 1) ? W.vfactor1 : W.vfactor0;
            const double v = small ? vs[W.first_slice + m] : v_slice[W.first_slice + m];
            double u = vf * tau_s * (0.0 + v);
            // potentialAction() correction (src/action.cpp:465-467)
            const double gvf = (m & 1) ? W.gvf1 : W.gvf0;
            if (g2_part && gvf > kCorrEps) {
                const double g = small ? gs[W.first_slice + m] : g2_slice[W.first_slice + m];
                u += gvf * tau_s * tau_s * tau_s * W.lambda * g;
            }
            tot += u;
        }
        tot = wave_sum(tot);
        if (lane == 0) {
            action[w] = tot;
            if (done) __threadfence_system();
        }
    }
    if (done) {
        __syncthreads();
        if (threadIdx.x == 0) publish_done(done, seq);
    }
}

// ---------------------------------------------------------------------------
// Vnn for a list of beads, old and new positions in one pass.
// Block (c, b): bead b, partners [c*chunk, (c+1)*chunk).
// ---------------------------------------------------------------------------
// One moved bead of a Vnn request: its slice, old and new positions and cells.
struct VnnSubject {
    int m, me;
    double ox, oy, oz, nx, ny, nz;
    int c_old, c_new, s1;
    bool has_new, factors, me_on, use_cells;
    const double *px, *py, *pz;
};

__device__ __forceinline__ VnnSubject vnn_subject(const VnnArgs& A, const VnnBead& q) {
    VnnSubject v;
    v.m = q.slice;
    v.me = q.ptcl;
    v.px = A.pos + (int64_t)v.m * A.nalloc;
    v.py = v.px + A.plane;
    v.pz = v.py + A.plane;
    const bool has_old = q.flags & 1;
    v.has_new = q.flags & 2;
    v.ox = has_old ? q.old_pos[0] : v.px[v.me];
    v.oy = has_old ? q.old_pos[1] : v.py[v.me];
    v.oz = has_old ? q.old_pos[2] : v.pz[v.me];
    v.nx = q.new_pos[0];
    v.ny = q.new_pos[1];
    v.nz = q.new_pos[2];
    v.use_cells = A.box.ngrid >= 4;
    v.c_old = v.c_new = 0;
    if (v.use_cells) {
        v.c_old = has_old ? cell_of(v.ox, v.oy, v.oz, A.box) : A.cell[v.m * A.nalloc + v.me];
        v.c_new = cell_of(v.nx, v.ny, v.nz, A.box);
    }
    v.s1 = bead_state(A.worm, v.m, v.me);
    v.factors = worm_touches_slice(A.worm, v.m);
    v.me_on = A.on[v.m * A.nalloc + v.me] != 0;
    return v;
}

// Partner j's terms of Vnn(bead) at the old and at the new position
// (src/action.cpp:1044-1060 over the interaction list, src/lookuptable.cpp:464-516).
template <int MODE>
__device__ __forceinline__ void vnn_partner(const VnnArgs& A, const VnnSubject& v, int j, const AzizDev& az,
                                            double inv_dr, double& acc_o, double& acc_n) {
    const bool on = (A.on[v.m * A.nalloc + j] != 0) && (j != v.me) && v.me_on;
    const double xj = v.px[j], yj = v.py[j], zj = v.pz[j];
    double f = 1.0;
    if (v.factors) f = pair_factor(v.s1, bead_state(A.worm, v.m, j));
    int cj = 0;
    if (v.use_cells) cj = A.cell[v.m * A.nalloc + j];
    {
        const double r2 = sep_r2<MODE>(v.ox, v.oy, v.oz, xj, yj, zj, A.box);
        const bool ok = on && (r2 < A.box.rc2) && (!v.use_cells || cells_adjacent(v.c_old, cj, A.box.ngrid));
        const double val = pair_value_from_r2<MODE>(r2, az, inv_dr);
        acc_o += ok ? f * val : 0.0;
    }
    if (v.has_new) {
        const double r2 = sep_r2<MODE>(v.nx, v.ny, v.nz, xj, yj, zj, A.box);
        const bool ok = on && (r2 < A.box.rc2) && (!v.use_cells || cells_adjacent(v.c_new, cj, A.box.ngrid));
        const double val = pair_value_from_r2<MODE>(r2, az, inv_dr);
        acc_n += ok ? f * val : 0.0;
    }
}

__device__ __forceinline__ void vnn_finalize_body(const VnnArgs& A, const VnnInline& req);

// FUSED: the partials go out with an agent-scope release through the arrival
// counter, and the workgroup that arrives last (every other one has read its
// positions by then) acquires them and runs the finalisation and the commit,
// so a Delta V is one kernel.  No workgroup waits for another.
template <int MODE, bool FUSED>
__device__ __forceinline__ void vnn_body(const VnnArgs& A, const VnnInline& req, const AzizDev& az, int c, int b) {
    static_assert(kVnnThreads == 256, "four waves per Vnn workgroup");
    __shared__ double red[kVnnThreads / 64][2];
    const VnnBead q = A.beads ? A.beads[b] : req.beads[b];
    const VnnSubject v = vnn_subject(A, q);
    const double inv_dr = 1.0 / az.dr;
    double acc_o = 0.0, acc_n = 0.0;
    const int j_end = min((c + 1) * A.chunk, A.nalloc);
    for (int j = c * A.chunk + threadIdx.x; j < j_end; j += kVnnThreads) vnn_partner<MODE>(A, v, j, az, inv_dr, acc_o, acc_n);
    // both sums in one pass: wave trees, then the four waves in order
    acc_o = wave_sum(acc_o);
    acc_n = wave_sum(acc_n);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) {
        red[wave][0] = acc_o;
        red[wave][1] = acc_n;
    }
    __syncthreads();
    if (!FUSED) {
        if (threadIdx.x < 2) {
            const int r = threadIdx.x;
            A.partials[((int64_t)b * A.chunks + c) * 2 + r] = ((red[0][r] + red[1][r]) + red[2][r]) + red[3][r];
        }
        return;
    }
    __shared__ int last;
    if (threadIdx.x == 0) {
        for (int r = 0; r < 2; ++r)
            A.partials[((int64_t)b * A.chunks + c) * 2 + r] = ((red[0][r] + red[1][r]) + red[2][r]) + red[3][r];
        const unsigned total = (unsigned)(gridDim.x * gridDim.y);
        const unsigned t = __hip_atomic_fetch_add(A.ticket, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        last = (t + 1 == total);
    }
    __syncthreads();
    if (!last) return;
    if (threadIdx.x == 0) __hip_atomic_store(A.ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    vnn_finalize_body(A, req);
}
template <int MODE, bool FUSED>
__global__ __launch_bounds__(kVnnThreads) void k_vnn(VnnArgs A, VnnInline req, AzizDev az) {
    vnn_body<MODE, FUSED>(A, req, az, blockIdx.x, blockIdx.y);
}

// The committed bead updates of a request, one lane per bead (the beads sit on
// distinct slices): Path::updateBead with LookupTable::updateBead.
__device__ __forceinline__ void vnn_commit(const VnnArgs& A, const VnnInline& req, int nthreads) {
    for (int k = threadIdx.x; k < A.n_beads; k += nthreads) {
        const VnnBead qb = A.beads ? A.beads[k] : req.beads[k];
        if (!(qb.flags & 4)) continue;
        const int mm = qb.slice, ii = qb.ptcl;
        const int cn = cell_of(qb.new_pos[0], qb.new_pos[1], qb.new_pos[2], A.box);
        cell_update_bead(mm, ii, cn, A.nalloc, A.num_slices, A.cell, A.label, A.numlab, A.hash);
        double* qx = A.pos + (int64_t)mm * A.nalloc + ii;
        qx[0] = qb.new_pos[0];
        qx[A.plane] = qb.new_pos[1];
        qx[2 * A.plane] = qb.new_pos[2];
    }
}

// ---------------------------------------------------------------------------
// Gradient correction (GSF / Li-Broughton): gradVSquared per slice and the
// three-body gradVnnSquared of moved beads.
// ---------------------------------------------------------------------------
// |F_i|^2 with F_i = sum_{j != i} gradV(r_i - r_j) over the slice's n beads, j in order.
template <int MODE, typename P>
__device__ __forceinline__ double force2(P px, P py, P pz, int n, int i, const BoxDev& box, const AzizDev& a,
                                         double inv_dr) {
    const double xi = px[i], yi = py[i], zi = pz[i];
    double fx = 0.0, fy = 0.0, fz = 0.0;
    for (int j = 0; j < n; ++j) {
        if (j == i) continue;
        double dx, dy, dz;
        if (MODE == 0) {
            dx = min_image_fast(xi - px[j], box.L, box.invL);
            dy = min_image_fast(yi - py[j], box.L, box.invL);
            dz = min_image_fast(zi - pz[j], box.L, box.invL);
        } else {
            dx = min_image_exact(xi - px[j], box.L, box.invL);
            dy = min_image_exact(yi - py[j], box.L, box.invL);
            dz = min_image_exact(zi - pz[j], box.L, box.invL);
        }
        const double g = grad_factor_from_r2<MODE>(norm2_exact(dx, dy, dz), a, inv_dr);
        fx += g * dx;
        fy += g * dy;
        fz += g * dz;
    }
    return norm2_exact(fx, fy, fz);
}

// gradVSquared(slice) (src/action.cpp:1188-1223): block (b, m) takes rows
// 256 b .. 256 b + 255 of slice m; one partial per block, summed in block order.
template <int MODE>
__global__ __launch_bounds__(256) void k_gradv2(const double* __restrict__ pos, int64_t plane, int nalloc,
                                                const int* __restrict__ nbs, BoxDev box, AzizDev az, int lds_beads,
                                                double* __restrict__ part) {
    extern __shared__ double lds[];
    __shared__ double red[4];
    double* lx = lds;
    double* ly = lds + lds_beads;
    double* lz = lds + 2 * lds_beads;
    const int m = blockIdx.y, b = blockIdx.x;
    const int n = nbs[m];
    const double* px = pos + (int64_t)m * nalloc;
    const double* py = px + plane;
    const double* pz = py + plane;
    const bool in_lds = n <= lds_beads;
    if (in_lds)
        for (int t = threadIdx.x; t < n; t += 256) {
            lx[t] = px[t];
            ly[t] = py[t];
            lz[t] = pz[t];
        }
    __syncthreads();
    const double inv_dr = 1.0 / az.dr;
    const int i = b * 256 + threadIdx.x;
    double f2 = 0.0;
    if (i < n)
        f2 = in_lds ? force2<MODE>(lx, ly, lz, n, i, box, az, inv_dr) : force2<MODE>(px, py, pz, n, i, box, az, inv_dr);
    const double s = block_sum<256>(f2, red);
    if (threadIdx.x == 0) part[(int64_t)m * gridDim.x + b] = s;
}

// The same sum split over C partner chunks when the slices alone leave CUs idle
// (N = 256: one workgroup per slice): block (b C + c, m) takes rows 256 b ...
// of slice m against partners [c Nc, (c+1) Nc) and writes the rows' partial
// forces; k_gradv2_sum adds the C partials of each row in chunk order.
template <int MODE>
__global__ __launch_bounds__(256) void k_gradv2_chunk(const double* __restrict__ pos, int64_t plane, int nalloc,
                                                      const int* __restrict__ nbs, BoxDev box, AzizDev az,
                                                      int lds_beads, int C, double* __restrict__ fpart) {
    extern __shared__ double lds[];
    double* lx = lds;
    double* ly = lds + lds_beads;
    double* lz = lds + 2 * lds_beads;
    const int m = blockIdx.y, b = blockIdx.x / C, c = blockIdx.x - b * C;
    const int n = nbs[m];
    const double* px = pos + (int64_t)m * nalloc;
    const double* py = px + plane;
    const double* pz = py + plane;
    const bool in_lds = n <= lds_beads;
    if (in_lds)
        for (int t = threadIdx.x; t < n; t += 256) {
            lx[t] = px[t];
            ly[t] = py[t];
            lz[t] = pz[t];
        }
    __syncthreads();
    const double inv_dr = 1.0 / az.dr;
    const int i = b * 256 + threadIdx.x;
    const int nc = (n + C - 1) / C;
    const int j0 = c * nc, j1 = min(n, j0 + nc);
    if (i >= n) return;
    const double* qx = in_lds ? lx : px;
    const double* qy = in_lds ? ly : py;
    const double* qz = in_lds ? lz : pz;
    const double xi = qx[i], yi = qy[i], zi = qz[i];
    double fx = 0.0, fy = 0.0, fz = 0.0;
    for (int j = j0; j < j1; ++j) {
        if (j == i) continue;
        double dx, dy, dz;
        if (MODE == 0) {
            dx = min_image_fast(xi - qx[j], box.L, box.invL);
            dy = min_image_fast(yi - qy[j], box.L, box.invL);
            dz = min_image_fast(zi - qz[j], box.L, box.invL);
        } else {
            dx = min_image_exact(xi - qx[j], box.L, box.invL);
            dy = min_image_exact(yi - qy[j], box.L, box.invL);
            dz = min_image_exact(zi - qz[j], box.L, box.invL);
        }
        const double g = grad_factor_from_r2<MODE>(norm2_exact(dx, dy, dz), az, inv_dr);
        fx += g * dx;
        fy += g * dy;
        fz += g * dz;
    }
    double* o = fpart + (((int64_t)m * C + c) * nalloc + i) * 3;
    o[0] = fx;
    o[1] = fy;
    o[2] = fz;
}

// Mixed mode, slices of at most 256 beads: each unordered pair once.  Chunk c
// takes offsets d in [c Nd, (c+1) Nd) of 1..n/2; lane i adds g s_ij to its own
// force and -g s_ij (gradV(r_j - r_i)) to partner j's entry in its wave's LDS
// array (within one wave the partners i + d are distinct, and every wave has
// its own array, so no two lanes touch one entry at once); the row's partial
// force is its own sum plus the four waves' entries in wave order.  Half the
// gradV evaluations of k_gradv2_chunk for the same partial layout.
__global__ __launch_bounds__(256) void k_gradv2_sym(const double* __restrict__ pos, int64_t plane, int nalloc,
                                                    const int* __restrict__ nbs, BoxDev box, AzizDev az, int C,
                                                    double* __restrict__ fpart) {
    __shared__ double lp[3][256];
    __shared__ double fj[4][3][256];
    const int m = blockIdx.y, c = blockIdx.x;
    const int n = nbs[m];
    const int i = threadIdx.x, wave = threadIdx.x >> 6;
    const double* px = pos + (int64_t)m * nalloc;
    if (i < n) {
        lp[0][i] = px[i];
        lp[1][i] = px[plane + i];
        lp[2][i] = px[2 * plane + i];
    }
#pragma unroll
    for (int w = 0; w < 4; ++w)
#pragma unroll
        for (int a = 0; a < 3; ++a) fj[w][a][i] = 0.0;
    __syncthreads();
    const double inv_dr = 1.0 / az.dr;
    const int nd = n / 2;
    const int per = (nd + C - 1) / C;
    const int d0 = 1 + c * per, d1 = min(nd + 1, d0 + per);
    double fx = 0.0, fy = 0.0, fz = 0.0;
    if (i < n) {
        const double xi = lp[0][i], yi = lp[1][i], zi = lp[2][i];
        for (int d = d0; d < d1; ++d) {
            if ((n & 1) == 0 && d == nd && i >= nd) continue;   // the d = n/2 pairs once each
            int j = i + d;
            if (j >= n) j -= n;
            const double dx = min_image_fast(xi - lp[0][j], box.L, box.invL);
            const double dy = min_image_fast(yi - lp[1][j], box.L, box.invL);
            const double dz = min_image_fast(zi - lp[2][j], box.L, box.invL);
            const double g = grad_factor_from_r2<0>(norm2_exact(dx, dy, dz), az, inv_dr);
            const double gx = g * dx, gy = g * dy, gz = g * dz;
            fx += gx;
            fy += gy;
            fz += gz;
            fj[wave][0][j] -= gx;
            fj[wave][1][j] -= gy;
            fj[wave][2][j] -= gz;
        }
    }
    __syncthreads();
    if (i < n) {
        double* o = fpart + (((int64_t)m * C + c) * nalloc + i) * 3;
        o[0] = (((fx + fj[0][0][i]) + fj[1][0][i]) + fj[2][0][i]) + fj[3][0][i];
        o[1] = (((fy + fj[0][1][i]) + fj[1][1][i]) + fj[2][1][i]) + fj[3][1][i];
        o[2] = (((fz + fj[0][2][i]) + fj[1][2][i]) + fj[2][2][i]) + fj[3][2][i];
    }
}

__global__ __launch_bounds__(256) void k_gradv2_sum(const int* __restrict__ nbs, int nalloc, int C,
                                                    const double* __restrict__ fpart, double* __restrict__ part) {
    __shared__ double red[4];
    const int m = blockIdx.y, b = blockIdx.x;
    const int i = b * 256 + threadIdx.x;
    double f2 = 0.0;
    if (i < nbs[m]) {
        double fx = 0.0, fy = 0.0, fz = 0.0;
        for (int c = 0; c < C; ++c) {
            const double* o = fpart + (((int64_t)m * C + c) * nalloc + i) * 3;
            fx += o[0];
            fy += o[1];
            fz += o[2];
        }
        f2 = norm2_exact(fx, fy, fz);
    }
    const double s = block_sum<256>(f2, red);
    if (threadIdx.x == 0) part[(int64_t)m * gridDim.x + b] = s;
}

int gradv2_chunks(int nalloc, int num_slices, int num_cus) {
    const int B = (nalloc + 255) / 256;
    const int64_t blocks = (int64_t)B * num_slices;
    const int64_t want = 4LL * std::max(1, num_cus);
    if (blocks >= want) return 1;
    return (int)std::min<int64_t>(16, (want + blocks - 1) / blocks);
}

__device__ __forceinline__ double group16_sum(double v) {
    v += __shfl_xor(v, 8, 16);
    v += __shfl_xor(v, 4, 16);
    v += __shfl_xor(v, 2, 16);
    v += __shfl_xor(v, 1, 16);
    return v;
}

// gradVnnSquared(bead1) (src/action.cpp:1282-1337, external gradV = 0) at the
// old (y even) or new (y odd) position of request bead y/2.  Every workgroup
// builds the bead's interaction list (the Vnn partner set, in index order) in
// LDS; workgroup c takes list entries bead2 = 16c .. 16c+15, 16 lanes each over
// bead3, and writes {sum of |F2|^2 + 2 F2.F3, F1} with F1 = -sum F2.
template <int MODE>
__device__ __forceinline__ void gradvnn_body(const VnnArgs& A, const VnnInline& req, const AzizDev& az, int lds_beads,
                                             double* __restrict__ part, int c, int y) {
    extern __shared__ double lds[];   // x | y | z (lds_beads each), then the list (int[nalloc])
    __shared__ int wcount[4];
    __shared__ int list_n;
    __shared__ double red[4][4];
    const int bead = y >> 1, which = y & 1;
    const VnnBead q = A.beads ? A.beads[bead] : req.beads[bead];
    const VnnSubject v = vnn_subject(A, q);
    double* out = part + (((int64_t)bead * 2 + which) * A.gnn_chunks + c) * 4;
    if (which == 1 && !v.has_new) {
        if (threadIdx.x < 4) out[threadIdx.x] = 0.0;
        return;
    }
    const int N = A.nalloc;
    const bool in_lds = N <= lds_beads;
    double* lx = lds;
    double* ly = lds + lds_beads;
    double* lz = lds + 2 * lds_beads;
    int* list = reinterpret_cast<int*>(lds + 3 * lds_beads);
    const double* px = in_lds ? lx : v.px;
    const double* py = in_lds ? ly : v.py;
    const double* pz = in_lds ? lz : v.pz;
    if (in_lds)
        for (int t = threadIdx.x; t < N; t += 256) {
            lx[t] = v.px[t];
            ly[t] = v.py[t];
            lz[t] = v.pz[t];
        }
    const double x1 = which ? v.nx : v.ox, y1 = which ? v.ny : v.oy, z1 = which ? v.nz : v.oz;
    const int c1 = which ? v.c_new : v.c_old;
    // interaction list in partner order: on, not itself, r^2 < rc^2, neighbouring cell
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    int count = 0;
    for (int j0 = 0; j0 < N; j0 += 256) {
        const int j = j0 + threadIdx.x;
        bool keep = false;
        if (j < N) {
            const double r2 = sep_r2<MODE>(x1, y1, z1, v.px[j], v.py[j], v.pz[j], A.box);
            keep = (A.on[v.m * N + j] != 0) && (j != v.me) && v.me_on && (r2 < A.box.rc2) &&
                   (!v.use_cells || cells_adjacent(c1, A.cell[v.m * N + j], A.box.ngrid));
        }
        const unsigned long long ball = __ballot(keep);
        if (lane == 0) wcount[wave] = __popcll(ball);
        __syncthreads();
        int off = count;
        for (int w = 0; w < wave; ++w) off += wcount[w];
        if (keep) list[off + __popcll(ball & ((1ull << lane) - 1ull))] = j;
        count += wcount[0] + wcount[1] + wcount[2] + wcount[3];
        __syncthreads();
    }
    if (threadIdx.x == 0) list_n = count;
    __syncthreads();
    const int nl = list_n;
    const double inv_dr = 1.0 / az.dr;
    const int grp = threadIdx.x >> 4, gl = threadIdx.x & 15;
    const int n2 = c * kGnnGroup + grp;
    double f3x = 0.0, f3y = 0.0, f3z = 0.0;
    double x2 = 0.0, y2 = 0.0, z2 = 0.0;
    if (n2 < nl) {
        const int j2 = list[n2];
        x2 = px[j2];
        y2 = py[j2];
        z2 = pz[j2];
        for (int m3 = gl; m3 < nl; m3 += 16) {
            if (m3 == n2) continue;
            const int j3 = list[m3];
            // Path::getSeparation(bead2, bead3) = r2 - r3, in the bin mode's arithmetic
            double dx, dy, dz;
            if (MODE == 0) {
                dx = min_image_fast(x2 - px[j3], A.box.L, A.box.invL);
                dy = min_image_fast(y2 - py[j3], A.box.L, A.box.invL);
                dz = min_image_fast(z2 - pz[j3], A.box.L, A.box.invL);
            } else {
                dx = min_image_exact(x2 - px[j3], A.box.L, A.box.invL);
                dy = min_image_exact(y2 - py[j3], A.box.L, A.box.invL);
                dz = min_image_exact(z2 - pz[j3], A.box.L, A.box.invL);
            }
            const double g = grad_factor_from_r2<MODE>(norm2_exact(dx, dy, dz), az, inv_dr);
            f3x += g * dx;
            f3y += g * dy;
            f3z += g * dz;
        }
    }
    f3x = group16_sum(f3x);
    f3y = group16_sum(f3y);
    f3z = group16_sum(f3z);
    double term = 0.0, f1x = 0.0, f1y = 0.0, f1z = 0.0;
    if (n2 < nl && gl == 0) {
        // beadSep = Path::getSeparation(bead2, bead1) = r2 - r1
        double dx, dy, dz;
        if (MODE == 0) {
            dx = min_image_fast(x2 - x1, A.box.L, A.box.invL);
            dy = min_image_fast(y2 - y1, A.box.L, A.box.invL);
            dz = min_image_fast(z2 - z1, A.box.L, A.box.invL);
        } else {
            dx = min_image_exact(x2 - x1, A.box.L, A.box.invL);
            dy = min_image_exact(y2 - y1, A.box.L, A.box.invL);
            dz = min_image_exact(z2 - z1, A.box.L, A.box.invL);
        }
        const double g = grad_factor_from_r2<MODE>(norm2_exact(dx, dy, dz), az, inv_dr);
        const double f2x = g * dx, f2y = g * dy, f2z = g * dz;
        term = norm2_exact(f2x, f2y, f2z) + 2.0 * 0.0 + 2.0 * ((f2x * f3x + f2y * f3y) + f2z * f3z);
        f1x = -f2x;
        f1y = -f2y;
        f1z = -f2z;
    }
    term = wave_sum(term);
    f1x = wave_sum(f1x);
    f1y = wave_sum(f1y);
    f1z = wave_sum(f1z);
    if (lane == 0) {
        red[wave][0] = term;
        red[wave][1] = f1x;
        red[wave][2] = f1y;
        red[wave][3] = f1z;
    }
    __syncthreads();
    if (threadIdx.x < 4) out[threadIdx.x] = ((red[0][threadIdx.x] + red[1][threadIdx.x]) + red[2][threadIdx.x]) +
                                            red[3][threadIdx.x];
}
template <int MODE>
__global__ __launch_bounds__(256) void k_gradvnn(VnnArgs A, VnnInline req, AzizDev az, int lds_beads,
                                                 double* __restrict__ part) {
    gradvnn_body<MODE>(A, req, az, lds_beads, part, blockIdx.x, blockIdx.y);
}

// k_vnn's and k_gradvnn's workgroups in one launch (a request with the
// gradient correction): the first chunks x n_beads workgroups are k_vnn's
// (c, b), the rest k_gradvnn's (c, 2 bead + which).  The two sets read the same
// positions and write disjoint partials, so the values are those of the two
// launches bit for bit, with one launch and one inter-kernel wait fewer.
template <int MODE>
__global__ __launch_bounds__(256) void k_vnn_gnn(VnnArgs A, VnnInline req, AzizDev az, int lds_beads,
                                                 double* __restrict__ part) {
    const int n_vnn = A.chunks * A.n_beads;
    const int w = blockIdx.x;
    if (w < n_vnn) {
        vnn_body<MODE, false>(A, req, az, w % A.chunks, w / A.chunks);
    } else {
        const int t = w - n_vnn;
        gradvnn_body<MODE>(A, req, az, lds_beads, part, t % A.gnn_chunks, t / A.gnn_chunks);
    }
}

// Per-bead Vnn (chunks in order), the two actions and dU (a fixed shuffle tree
// over beads), then the committed bead updates, one lane per bead (the beads sit
// on distinct slices).
__device__ __forceinline__ void vnn_finalize_body(const VnnArgs& A, const VnnInline& req) {
    __shared__ double red[3][4];
    double ao = 0.0, an = 0.0, du = 0.0;
    const double tau_s = A.shift * A.tau;
    for (int k = threadIdx.x; k < A.n_beads; k += 256) {
        double vo = 0.0, vn = 0.0;
        for (int cc = 0; cc < A.chunks; ++cc) {
            vo += A.partials[((int64_t)k * A.chunks + cc) * 2 + 0];
            vn += A.partials[((int64_t)k * A.chunks + cc) * 2 + 1];
        }
        A.per_bead[2 * k + 0] = vo;
        A.per_bead[2 * k + 1] = vn;
        const int slice = A.beads ? A.beads[k].slice : req.beads[k].slice;
        const double co = ((slice & 1) ? A.vfactor1 : A.vfactor0) * tau_s;
        // potentialAction(bead) correction at shift 1 (src/action.cpp:497-503):
        // gradVnnSquared = sum of the chunks' three-body terms + |F1|^2
        double cor[2] = {0.0, 0.0};
        const double gvf = (slice & 1) ? A.gvf1 : A.gvf0;
        if (A.gnn_part && (A.shift == 1 || A.cor_any_shift) && gvf > kCorrEps) {
            // both positions' chunk terms, in chunk order per sum; four chunks'
            // loads are issued together before their additions
            const double* g0 = A.gnn_part + ((int64_t)k * 2 + 0) * A.gnn_chunks * 4;
            const double* g1 = g0 + (int64_t)A.gnn_chunks * 4;
            double t[2] = {0.0, 0.0}, fx[2] = {0.0, 0.0}, fy[2] = {0.0, 0.0}, fz[2] = {0.0, 0.0};
            int cc = 0;
            for (; cc + 4 <= A.gnn_chunks; cc += 4) {
                double v0[16], v1[16];
#pragma unroll
                for (int u = 0; u < 16; ++u) {
                    v0[u] = g0[4 * cc + u];
                    v1[u] = g1[4 * cc + u];
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    t[0] += v0[4 * u + 0];
                    fx[0] += v0[4 * u + 1];
                    fy[0] += v0[4 * u + 2];
                    fz[0] += v0[4 * u + 3];
                    t[1] += v1[4 * u + 0];
                    fx[1] += v1[4 * u + 1];
                    fy[1] += v1[4 * u + 2];
                    fz[1] += v1[4 * u + 3];
                }
            }
            for (; cc < A.gnn_chunks; ++cc)
                for (int w = 0; w < 2; ++w) {
                    const double* g = w ? g1 : g0;
                    t[w] += g[4 * cc + 0];
                    fx[w] += g[4 * cc + 1];
                    fy[w] += g[4 * cc + 2];
                    fz[w] += g[4 * cc + 3];
                }
            for (int w = 0; w < 2; ++w)
                cor[w] = gvf * tau_s * tau_s * tau_s * A.lambda * (t[w] + norm2_exact(fx[w], fy[w], fz[w]));
        }
        if (A.per_bead_cor) {
            A.per_bead_cor[2 * k + 0] = cor[0];
            A.per_bead_cor[2 * k + 1] = cor[1];
        }
        // the bead actions carry the correction at shift 1 only (src/action.cpp:500)
        if (A.shift != 1) cor[0] = cor[1] = 0.0;
        ao += co * vo + cor[0];
        an += co * vn + cor[1];
        du += co * (vn - vo) + (cor[1] - cor[0]);
    }
    ao = wave_sum(ao);
    an = wave_sum(an);
    du = wave_sum(du);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) {
        red[0][wave] = ao;
        red[1][wave] = an;
        red[2][wave] = du;
    }
    __syncthreads();
    if (threadIdx.x == 0)
        for (int r = 0; r < 3; ++r) A.out[r] = ((red[r][0] + red[r][1]) + red[r][2]) + red[r][3];
    if (A.commit_in_kernel) vnn_commit(A, req, 256);
    if (A.done) {
        __syncthreads();
        if (threadIdx.x == 0) publish_done(A.done, A.seq);
    }
}

__global__ __launch_bounds__(256) void k_vnn_finalize(VnnArgs A, VnnInline req) { vnn_finalize_body(A, req); }

// Delta V of a short request with no hand-off between workgroups: workgroup b
// sums bead b's Vnn over all its partners (the slice holds at most kVnnChunk
// beads) at the old position on its first 256 lanes and at the new position on
// the other 256, so each lane's chain is one pair term per partner; it commits
// bead b when asked (every other workgroup reads another slice: the beads sit on
// distinct slices) and publishes {Vnn_old, Vnn_new} with the call's sequence
// number straight into the caller's mapped block, slot b.  The caller adds the
// beads' weighted terms in bead order once every slot carries the sequence
// number.  A lane's partners, and so its sum, are those of the two-position
// kernels' lanes, and the waves are added in the same order.
template <int MODE>
__global__ __launch_bounds__(kVnnDirectThreads) void k_vnn_direct(VnnArgs A, VnnInline req, AzizDev az,
                                                                  VnnDirectOut* out) {
    __shared__ double red[kVnnDirectThreads / 64];
    const int b = blockIdx.x;
    const VnnBead q = req.beads[b];
    const VnnSubject v = vnn_subject(A, q);
    const double inv_dr = 1.0 / az.dr;
    const bool at_new = threadIdx.x >= kVnnThreads;
    const int j0 = threadIdx.x - (at_new ? kVnnThreads : 0);
    const double sx = at_new ? v.nx : v.ox, sy = at_new ? v.ny : v.oy, sz = at_new ? v.nz : v.oz;
    const int cs = at_new ? v.c_new : v.c_old;
    double acc = 0.0;
    for (int j = j0; j < A.nalloc; j += kVnnThreads) {
        const bool on = (A.on[v.m * A.nalloc + j] != 0) && (j != v.me) && v.me_on;
        const double xj = v.px[j], yj = v.py[j], zj = v.pz[j];
        double f = 1.0;
        if (v.factors) f = pair_factor(v.s1, bead_state(A.worm, v.m, j));
        int cj = 0;
        if (v.use_cells) cj = A.cell[v.m * A.nalloc + j];
        const double r2 = sep_r2<MODE>(sx, sy, sz, xj, yj, zj, A.box);
        const bool ok = on && (r2 < A.box.rc2) && (!v.use_cells || cells_adjacent(cs, cj, A.box.ngrid));
        const double val = pair_value_from_r2<MODE>(r2, az, inv_dr);
        acc += ok ? f * val : 0.0;
    }
    acc = wave_sum(acc);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) red[wave] = acc;
    __syncthreads();   // also: every lane's partner reads are done before the commit
    if (threadIdx.x == 0) {
        VnnDirectOut* o = out + b;
        o->v_old = ((red[0] + red[1]) + red[2]) + red[3];
        o->v_new = ((red[4] + red[5]) + red[6]) + red[7];
        if (q.flags & 4) {
            const int cn = cell_of(q.new_pos[0], q.new_pos[1], q.new_pos[2], A.box);
            cell_update_bead(q.slice, q.ptcl, cn, A.nalloc, A.num_slices, A.cell, A.label, A.numlab, A.hash);
            double* qx = A.pos + (int64_t)q.slice * A.nalloc + q.ptcl;
            qx[0] = q.new_pos[0];
            qx[A.plane] = q.new_pos[1];
            qx[2 * A.plane] = q.new_pos[2];
        }
        publish_done(&o->seq, A.seq);
    }
}

// ---------------------------------------------------------------------------
// Sequential bead updates (any slices, list order) -- Path::updateBead.
// ---------------------------------------------------------------------------
// distinct != 0: the beads sit on distinct slices (their cell tables are
// disjoint), one lane each; otherwise one lane in list order.
__global__ void k_update_beads(const int* __restrict__ sp, const double* __restrict__ p3, int n, double* pos,
                               int64_t plane, int nalloc, int num_slices, int* cell, int* label, int* numlab,
                               int* hash, BoxDev box, int distinct) {
    int k0 = 0, step = 1;
    if (distinct) {
        k0 = blockIdx.x * blockDim.x + threadIdx.x;
        step = gridDim.x * blockDim.x;
    } else if (blockIdx.x != 0 || threadIdx.x != 0) {
        return;
    }
    for (int k = k0; k < n; k += step) {
        const int m = sp[2 * k], i = sp[2 * k + 1];
        const int cn = cell_of(p3[3 * k], p3[3 * k + 1], p3[3 * k + 2], box);
        cell_update_bead(m, i, cn, nalloc, num_slices, cell, label, numlab, hash);
        double* q = pos + (int64_t)m * nalloc + i;
        q[0] = p3[3 * k];
        q[plane] = p3[3 * k + 1];
        q[2 * plane] = p3[3 * k + 2];
    }
}

// ---------------------------------------------------------------------------
// LookupTable::updateGrid (src/lookuptable.cpp:314-349): cells of every bead,
// then per slice a sequential label pass in particle order.
// ---------------------------------------------------------------------------
__global__ void k_cell_assign(const double* __restrict__ pos, int64_t plane, int total, int* __restrict__ cell,
                              const uint8_t* __restrict__ on, BoxDev box) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    cell[t] = on[t] ? cell_of(pos[t], pos[t + plane], pos[t + 2 * plane], box) : -1;
}

__global__ void k_cell_labels(const int* __restrict__ cell, int* __restrict__ label, int* __restrict__ numlab,
                              int* __restrict__ hash, int nalloc, int num_slices, int cells) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= num_slices) return;
    for (int c = 0; c < cells; ++c) numlab[c * num_slices + m] = 0;
    for (int i = 0; i < nalloc; ++i) {
        const int c = cell[m * nalloc + i];
        if (c < 0) {
            label[m * nalloc + i] = -1;
            continue;
        }
        const int nl = numlab[c * num_slices + m];
        hash[((int64_t)c * num_slices + m) * nalloc + nl] = i;
        label[m * nalloc + i] = nl;
        numlab[c * num_slices + m] = nl + 1;
    }
}

// The same labels with one wave per slice: the slice's beads in order, 64 at a
// time; within a group of 64, the lanes of each cell (one ballot per distinct
// cell) take consecutive labels after the cell's running count, kept in LDS.
// Label order, hash and counts equal the serial insertion order above.
constexpr int kCellLabelsLds = 8192;   // cells counted in LDS (beyond: the serial kernel)
__global__ __launch_bounds__(64) void k_cell_labels_wave(const int* __restrict__ cell, int* __restrict__ label,
                                                         int* __restrict__ numlab, int* __restrict__ hash,
                                                         int nalloc, int num_slices, int cells) {
    __shared__ int cnt[kCellLabelsLds];
    const int m = blockIdx.x, lane = threadIdx.x;
    for (int c = lane; c < cells; c += 64) cnt[c] = 0;
    __syncthreads();
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));   // lanes < this one
    for (int base = 0; base < nalloc; base += 64) {
        const int i = base + lane;
        const int c = (i < nalloc) ? cell[m * nalloc + i] : -1;
        uint64_t todo = __builtin_amdgcn_ballot_w64(c >= 0);
        int lab = -1;
        while (todo) {
            const int leader = __builtin_ctzll(todo);
            const int cl = __builtin_amdgcn_readlane(c, leader);
            const uint64_t grp = __builtin_amdgcn_ballot_w64(c == cl);
            const int start = cnt[cl];
            if (c == cl) lab = start + __builtin_popcountll(grp & below);
            __syncthreads();   // every lane has read cnt[cl] before lane `leader` advances it
            if (lane == leader) cnt[cl] = start + __builtin_popcountll(grp);
            __syncthreads();
            todo &= ~grp;
        }
        if (i < nalloc) {
            label[m * nalloc + i] = lab;
            if (c >= 0) hash[((int64_t)c * num_slices + m) * nalloc + lab] = i;
        }
    }
    for (int c = lane; c < cells; c += 64) numlab[c * num_slices + m] = cnt[c];
}

// LookupTable::updateInteractionList (src/lookuptable.cpp:464-516), one bead,
// walking the 27 de-duplicated neighbour cells in order.
__global__ void k_interaction_list(int m, int me, const double* __restrict__ pos, int64_t plane, int nalloc,
                                   int num_slices, const int* __restrict__ cell, const int* __restrict__ numlab,
                                   const int* __restrict__ hash, const int* __restrict__ grid_nn, BoxDev box,
                                   int* __restrict__ out_list, double* __restrict__ out_sep,
                                   int* __restrict__ out_count) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const double* px = pos + (int64_t)m * nalloc;
    const double* py = px + plane;
    const double* pz = py + plane;
    const int c1 = cell[m * nalloc + me];
    int count = 0;
    for (int nn = 0; nn < 27; ++nn) {
        const int cn = grid_nn[c1 * 27 + nn];
        if (cn < 0) continue;
        const int maxnl = numlab[cn * num_slices + m];
        for (int lab = 0; lab < maxnl; ++lab) {
            const int j = hash[((int64_t)cn * num_slices + m) * nalloc + lab];
            if (j == me) continue;
            const double dx = min_image_exact(px[j] - px[me], box.L, box.invL);
            const double dy = min_image_exact(py[j] - py[me], box.L, box.invL);
            const double dz = min_image_exact(pz[j] - pz[me], box.L, box.invL);
            if (norm2_exact(dx, dy, dz) < box.rc2) {
                out_list[2 * count] = m;
                out_list[2 * count + 1] = j;
                out_sep[3 * count] = dx;
                out_sep[3 * count + 1] = dy;
                out_sep[3 * count + 2] = dz;
                ++count;
            }
        }
    }
    *out_count = count;
}

// ---------------------------------------------------------------------------
// Separation histogram of one slice (src/action.cpp:216-224), integer bins.
// ---------------------------------------------------------------------------
__global__ void k_sep_hist(const double* __restrict__ pos, int64_t plane, int m, int n, int nalloc, BoxDev box,
                           double dsep, int* __restrict__ hist) {
    const double* px = pos + (int64_t)m * nalloc;
    const double* py = px + plane;
    const double* pz = py + plane;
    const int64_t total = (int64_t)n * n;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(t / n), j = (int)(t % n);
        if (j <= i) continue;
        const double dx = min_image_exact(px[j] - px[i], box.L, box.invL);
        const double dy = min_image_exact(py[j] - py[i], box.L, box.invL);
        const double dz = min_image_exact(pz[j] - pz[i], box.L, box.invL);
        const int nr = (int)(__builtin_sqrt(norm2_exact(dx, dy, dz)) / dsep);
        if (nr >= 0 && nr < 50) atomicAdd(&hist[nr], 1);
    }
}

// ---------------------------------------------------------------------------
// Kinetic link sum: sum over active beads of |putInBC(r(next) - r)|^2.
// next [M][nalloc][2]; a bead with next = XXX contributes 0.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_kinetic(const double* __restrict__ pos, int64_t plane, int nalloc,
                                                 int num_slices, const int* __restrict__ nbs,
                                                 const int* __restrict__ next, BoxDev box,
                                                 double* __restrict__ partials) {
    __shared__ double red[4];
    const int total = num_slices * nalloc;
    double acc = 0.0;
    for (int t = blockIdx.x * 256 + threadIdx.x; t < total; t += gridDim.x * 256) {
        const int m = t / nalloc, i = t % nalloc;
        if (i >= nbs[m]) continue;
        int nm, ni;
        if (next) {
            nm = next[2 * t];
            ni = next[2 * t + 1];
        } else {
            nm = (m + 1 == num_slices) ? 0 : m + 1;
            ni = i;
        }
        if (nm < 0 && ni < 0) continue;
        const int64_t a = (int64_t)m * nalloc + i, b = (int64_t)nm * nalloc + ni;
        const double dx = min_image_exact(pos[b] - pos[a], box.L, box.invL);
        const double dy = min_image_exact(pos[b + plane] - pos[a + plane], box.L, box.invL);
        const double dz = min_image_exact(pos[b + 2 * plane] - pos[a + 2 * plane], box.L, box.invL);
        acc += norm2_exact(dx, dy, dz);
    }
    const double s = block_sum<256>(acc, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// out[0] = sum of n partials in order
__global__ void k_sum_partials(const double* __restrict__ partials, int n, double* __restrict__ out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double tot = 0.0;
    for (int k = 0; k < n; ++k) tot += partials[k];
    out[0] = tot;
}

// ---------------------------------------------------------------------------
// Batched PotentialBase::V / gradV / grad2V on separations (table semantics).
// what: 0 V, 1 gradV, 2 grad2V.  Tables d1/d2 hold dV/dr and d2V/dr2.
// ---------------------------------------------------------------------------
template <int MODE>
__global__ void k_aziz_batch(const double* __restrict__ sep, int n, AzizDev az, const double* __restrict__ d1,
                             const double* __restrict__ d2, int what, double* __restrict__ out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const double x = sep[3 * t], y = sep[3 * t + 1], z = sep[3 * t + 2];
    const double r2 = norm2_exact(x, y, z);
    const double r = __builtin_sqrt(r2);
    const double q = r / az.dr;
    const bool valid = q < (double)az.table_length && (int)q > 0 && (int)q >= az.k_eps;
    const int k = valid ? (int)q : 0;
    if (what == 0) {
        double v = 0.0;
        if (valid) v = (MODE == 2) ? az.table[k] : (MODE == 1 ? aziz_v64(k, az) : aziz_vmix(k, az));
        out[t] = v;
    } else if (what == 1) {
        const double g = valid ? d1[k] / r : 0.0;
        out[3 * t] = valid ? g * x : 0.0;
        out[3 * t + 1] = valid ? g * y : 0.0;
        out[3 * t + 2] = valid ? g * z : 0.0;
    } else {
        out[t] = valid ? d2[k] : 0.0;
    }
}

// ---------------------------------------------------------------------------
// Host-side launchers (C++ linkage, used by pimc_capi.hip)
// ---------------------------------------------------------------------------
hipError_t launch_pair_action(int mode, hipStream_t st, const RowTile* tiles, int n_tiles, const int* ranges, int G,
                              const SliceDesc* slices, int n_slices, const WalkerDesc* walkers, int n_walkers,
                              const double* pos, const AzizDev& az, int lds_beads, double* partials, double* v_slice,
                              double* action, const KernelTimer* timer, bool finalize, unsigned long long* done,
                              unsigned long long seq, const double* g2_part, int g2_blocks, double* g2_slice,
                              const int* perm) {
    (void)hipGetLastError();  // drop any stale error of an earlier call
    if (n_tiles <= 0 || G <= 0) return hipSuccess;
    const bool spatial = perm != nullptr;
    dim3 g(G), b(kPairThreads);
    // LDS plane stride: the smallest of 256 / 512 / 1024 / 2048 beads that holds
    // the largest staged slice (a compile-time stride lets the three plane
    // reads share one address)
    const int PS = lds_beads <= 0 ? 0 : lds_beads <= 256 ? 256 : lds_beads <= 512 ? 512 : lds_beads <= 1024 ? 1024 : 2048;
    if (lds_beads > kLdsBeads) return hipErrorInvalidValue;
    auto go = [&](auto kern) { hipLaunchKernelGGL(kern, g, b, 0, st, tiles, ranges, walkers, pos, az, partials); };
    auto by_ps = [&](auto mode_tag) {
        constexpr int MODE = decltype(mode_tag)::value;
        if (MODE == 0 && spatial) {
            hipLaunchKernelGGL(k_pair_window, g, b, 0, st, tiles, ranges, walkers, pos, az, partials, perm);
            return;
        }
        if (PS == 0)
            go(k_pair_action<MODE, 0>);
        else if (PS == 256)
            go(k_pair_action<MODE, 256>);
        else if (PS == 512)
            go(k_pair_action<MODE, 512>);
        else if (PS == 1024)
            go(k_pair_action<MODE, 1024>);
        else
            go(k_pair_action<MODE, 2048>);
    };
    if (timer && timer->start) (void)hipEventRecord(timer->start, st);
    if (mode == 0)
        by_ps(std::integral_constant<int, 0>{});
    else if (mode == 1)
        by_ps(std::integral_constant<int, 1>{});
    else
        by_ps(std::integral_constant<int, 2>{});
    if (timer && timer->stop) (void)hipEventRecord(timer->stop, st);
    if (finalize)
        hipLaunchKernelGGL(k_pair_finalize, dim3(1), dim3(kFinThreads), 0, st, n_tiles, slices, n_slices, walkers,
                           n_walkers, partials, v_slice, action, done, seq, g2_part, g2_blocks, g2_slice);
    return hipGetLastError();
}

hipError_t launch_slice_order(hipStream_t st, const SliceDesc* slices, int n_slices, const double* pos, double invL,
                              int* perm) {
    (void)hipGetLastError();  // drop any stale error of an earlier call
    if (n_slices <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_slice_order, dim3(n_slices), dim3(kOrderThreads), 0, st, slices, pos, invL, perm);
    return hipGetLastError();
}

int pair_grid_for(int lds_beads, int num_cus) {
    // Sixteen workgroups per CU, more than LDS holds at once for large slices:
    // short per-workgroup ranges overlap one workgroup's slice staging with
    // another's pairs and shorten the tail (measured: N = 256 x 167 slices
    // 21.5 us against 26.2 us at four per CU, N = 1024 190 us against 233 us,
    // N = 2048 x 250 1.07 ms against 1.14 ms).
    (void)lds_beads;
    return 16 * std::max(1, num_cus);
}

int64_t append_row_tiles(int slice_index, const SliceDesc& sd, int pairs_per_lane, std::vector<RowTile>& tiles,
                         bool window) {
    const int n = sd.n;
    if (n < 2) return 0;
    auto tile = [&](int i0, int i_end, int d0, int dn) {
        tiles.push_back(RowTile{slice_index, i0, i_end, d0, dn, n, sd.slice, sd.walker, sd.pos, sd.plane});
    };
    const bool win = window && n >= kWinMinBeads;
    const int R = win ? kWinRows : n < kPairThreads ? n : kPairThreads;
    const int G = kPairThreads / R;
    int dn = std::max(1, pairs_per_lane) * G;    // offsets per tile
    if (win) dn = std::min(dn, kWinOffsets);
    if (!win && n < kPairThreads) {
        // one row block: offsets 1..n/2 in one run, d = n/2 of an even n masked to
        // rows i < n/2 inside the kernel (a tile of its own would mostly idle)
        const int dmax = n / 2;
        for (int d0 = 1; d0 <= dmax; d0 += dn) tile(0, n, d0, std::min(dn, dmax - d0 + 1));
        return (int64_t)n * (n - 1) / 2;
    }
    // row blocks of R against offsets 1..n/2; the d = n/2 pairs of an even n
    // count once: rows >= n/2 drop that offset, in the last iteration of the
    // block's last tile
    const int dmax = n / 2;
    for (int i0 = 0; i0 < n; i0 += R) {
        const int i_end = std::min(i0 + R, n);
        for (int d0 = 1; d0 <= dmax; d0 += dn) tile(i0, i_end, d0, std::min(dn, dmax - d0 + 1));
    }
    return (int64_t)n * (n - 1) / 2;
}

// A tile costs its pairs plus a fixed part (descriptor, slice staging, block
// sum): 128 walkers of N = 16 (120 pairs per tile) ran 30 k tiles in 73 us,
// against 6 us for their pairs at the N = 256 rate, about 1200 pair
// evaluations per tile.  Ranges balanced by this cost keep a workgroup of many
// small slices from finishing last in a batch of uneven walkers; over 1024 /
// 2048 / 4096 the 512-walker batch took 2.00 / 1.85 / 2.00 ms (3.03 ms when
// balanced by pairs alone).
constexpr int64_t kTileCost = 2048;

void balance_row_tiles(const std::vector<RowTile>& tiles, int G, std::vector<int>& ranges) {
    const int T = (int)tiles.size();
    ranges.assign(G + 1, T);
    ranges[0] = 0;
    int64_t tot = 0;
    for (const auto& t : tiles) tot += row_tile_pairs(t) + kTileCost;
    int64_t acc = 0;
    int g = 1;
    for (int k = 0; k < T && g < G; ++k) {
        acc += row_tile_pairs(tiles[k]) + kTileCost;
        while (g < G && acc * G >= tot * g) ranges[g++] = k + 1;
    }
}

hipError_t launch_vnn_direct(int mode, hipStream_t st, const VnnArgs& A, const VnnInline& req, const AzizDev& az,
                             VnnDirectOut* out, const KernelTimer* timer) {
    (void)hipGetLastError();  // drop any stale error of an earlier call
    if (A.n_beads <= 0) return hipSuccess;
    if (A.n_beads > kVnnInline || A.nalloc > kVnnChunk || A.beads) return hipErrorInvalidValue;
    dim3 g(A.n_beads), b(kVnnDirectThreads);
    if (timer && timer->start) (void)hipEventRecord(timer->start, st);
    if (mode == 0)
        hipLaunchKernelGGL(k_vnn_direct<0>, g, b, 0, st, A, req, az, out);
    else if (mode == 1)
        hipLaunchKernelGGL(k_vnn_direct<1>, g, b, 0, st, A, req, az, out);
    else
        hipLaunchKernelGGL(k_vnn_direct<2>, g, b, 0, st, A, req, az, out);
    if (timer && timer->stop) (void)hipEventRecord(timer->stop, st);
    return hipGetLastError();
}

const void* vnn_direct_stub(int mode) {
    if (mode == 0) return reinterpret_cast<const void*>(&k_vnn_direct<0>);
    if (mode == 1) return reinterpret_cast<const void*>(&k_vnn_direct<1>);
    return reinterpret_cast<const void*>(&k_vnn_direct<2>);
}

hipError_t launch_vnn(int mode, hipStream_t st, const VnnArgs& A, const VnnInline& req, const AzizDev& az,
                      const KernelTimer* timer, double* gnn_part, int lds_beads, const KernelTimer* gnn_timer) {
    (void)hipGetLastError();  // drop any stale error of an earlier call
    if (A.n_beads <= 0) return hipSuccess;
    dim3 g(A.chunks, A.n_beads), b(kVnnThreads);
    auto go = [&](auto kern) { hipLaunchKernelGGL(kern, g, b, 0, st, A, req, az); };
    const bool fused = A.ticket && !gnn_part;
    if (timer && timer->start) (void)hipEventRecord(timer->start, st);
    if (fused) {
        if (mode == 0)
            go(k_vnn<0, true>);
        else if (mode == 1)
            go(k_vnn<1, true>);
        else
            go(k_vnn<2, true>);
        if (timer && timer->stop) (void)hipEventRecord(timer->stop, st);
        return hipGetLastError();
    }
    const bool timed = (timer && timer->start) || (gnn_timer && gnn_timer->start);
    if (gnn_part && !timed) {
        // one launch for both workgroup sets (a timed call keeps two, so each
        // kernel has its own duration)
        const size_t lds = sizeof(double) * 3 * (size_t)lds_beads + sizeof(int) * (size_t)A.nalloc;
        dim3 gg(A.chunks * A.n_beads + A.gnn_chunks * 2 * A.n_beads);
        if (mode == 0)
            hipLaunchKernelGGL(k_vnn_gnn<0>, gg, dim3(256), lds, st, A, req, az, lds_beads, gnn_part);
        else if (mode == 1)
            hipLaunchKernelGGL(k_vnn_gnn<1>, gg, dim3(256), lds, st, A, req, az, lds_beads, gnn_part);
        else
            hipLaunchKernelGGL(k_vnn_gnn<2>, gg, dim3(256), lds, st, A, req, az, lds_beads, gnn_part);
        hipLaunchKernelGGL(k_vnn_finalize, dim3(1), dim3(256), 0, st, A, req);
        return hipGetLastError();
    }
    if (mode == 0)
        go(k_vnn<0, false>);
    else if (mode == 1)
        go(k_vnn<1, false>);
    else
        go(k_vnn<2, false>);
    if (timer && timer->stop) (void)hipEventRecord(timer->stop, st);
    if (gnn_part) {
        VnnArgs B = A;
        const size_t lds = sizeof(double) * 3 * (size_t)lds_beads + sizeof(int) * (size_t)A.nalloc;
        dim3 gg(A.gnn_chunks, 2 * A.n_beads);
        if (gnn_timer && gnn_timer->start) (void)hipEventRecord(gnn_timer->start, st);
        if (mode == 0)
            hipLaunchKernelGGL(k_gradvnn<0>, gg, dim3(256), lds, st, B, req, az, lds_beads, gnn_part);
        else if (mode == 1)
            hipLaunchKernelGGL(k_gradvnn<1>, gg, dim3(256), lds, st, B, req, az, lds_beads, gnn_part);
        else
            hipLaunchKernelGGL(k_gradvnn<2>, gg, dim3(256), lds, st, B, req, az, lds_beads, gnn_part);
        if (gnn_timer && gnn_timer->stop) (void)hipEventRecord(gnn_timer->stop, st);
    }
    hipLaunchKernelGGL(k_vnn_finalize, dim3(1), dim3(256), 0, st, A, req);
    return hipGetLastError();
}

hipError_t launch_gradv2(int mode, hipStream_t st, const double* pos, int64_t plane, int nalloc, int num_slices,
                         const int* nbs, const BoxDev& box, const AzizDev& az, int lds_beads, double* part,
                         int* blocks_per_slice, int chunks, double* fpart) {
    (void)hipGetLastError();
    const int B = (nalloc + 255) / 256;
    *blocks_per_slice = B;
    const size_t lds = sizeof(double) * 3 * (size_t)lds_beads;
    if (chunks > 1 && fpart && mode == 0 && nalloc <= 256) {
        hipLaunchKernelGGL(k_gradv2_sym, dim3(chunks, num_slices), dim3(256), 0, st, pos, plane, nalloc, nbs, box, az,
                           chunks, fpart);
        hipLaunchKernelGGL(k_gradv2_sum, dim3(B, num_slices), dim3(256), 0, st, nbs, nalloc, chunks, fpart, part);
        return hipGetLastError();
    }
    if (chunks > 1 && fpart) {
        dim3 gc(B * chunks, num_slices), b(256);
        if (mode == 0)
            hipLaunchKernelGGL(k_gradv2_chunk<0>, gc, b, lds, st, pos, plane, nalloc, nbs, box, az, lds_beads, chunks,
                               fpart);
        else if (mode == 1)
            hipLaunchKernelGGL(k_gradv2_chunk<1>, gc, b, lds, st, pos, plane, nalloc, nbs, box, az, lds_beads, chunks,
                               fpart);
        else
            hipLaunchKernelGGL(k_gradv2_chunk<2>, gc, b, lds, st, pos, plane, nalloc, nbs, box, az, lds_beads, chunks,
                               fpart);
        hipLaunchKernelGGL(k_gradv2_sum, dim3(B, num_slices), b, 0, st, nbs, nalloc, chunks, fpart, part);
        return hipGetLastError();
    }
    dim3 g(B, num_slices), b(256);
    if (mode == 0)
        hipLaunchKernelGGL(k_gradv2<0>, g, b, lds, st, pos, plane, nalloc, nbs, box, az, lds_beads, part);
    else if (mode == 1)
        hipLaunchKernelGGL(k_gradv2<1>, g, b, lds, st, pos, plane, nalloc, nbs, box, az, lds_beads, part);
    else
        hipLaunchKernelGGL(k_gradv2<2>, g, b, lds, st, pos, plane, nalloc, nbs, box, az, lds_beads, part);
    return hipGetLastError();
}

hipError_t launch_update_beads(hipStream_t st, const int* sp, const double* p3, int n, double* pos, int64_t plane,
                               int nalloc, int num_slices, int* cell, int* label, int* numlab, int* hash,
                               const BoxDev& box, bool distinct) {
    (void)hipGetLastError();  // drop any stale error of an earlier call
    const int grid = distinct ? (n + 255) / 256 : 1;
    hipLaunchKernelGGL(k_update_beads, dim3(grid), dim3(distinct ? 256 : 64), 0, st, sp, p3, n, pos, plane, nalloc,
                       num_slices, cell, label, numlab, hash, box, distinct ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_cell_build(hipStream_t st, const double* pos, int64_t plane, int nalloc, int num_slices,
                             const uint8_t* on, int* cell, int* label, int* numlab, int* hash, int cells,
                             const BoxDev& box) {
    (void)hipGetLastError();  // drop any stale error of an earlier call
    const int total = nalloc * num_slices;
    hipLaunchKernelGGL(k_cell_assign, dim3((total + 255) / 256), dim3(256), 0, st, pos, plane, total, cell, on, box);
    if (cells <= kCellLabelsLds)
        hipLaunchKernelGGL(k_cell_labels_wave, dim3(num_slices), dim3(64), 0, st, cell, label, numlab, hash, nalloc,
                           num_slices, cells);
    else
        hipLaunchKernelGGL(k_cell_labels, dim3((num_slices + 63) / 64), dim3(64), 0, st, cell, label, numlab, hash,
                           nalloc, num_slices, cells);
    return hipGetLastError();
}

hipError_t launch_interaction_list(hipStream_t st, int m, int me, const double* pos, int64_t plane, int nalloc,
                                   int num_slices, const int* cell, const int* numlab, const int* hash,
                                   const int* grid_nn, const BoxDev& box, int* out_list, double* out_sep,
                                   int* out_count) {
    (void)hipGetLastError();  // drop any stale error of an earlier call
    hipLaunchKernelGGL(k_interaction_list, dim3(1), dim3(64), 0, st, m, me, pos, plane, nalloc, num_slices, cell,
                       numlab, hash, grid_nn, box, out_list, out_sep, out_count);
    return hipGetLastError();
}

hipError_t launch_sep_hist(hipStream_t st, const double* pos, int64_t plane, int m, int n, int nalloc,
                           const BoxDev& box, double dsep, int* hist) {
    (void)hipGetLastError();  // drop any stale error of an earlier call
    int64_t total = (int64_t)n * n;
    int blocks = (int)((total + 255) / 256);
    if (blocks > 4096) blocks = 4096;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_sep_hist, dim3(blocks), dim3(256), 0, st, pos, plane, m, n, nalloc, box, dsep, hist);
    return hipGetLastError();
}

hipError_t launch_kinetic(hipStream_t st, const double* pos, int64_t plane, int nalloc, int num_slices,
                          const int* nbs, const int* next, const BoxDev& box, double* partials, int max_blocks,
                          double* out) {
    (void)hipGetLastError();  // drop any stale error of an earlier call
    int total = nalloc * num_slices;
    int blocks = (total + 255) / 256;
    if (blocks > max_blocks) blocks = max_blocks;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_kinetic, dim3(blocks), dim3(256), 0, st, pos, plane, nalloc, num_slices, nbs, next, box,
                       partials);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(64), 0, st, partials, blocks, out);
    return hipGetLastError();
}

hipError_t launch_aziz_batch(int mode, hipStream_t st, const double* sep, int n, const AzizDev& az, const double* d1,
                             const double* d2, int what, double* out) {
    (void)hipGetLastError();  // drop any stale error of an earlier call
    if (n <= 0) return hipSuccess;
    dim3 g((n + 255) / 256), b(256);
    if (mode == 0)
        hipLaunchKernelGGL(k_aziz_batch<0>, g, b, 0, st, sep, n, az, d1, d2, what, out);
    else if (mode == 1)
        hipLaunchKernelGGL(k_aziz_batch<1>, g, b, 0, st, sep, n, az, d1, d2, what, out);
    else
        hipLaunchKernelGGL(k_aziz_batch<2>, g, b, 0, st, sep, n, az, d1, d2, what, out);
    return hipGetLastError();
}

}  // namespace pimc
