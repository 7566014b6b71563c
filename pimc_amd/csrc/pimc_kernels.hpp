// pimc_kernels.hpp -- kernel-side data descriptors shared by the host code and
// the kernels (pair action, Delta V, cell lists, structure factors).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "pimc_device.hpp"

namespace pimc {

constexpr int kPairThreads = 256;   // workgroup of the all-pairs kernel
constexpr int kVnnThreads = 256;    // workgroup of the Vnn (Delta V) kernel
constexpr int kVnnDirectThreads = 2 * kVnnThreads;   // k_vnn_direct: old and new position
constexpr int kLdsBeads = 2048;     // slice positions staged in LDS up to this size

// One imaginary-time slice of one walker.
struct SliceDesc {
    int n;          // active beads on the slice (numBeadsAtSlice, left packed)
    int slice;      // slice index m within its walker
    int walker;     // walker index
    int first_item; // first work item of this slice (items are grouped by slice)
    int64_t pos;    // offset of x of bead 0 of this slice in the SoA position array
    int64_t plane;  // distance between the x, y and z planes of this walker
};

// One walker (chain): box, table length and action weights.
struct WalkerDesc {
    BoxDev box;
    int table_length;
    int num_slices;
    int first_slice;   // index of its slice 0 in the SliceDesc array
    int shift;
    double vfactor0, vfactor1, tau;
    double gvf0, gvf1, lambda;   // gradient correction gradVFactor[2] and lambda (0: none)
    WormDev worm;
};

// A gradVFactor counts as finite above EPS (src/action.cpp:466, :500).
constexpr double kCorrEps = 1.0e-7;

// Every i<j pair of a slice with n beads is visited once as (i, j = (i + d) mod n)
// with offsets 1 <= d <= (n-1)/2 for all rows i, plus, for even n, d = n/2 for
// rows i < n/2.  A tile is rows [i0, i_end) against offsets [d0, d0 + dn).  A
// workgroup's 256 lanes hold one row each (n >= 256) or form 256/n groups of n
// rows (n < 256), group g taking offsets d0 + g, d0 + g + 256/n, ...; a lane's
// row bead stays in registers while its partner index steps through the slice.
// The d = n/2 pairs of an even n stay in the tile holding that offset (always
// the last iteration of its lane group), where rows >= n/2 drop them.
// A tile carries its slice's fields, so a workgroup reaches the positions and
// the walker in one dependent load after the tile's.
struct RowTile {
    int slice;      // SliceDesc index (the partial's slice)
    int i0, i_end;  // rows
    int d0, dn;     // partner offsets
    int n;          // active beads on the slice
    int m;          // slice index within its walker
    int walker;     // WalkerDesc index
    int64_t pos;    // offset of x of bead 0 in the SoA position array
    int64_t plane;  // distance between the x, y and z planes
};

// Appends the tiles of slice `slice_index` (described by sd) with about
// pairs_per_lane pairs per lane each; returns the slice's pair count n(n-1)/2.
// window: the tiles of the window form (k_pair_window: 32 rows x at most 1024
// offsets on slices of 64 beads and more).
int64_t append_row_tiles(int slice_index, const SliceDesc& sd, int pairs_per_lane, std::vector<RowTile>& tiles,
                         bool window = false);
// Splits tiles [0, T) into G contiguous ranges of about equal pair count:
// workgroup g of the persistent launch takes tiles [ranges[g], ranges[g+1]).
void balance_row_tiles(const std::vector<RowTile>& tiles, int G, std::vector<int>& ranges);
// Pair evaluations of a tile (masked d = n/2 lanes included: the balancing weight).
inline int64_t row_tile_pairs(const RowTile& t) { return (int64_t)(t.i_end - t.i0) * t.dn; }

// Virial pair terms of a closed path (pimc_virial.hip): per-bead terms
// F.r, (T F).r, F.delta, (T F).delta, |F|^2, -vel1.vel2, -vel2.vel2.
constexpr int kVirialTerms = 7;
struct VirialArgs {
    const double* pos;
    int64_t plane;
    int nalloc, num_slices;
    const int* nbs;      // numBeadsAtSlice [M]
    const int* next;     // next links [M][nalloc][2], or null: (m+1 mod M, i)
    const int* prev;     // prev links [M][nalloc][2], or null: (m-1 mod M, i)
    int window;          // virialWindow >= 1
    int t_even, t_odd;   // evaluate the T-matrix on even / odd slices (finite gradVFactor)
    int lds_beads;       // slices with at most this many beads are staged in LDS
    BoxDev box;
    double* part;        // [M][blocks][kVirialTerms]
    int chunks;          // partner chunks (1: one pass per row block)
    double* fpart;       // chunk partials [M][chunks][nalloc][9] = F, T (chunks > 1)
};

// A Vnn evaluation request for one bead: partners are the active beads of its slice.
struct VnnBead {
    int walker;
    int slice;      // slice index m
    int ptcl;
    int flags;      // bit0: has old position, bit1: has new position, bit2: commit
    double old_pos[3];
    double new_pos[3];
};

constexpr int kVnnChunk = 1024;     // partners per Vnn workgroup

struct VnnArgs {
    const VnnBead* beads;  // device list, or nullptr when the beads travel inline
    int n_beads;
    int chunks;        // blocks per bead
    int chunk;         // partners per block
    double* pos;       // SoA planes of the walker
    int64_t plane;
    int nalloc, num_slices;
    const uint8_t* on; // [M][nalloc]
    int* cell;         // [M][nalloc] packed cell numbers
    int* label;
    int* numlab;
    int* hash;
    int commit_in_kernel;
    BoxDev box;
    WormDev worm;
    int shift;
    double vfactor0, vfactor1, tau;
    double* partials;  // [n_beads][chunks][2]
    double* per_bead;  // [n_beads][2]
    double* out;       // {A_old, A_new, dU} (host-mapped pinned memory)
    unsigned long long* done;   // host-mapped: set to seq once out is written (or null)
    unsigned long long seq;
    // gradient correction: gradVnnSquared partials [n_beads][2][gnn_chunks][4] (null: none)
    const double* gnn_part;
    int gnn_chunks;
    double gvf0, gvf1, lambda;
    int cor_any_shift;     // evaluate the correction at any shift (potentialActionCorrection)
    double* per_bead_cor;  // [n_beads][2] corrections at the old / new position, or null
    // arrival counter of the one-kernel form (zero between calls), or null: the
    // last workgroup of k_vnn to finish runs the finalisation itself
    unsigned* ticket;
};

constexpr int kGnnGroup = 16;   // bead2 per three-body workgroup (16 lanes each)

// Requests of up to kVnnInline beads travel in the kernel arguments.
constexpr int kVnnInline = 32;
struct VnnInline {
    VnnBead beads[kVnnInline];
};

// One moved bead's result of the direct Delta V form, written by its workgroup
// into the caller's mapped block: Vnn at the old and the new position, then the
// call's sequence number once both are visible.
struct VnnDirectOut {
    double v_old, v_new;
    unsigned long long seq;
    unsigned long long pad;
};

struct RhoArgs {
    const double* pos;     // SoA planes [3][M][nalloc]
    int64_t plane;
    int nalloc, num_slices;
    const int* nbs;        // beads per slice
    double invL;
    int nq;
    const int* qn;         // [nq][3] integer lattice components (lattice path)
    const double* q;       // [nq][3] (general path)
    int nmax[3];
    double* rho;           // [M][nq][2] FP64
};

// Publishes seq to a host-mapped completion word after the calling thread's
// earlier writes (the host spins on it instead of a stream synchronisation).
__device__ __forceinline__ void publish_done(unsigned long long* done, unsigned long long seq) {
    __threadfence_system();
    __hip_atomic_store(done, seq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Start/stop events recorded on the stream right before and after one kernel's
// launch, so their elapsed time is that kernel's duration; null: not timed.
struct KernelTimer {
    hipEvent_t start = nullptr, stop = nullptr;
};

// Workgroup b of a launch is dispatched to XCD b % 8.  Work item of workgroup b
// such that each XCD walks one contiguous run of the G items (items are grouped
// by slice, so a slice's beads are fetched into one XCD's L2 instead of eight).
__device__ __forceinline__ int xcd_contiguous_item(int b, int G) {
    const int x = b & 7, local = b >> 3;
    const int per = G >> 3, rem = G & 7;
    return x < rem ? x * (per + 1) + local : rem * (per + 1) + (x - rem) * per + local;
}

// ---- launchers (defined in pimc_pair.hip / pimc_sq.hip) ----
// The persistent row-tile pair kernel over G workgroups, then (finalize != 0)
// the single-workgroup slice and walker sums.
// done (host-mapped, or null) is set to seq after action has been written.
// g2_part (or null): gradVSquared partials [n_slices][g2_blocks] whose slice sums
// (written to g2_slice) enter the walkers' actions with the gradient correction.
// perm (or null): mixed mode, the window form (k_pair_window, tiles from
// append_row_tiles(..., window = true)) over slices in the spatial order
// launch_slice_order wrote (slice s at perm + slices[s].pos).
hipError_t launch_pair_action(int mode, hipStream_t st, const RowTile* tiles, int n_tiles, const int* ranges, int G,
                              const SliceDesc* slices, int n_slices, const WalkerDesc* walkers, int n_walkers,
                              const double* pos, const AzizDev& az, int lds_beads, double* partials, double* v_slice,
                              double* action, const KernelTimer* timer = nullptr, bool finalize = true,
                              unsigned long long* done = nullptr, unsigned long long seq = 0,
                              const double* g2_part = nullptr, int g2_blocks = 0, double* g2_slice = nullptr,
                              const int* perm = nullptr);
// Spatial order of every slice (slices of at most kLdsBeads beads, one box):
// perm[slices[s].pos + t], t < slices[s].n, lists slice s's beads by the index of
// their cell along a Hilbert curve through a 32^3 grid over the box, ties by
// bead index (perm spans the x plane of the position array).
hipError_t launch_slice_order(hipStream_t st, const SliceDesc* slices, int n_slices, const double* pos, double invL,
                              int* perm);
// Workgroups of the persistent pair kernel for a slice staging of lds_beads beads.
int pair_grid_for(int lds_beads, int num_cus);
hipError_t launch_vnn(int mode, hipStream_t st, const VnnArgs& A, const VnnInline& req, const AzizDev& az,
                      const KernelTimer* timer = nullptr, double* gnn_part = nullptr, int lds_beads = 0,
                      const KernelTimer* gnn_timer = nullptr);
// The direct form (k_vnn_direct): at most kVnnInline beads travelling inline, at
// most kVnnChunk beads per slice; out [n_beads] is host-mapped.
hipError_t launch_vnn_direct(int mode, hipStream_t st, const VnnArgs& A, const VnnInline& req, const AzizDev& az,
                             VnnDirectOut* out, const KernelTimer* timer = nullptr);
// k_vnn_direct<mode>'s host stub, and its argument block as the AMDGPU ABI lays
// it out (each by-value argument at its alignment), for a dispatch packet
const void* vnn_direct_stub(int mode);
constexpr uint32_t karg_align(uint32_t o, uint32_t a) { return (o + a - 1) / a * a; }
constexpr uint32_t kVdirOffReq = karg_align(sizeof(VnnArgs), alignof(VnnInline));
constexpr uint32_t kVdirOffAz = karg_align(kVdirOffReq + sizeof(VnnInline), alignof(AzizDev));
constexpr uint32_t kVdirOffOut = karg_align(kVdirOffAz + sizeof(AzizDev), alignof(VnnDirectOut*));
constexpr uint32_t kVdirBytes = kVdirOffOut + sizeof(VnnDirectOut*);
// gradVSquared(slice) partials [M][blocks_per_slice] for every slice.
hipError_t launch_gradv2(int mode, hipStream_t st, const double* pos, int64_t plane, int nalloc, int num_slices,
                         const int* nbs, const BoxDev& box, const AzizDev& az, int lds_beads, double* part,
                         int* blocks_per_slice, int chunks = 1, double* fpart = nullptr);
// Partner chunks of the gradVSquared sum that fill the CUs (1: one pass).
int gradv2_chunks(int nalloc, int num_slices, int num_cus);
// Device-resident chain (pimc_dchain.hip): one workgroup runs n_moves moves.
constexpr int kChainThreads = 768;      // 12 waves: three per SIMD
constexpr int kChainThreadsCor = 512;   // with the gradient correction: two per SIMD
constexpr int kChainMaxBeads = 512;   // beads one move may list (M for CenterOfMass)
constexpr int kChainMaxWind = 343;    // winding sectors, (2 max_wind + 1)^3 with max_wind <= 3
constexpr int kChainMaxList = 8192;   // beads per slice the correction's list mask covers
struct ChainRunArgs {
    double* pos;
    int64_t plane;
    int nalloc, num_slices;
    int *cell, *label, *numlab, *hash;
    BoxDev box;
    double side, tau, lambda, vf0, vf1;
    double gvf0, gvf1;   // gradVFactor of the GSF / Li-Broughton correction (0: none)
    double sqrtLambdaTau, sqrt2LambdaTau, com_delta;
    int Mbar, levels, var_updates, num_wind;
    const int* wind;     // [num_wind][3]
    uint32_t* mt;        // MTRand state [624], left, next (read and written back)
    const int* kinds;    // [n_moves]: 0 staging, 1 bisection, 2 centre of mass
    int n_moves;
    int* accepted;       // [n_moves]
    double* info;        // [n_moves][3]: attempted, last Delta, length
    int* near;           // decisions inside the guard band
};
// args / azs: device arrays, one entry per chain (workgroup).
hipError_t launch_chain_run(int mode, bool cor, hipStream_t st, int n_chains, const ChainRunArgs* args,
                            const AzizDev* azs);
hipError_t launch_virial(int mode, hipStream_t st, const VirialArgs& A, const AzizDev& az, int* blocks_per_slice);
hipError_t launch_update_beads(hipStream_t st, const int* sp, const double* p3, int n, double* pos, int64_t plane,
                               int nalloc, int num_slices, int* cell, int* label, int* numlab, int* hash,
                               const BoxDev& box, bool distinct);
hipError_t launch_cell_build(hipStream_t st, const double* pos, int64_t plane, int nalloc, int num_slices,
                             const uint8_t* on, int* cell, int* label, int* numlab, int* hash, int cells,
                             const BoxDev& box);
hipError_t launch_interaction_list(hipStream_t st, int m, int me, const double* pos, int64_t plane, int nalloc,
                                   int num_slices, const int* cell, const int* numlab, const int* hash,
                                   const int* grid_nn, const BoxDev& box, int* out_list, double* out_sep,
                                   int* out_count);
hipError_t launch_sep_hist(hipStream_t st, const double* pos, int64_t plane, int m, int n, int nalloc,
                           const BoxDev& box, double dsep, int* hist);
hipError_t launch_kinetic(hipStream_t st, const double* pos, int64_t plane, int nalloc, int num_slices,
                          const int* nbs, const int* next, const BoxDev& box, double* partials, int max_blocks,
                          double* out);
hipError_t launch_aziz_batch(int mode, hipStream_t st, const double* sep, int n, const AzizDev& az, const double* d1,
                             const double* d2, int what, double* out);
hipError_t launch_rho(bool lattice, hipStream_t st, const RhoArgs& A);
hipError_t launch_isf(hipStream_t st, const double* rho, int nq, int M, double inv_n, double* fqt, int only_tau0);
hipError_t launch_ssf_pairs(hipStream_t st, const double* pos, int64_t plane, int nalloc, int M, const int* nbs,
                            const double* q, int nq, const BoxDev& box, double* partial, double inv_n, double* sq);
int rho_max_nmax();

}  // namespace pimc
