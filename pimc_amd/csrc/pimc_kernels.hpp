// pimc_kernels.hpp -- kernel-side data descriptors shared by the host code and
// the kernels (pair action, Delta V, cell lists, structure factors).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pimc_device.hpp"

namespace pimc {

constexpr int kPairThreads = 256;   // workgroup of the all-pairs kernel
constexpr int kVnnThreads = 256;    // workgroup of the Vnn (Delta V) kernel
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
    WormDev worm;
};

// A contiguous range of flattened pair indices of one slice.  Pair p of a slice
// with n beads is (i, j = (i + d) mod n) with d = 1 + p / n, i = p % n,
// 0 <= p < n * (n/2); for even n the pairs d = n/2, i >= n/2 repeat and are masked.
struct PairItem {
    int slice;      // SliceDesc index
    int p_begin;
    int p_count;
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
};

// Requests of up to kVnnInline beads travel in the kernel arguments.
constexpr int kVnnInline = 32;
struct VnnInline {
    VnnBead beads[kVnnInline];
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

// ---- launchers (defined in pimc_pair.hip / pimc_sq.hip) ----
hipError_t launch_pair_action(int mode, hipStream_t st, const PairItem* items, int n_items, const SliceDesc* slices,
                              int n_slices, const WalkerDesc* walkers, int n_walkers, const double* pos,
                              const AzizDev& az, int lds_beads, double* partials, double* v_slice, double* action);
hipError_t launch_vnn(int mode, hipStream_t st, const VnnArgs& A, const VnnInline& req, const AzizDev& az);
hipError_t launch_update_beads(hipStream_t st, const int* sp, const double* p3, int n, double* pos, int64_t plane,
                               int nalloc, int num_slices, int* cell, int* label, int* numlab, int* hash,
                               const BoxDev& box);
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
