// pimc_host.hpp -- host-side context of the C-ABI (one chain resident in HBM).
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/pimc_gpu.h"
#include "pimc_direct.hpp"
#include "pimc_kernels.hpp"

namespace pimc {

struct AzizParams {
    double eps, rm, D, alpha, beta, C6, C8, C10, A;
};

bool aziz_params(int year, AzizParams* p);
double aziz_value(const AzizParams& p, double r, int what);  // valueV / dVdr / d2Vdr2
AzizDev make_aziz_dev(const AzizParams& p, double max_sep);
void build_table(const AzizParams& p, int length, int what, std::vector<double>& out);
double aziz_tail(const AzizParams& p, double rc);

// q = (2 pi / L) n with integer n, |n_d| <= rho_max_nmax(): the lattice
// wavevectors whose S(q) and F(q,tau) use the density-mode kernels.
bool lattice_q(const double* q, double L, int* n3);

struct TableCache;
TableCache& table_cache();
const double* cached_table(int year, int what, int length, std::string* err);

}  // namespace pimc

struct pimc_ctx {
    pimc_desc desc{};
    pimc::AzizParams params{};
    pimc::AzizDev az{};
    pimc::BoxDev box{};
    pimc::WormDev worm{};
    double max_sep = 0, tailV = 0;
    int64_t plane = 0;
    int cells = 1;
    hipStream_t stream = nullptr;
    bool have_path = false;
    uint64_t path_version = 0;    // bumped by every call that changes positions, mask, worm or action factors
    std::string err;

    // device state
    double* d_pos = nullptr;      // [3][M][N]
    uint8_t* d_on = nullptr;      // [M][N]
    int* d_nbs = nullptr;         // [M]
    int* d_cell = nullptr;        // [M][N] packed cell number (-1: not in the table)
    int* d_label = nullptr;       // [M][N]
    int* d_numlab = nullptr;      // [cells][M]
    int* d_hash = nullptr;        // [cells][M][N]
    int* d_grid_nn = nullptr;     // [cells][27]
    pimc::SliceDesc* d_slices = nullptr;
    pimc::WalkerDesc* d_walker = nullptr;
    pimc::RowTile* d_tiles = nullptr;
    int* d_ranges = nullptr;      // [n_groups + 1] tile ranges of the persistent pair kernel
    double* d_partials = nullptr; // one per tile
    int n_tiles = 0, cap_tiles = 0;
    int n_groups = 0, cap_groups = 0;
    int num_cus = 256;
    int device = 0;               // the HIP device current when the context was created
    double* d_vslice = nullptr;
    double* d_out = nullptr;
    double* h_map = nullptr;      // 64 doubles of mapped, coherent pinned host memory
    double* d_map = nullptr;      // its device address
    // completion word (slot kDoneSlot of the mapped block): the last kernel of a
    // call publishes the call's sequence number there
    static constexpr int kDoneSlot = 63;
    unsigned long long seq = 0;
    unsigned long long* d_done() { return reinterpret_cast<unsigned long long*>(d_map + kDoneSlot); }
    int wait_done(unsigned long long s);   // spin, then fall back to a stream sync
    // the direct Delta V form's per-bead results (mapped, coherent pinned block)
    pimc::VnnDirectOut* h_vdir = nullptr;
    pimc::VnnDirectOut* d_vdir = nullptr;
    int wait_direct(int n, unsigned long long s, bool packet);   // every slot of n carries s
    // short Delta V requests dispatched through the device's AQL queue (the
    // kernel-argument ring and k_vnn_direct per mode; direct_state -1: not set
    // up yet, 0: launched through HIP, 1: packets)
    pimc::DirectRing dring;
    pimc::DirectKernel dk_vdir[3];
    int direct_state = -1;
    long direct_count = 0;   // requests dispatched as packets
    bool direct_ready(int mode);
    // Every entry point returns only after its work on the stream has finished
    // or its last kernel has published its result (nothing of that kernel is
    // left but its exit), so the stream is quiet between calls -- unless a call
    // failed part-way.  Failures are counted; the first packet after one waits
    // for the stream.
    unsigned long long fail_count = 0, fail_seen = 0;
    bool stream_quiet();
    int lds_beads = 0;            // slice positions staged in LDS (0: read from HBM)
    int pair_lds = 0;             // the pair kernel's staging (lds_beads, or 0 when tuned off)
    // launch shape of the pair kernel (pimc_set_launch_shape / pimc_autotune; 0 / -1: automatic)
    int tune_ppl = 0, tune_gmul = 0, tune_lds = -1;
    bool pair_spatial = false;   // pair kernel stages slices in spatial order (mixed mode)
    int* d_perm = nullptr;        // [M][N] spatial order of each slice (pair_spatial)
    uint64_t perm_version = ~0ull;   // path_version d_perm was computed for (~0: none)
    // action factors (LocalAction VFactor, gradVFactor, estimator period); the
    // per-call vfactor argument overrides VFactor when given
    double vf_default[2] = {1.0, 1.0};
    double gvf[2] = {0.0, 0.0};
    int period = 1;
    bool correction() const { return gvf[0] > pimc::kCorrEps || gvf[1] > pimc::kCorrEps; }
    double* d_g2part = nullptr;   // gradVSquared partials [M][g2_blocks]
    double* d_g2slice = nullptr;  // gradVSquared(slice) [M]
    int g2_blocks = 0;
    double* d_gnn_part = nullptr; // gradVnnSquared partials [cap_req][2][chunks][4]
    int cap_gnn = 0;
    pimc::WalkerDesc walker_cache{};
    bool walker_valid = false;
    int* d_hist = nullptr;
    double* d_kin_partials = nullptr;
    // requests
    int cap_req = 0;
    pimc::VnnBead* d_req_beads = nullptr;
    pimc::VnnBead* h_req_beads = nullptr;
    int* d_req_sp = nullptr;
    double* d_req_pos = nullptr;
    int* h_req_sp = nullptr;
    double* h_req_pos = nullptr;
    double* d_per_bead = nullptr;
    double* h_per_bead = nullptr;
    double* d_g2f = nullptr;            // gradVSquared partial forces [M][chunks][N][3]
    size_t cap_g2f = 0;
    double* d_virial = nullptr;         // virial per-workgroup sums [M][blocks][7]
    double* d_vf = nullptr;             // virial chunk partials [M][chunks][N][9]
    size_t cap_vf = 0;
    size_t cap_virial = 0;
    // the virial estimator's links (next, then prev) on the device, and the next
    // links and beads per slice they were checked against: a call with the same
    // links on the same beads uploads nothing
    int* d_links = nullptr;
    size_t cap_links = 0;
    std::vector<int32_t> links_next;
    std::vector<int> links_nbs;
    bool links_ok = false;
    // the energy estimator's next links on the device and the links they were
    // checked against (the same links again upload nothing)
    int* d_enext = nullptr;
    size_t cap_enext = 0;
    std::vector<int32_t> enext;
    bool enext_ok = false;
    double* d_per_bead_cor = nullptr;   // per-bead corrections [cap_req][2]
    double* h_per_bead_cor = nullptr;
    double* d_vnn_partials = nullptr;
    unsigned* d_vnn_ticket = nullptr;   // arrival counter of the one-kernel Delta V (zero between calls)
    // structure factors
    int cap_q = 0;
    double* d_q = nullptr;
    int* d_qn = nullptr;
    double* d_rho = nullptr;
    size_t cap_rho = 0;
    double* d_fqt = nullptr;
    size_t cap_fqt = 0;
    // F(q,tau) of long lattice q lists into page-locked memory: a second stream
    // copies each chunk's rows out while the next chunk computes
    hipStream_t copy_stream = nullptr;
    std::vector<hipEvent_t> chunk_ev;
    long isf_chunks = 0;               // chunks computed that way (pimc_kernel_time "isf_chunks")
    // staging
    double* h_stage = nullptr;    // >= 3*M*N doubles, pinned
    double* h_out = nullptr;      // 64 doubles, pinned
    // host mirrors
    std::vector<int> h_nbs;
    std::vector<uint8_t> h_on;
    std::vector<int> grid_nn;

    std::vector<void*> dev_allocs, host_allocs;

    // timing: a pool of reusable event pairs; durations are summed lazily, when
    // the pool is exhausted or the totals are read, so timed calls add only the
    // event records
    bool timing = false;
    unsigned timing_kinds = ~0u;        // kinds bracketed while timing (bit k: kKernelNames[k])
    struct Pending {
        int kind;
        int ev;    // index into ev_pool
    };
    std::vector<hipEvent_t> ev_pool;   // 2 per slot
    std::vector<Pending> pending;
    int ev_next = 0;
    // pair_action, delta_action, rho, ssf, isf, gradv2, gradvnn, virial, chain_run
    static constexpr int kKinds = 9;
    double kernel_ms[kKinds] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    long kernel_count[kKinds] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    int timed_slot(int kind);          // -1 when timing is off
    bool ensure_event_pool();
    void timed_begin(int kind);
    void timed_end();
    pimc::KernelTimer kernel_timer(int kind);   // events for one kernel's own duration
    void resolve_timing();             // sums the pending durations (waits for them)
    int pending_end = -1;              // slot opened by timed_begin

    ~pimc_ctx();
};

namespace pimc {
// The context's Aziz constants with the table pointers its mode needs.
AzizDev context_aziz(pimc_ctx* c, int* rc);
}  // namespace pimc
