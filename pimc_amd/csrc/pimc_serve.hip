// pimc_serve.hip -- the serving form: Delta V and S(q) requests from many
// concurrent Markov chains with continuous batching.
//
// * Chains are admitted into a pre-allocated HBM arena (the budget) by first fit;
//   a chain that does not fit waits until another chain is removed.
// * Worldlines stay resident; a request carries only its bead list, the proposed
//   positions and the uniform deviate of its Metropolis test
//   (rand() < exp(-(A_new - A_old)), src/move.cpp:1248, :1394).
// * A scheduler thread packs every ready request (at most one in flight per
//   chain, so a chain's moves stay ordered) into a batch, uploads one
//   descriptor, and launches on one of up to four streams (default priority)
//   while earlier batches still run, so batches overlap on the device.
// * Each request is reduced by its own fixed-order tree and decided and
//   committed on the device, so its result does not depend on the batch it
//   travelled in.
// * An S(q) request (StaticStructureFactorEstimator::accumulate on a resident
//   chain) takes a batch slot of its own when one is free: the density-mode
//   kernels run over its q list in runs that fit the slot's rho buffer and write
//   S(q) into the slot's mapped result block.  It is queued with the chain's
//   moves, so it sees every move submitted before it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <random>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "pimc_host.hpp"

namespace pimc {

struct ServeChainDev {
    double* pos;        // SoA planes x|y|z of [M][N]
    int64_t plane;
    const uint8_t* on;  // [M][N]
    int* cell;          // [M][N] packed cell numbers
    int* nbs;           // [M] beads per slice (N: every bead is on)
    int nalloc, num_slices;
    int table_length;
    BoxDev box;
    double tau;
};

struct ServeReqDev {
    int chain;          // index into the chain table
    int n_beads;
    int first_bead;     // into the batch bead array
    int chunks;         // partner chunks per bead
    int first_group;
    int commit_mode;    // 0: evaluate only, 1: Metropolis with u, 2: always commit
    int shift;
    int out_slot;
    double u;
    double vfactor0, vfactor1;
};

struct ServeGroupDev {
    int req;
    int bead;           // index into the batch bead array
    int chunk;
};

// One request's result in the slot's mapped block: A_old, A_new, dU, accepted,
// then the batch's sequence number once the four are visible.
struct ServeOut {
    double v[4];
    unsigned long long seq;
    unsigned long long pad;
};

template <int MODE>
__global__ __launch_bounds__(kVnnThreads) void k_serve_vnn(const ServeGroupDev* __restrict__ groups,
                                                           const ServeReqDev* __restrict__ reqs,
                                                           const VnnBead* __restrict__ beads,
                                                           const ServeChainDev* __restrict__ chains, AzizDev az,
                                                           int chunk, double* __restrict__ partials) {
    __shared__ double red[kVnnThreads / 64];
    const ServeGroupDev g = groups[blockIdx.x];
    const ServeReqDev& R = reqs[g.req];
    const ServeChainDev C = chains[R.chain];
    const VnnBead q = beads[g.bead];
    AzizDev a = az;
    a.table_length = C.table_length;
    const double inv_dr = 1.0 / a.dr;
    const int m = q.slice, me = q.ptcl;
    const double* px = C.pos + (int64_t)m * C.nalloc;
    const double* py = px + C.plane;
    const double* pz = py + C.plane;
    const double ox = px[me], oy = py[me], oz = pz[me];
    const double nx = q.new_pos[0], ny = q.new_pos[1], nz = q.new_pos[2];
    const BoxDev box = C.box;
    const bool use_cells = box.ngrid >= 4;
    int c_old = 0, c_new = 0;
    if (use_cells) {
        c_old = C.cell[m * C.nalloc + me];
        c_new = cell_of(nx, ny, nz, box);
    }
    const bool me_on = C.on[m * C.nalloc + me] != 0;
    double acc_o = 0.0, acc_n = 0.0;
    const int j_end = min((g.chunk + 1) * chunk, C.nalloc);
    for (int j = g.chunk * chunk + threadIdx.x; j < j_end; j += kVnnThreads) {
        const bool on = (C.on[m * C.nalloc + j] != 0) && (j != me) && me_on;
        const double xj = px[j], yj = py[j], zj = pz[j];
        const int cj = use_cells ? C.cell[m * C.nalloc + j] : 0;
        {
            const double r2 = sep_r2<MODE>(ox, oy, oz, xj, yj, zj, box);
            const bool ok = on && (r2 < box.rc2) && (!use_cells || cells_adjacent(c_old, cj, box.ngrid));
            const double v = pair_value_from_r2<MODE>(r2, a, inv_dr);
            acc_o += ok ? v : 0.0;
        }
        {
            const double r2 = sep_r2<MODE>(nx, ny, nz, xj, yj, zj, box);
            const bool ok = on && (r2 < box.rc2) && (!use_cells || cells_adjacent(c_new, cj, box.ngrid));
            const double v = pair_value_from_r2<MODE>(r2, a, inv_dr);
            acc_n += ok ? v : 0.0;
        }
    }
    const double so = block_sum<kVnnThreads>(acc_o, red);
    const double sn = block_sum<kVnnThreads>(acc_n, red);
    if (threadIdx.x == 0) {
        partials[2 * (int64_t)blockIdx.x + 0] = so;
        partials[2 * (int64_t)blockIdx.x + 1] = sn;
    }
}

// Publishes an S(q) batch's sequence number once every kernel before it on the
// stream has finished (stream order), so the worker reads a host cache line
// instead of querying an event.
__global__ void k_serve_publish(unsigned long long* done, unsigned long long seq) {
    if (threadIdx.x == 0) publish_done(done, seq);
}

// One wave per request: per-bead sums (chunks in order), the actions and dU
// (fixed shuffle tree), the Metropolis decision, the commit, then the result
// and the batch's sequence number into the request's record of the slot's
// mapped block (the worker retires a batch once every record carries it).
__global__ __launch_bounds__(64) void k_serve_finalize(const ServeReqDev* __restrict__ reqs,
                                                       const VnnBead* __restrict__ beads,
                                                       const ServeChainDev* __restrict__ chains,
                                                       const double* __restrict__ partials,
                                                       ServeOut* __restrict__ out, unsigned long long seq) {
    const ServeReqDev R = reqs[blockIdx.x];
    const ServeChainDev C = chains[R.chain];
    const int lane = threadIdx.x;
    const double tau_s = R.shift * C.tau;
    double ao = 0.0, an = 0.0, du = 0.0;
    for (int k = lane; k < R.n_beads; k += 64) {
        double vo = 0.0, vn = 0.0;
        const int g0 = R.first_group + k * R.chunks;
        for (int c = 0; c < R.chunks; ++c) {
            vo += partials[2 * (int64_t)(g0 + c)];
            vn += partials[2 * (int64_t)(g0 + c) + 1];
        }
        const double co = ((beads[R.first_bead + k].slice & 1) ? R.vfactor1 : R.vfactor0) * tau_s;
        ao += co * vo;
        an += co * vn;
        du += co * (vn - vo);
    }
    ao = wave_sum(ao);
    an = wave_sum(an);
    du = wave_sum(du);
    const bool accept = R.commit_mode == 2 || (R.commit_mode == 1 && R.u < exp(-du));
    if (accept) {
        for (int k = lane; k < R.n_beads; k += 64) {
            const VnnBead qb = beads[R.first_bead + k];
            const int64_t bi = (int64_t)qb.slice * C.nalloc + qb.ptcl;
            C.pos[bi] = qb.new_pos[0];
            C.pos[bi + C.plane] = qb.new_pos[1];
            C.pos[bi + 2 * C.plane] = qb.new_pos[2];
            if (C.box.ngrid >= 4) C.cell[bi] = cell_of(qb.new_pos[0], qb.new_pos[1], qb.new_pos[2], C.box);
        }
    }
    if (lane == 0) {
        ServeOut* o = out + R.out_slot;
        o->v[0] = ao;
        o->v[1] = an;
        o->v[2] = du;
        o->v[3] = accept ? 1.0 : 0.0;
        publish_done(&o->seq, seq);
    }
}

__global__ void k_serve_cells(const double* __restrict__ pos, int64_t plane, int total, int* __restrict__ cell,
                              BoxDev box) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < total) cell[t] = cell_of(pos[t], pos[t + plane], pos[t + 2 * plane], box);
}

}  // namespace pimc

using namespace pimc;
using Clock = std::chrono::steady_clock;
constexpr size_t kZeroCopyBytes = 16384;   // batches read in place from mapped host memory up to this size
// Delta V batches whose descriptors fit this many bytes are written through the
// BAR into the slot's device block and sent as AQL packets on the slot's own
// queue (larger ones: one DMA copy and two HIP launches on the slot's stream)
constexpr size_t kPacketDescBytes = 65536;

namespace {

constexpr int kSlots = 4;            // batches in flight
constexpr int kMaxReqPerBatch = 4096;
constexpr int kMaxSsfQ = 1 << 16;                      // q-vectors of one S(q) request
constexpr size_t kSsfRhoBytes = size_t(32) << 20;      // rho_q(m) buffer of a slot
constexpr int kMaxBeadsPerBatch = 1 << 17;
constexpr int kMaxGroupsPerBatch = 1 << 18;

struct Arena {
    char* base = nullptr;
    size_t size = 0;
    std::map<size_t, size_t> free_;   // offset -> length
    bool alloc(size_t n, size_t* off) {
        n = (n + 255) & ~size_t(255);
        for (auto it = free_.begin(); it != free_.end(); ++it)
            if (it->second >= n) {
                *off = it->first;
                size_t rest = it->second - n, at = it->first + n;
                free_.erase(it);
                if (rest) free_[at] = rest;
                return true;
            }
        return false;
    }
    void release(size_t off, size_t n) {
        n = (n + 255) & ~size_t(255);
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    size_t free_bytes() const {
        size_t s = 0;
        for (auto& kv : free_) s += kv.second;
        return s;
    }
};

struct Chain {
    pimc_desc d{};
    int state = 0;                    // 0 waiting for memory, 1 resident, 2 removed
    size_t off = 0, bytes = 0;
    ServeChainDev dev{};
    int dev_index = -1;
    bool busy = false;
    std::deque<uint64_t> queue;
    std::vector<double> pending_beads;  // positions while waiting for admission
};

struct Request {
    uint64_t id = 0;
    int kind = 0;                     // 0: Delta V, 1: S(q)
    int chain = -1;
    int shift = 1;
    double vf[2] = {1.0, 1.0};
    int commit_mode = 0;
    double u = 0.0;
    std::vector<VnnBead> beads;
    std::vector<int> qn;              // S(q): lattice components [nq][3]
    std::vector<double> sq;           // S(q): result [nq]
    Clock::time_point t_submit, t_done;
    bool done = false;
    int tag = -1;                     // load-generator thread that submitted it (-1: none)
    double out[4] = {0, 0, 0, 0};
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t ev = nullptr;
    bool busy = false;
    char* h_desc = nullptr;           // pinned, mapped: reqs | beads | groups
    char* m_desc = nullptr;           // the device's address of h_desc (read in place)
    char* d_desc = nullptr;
    double* d_partials = nullptr;
    ServeOut* h_out = nullptr;        // mapped pinned results [kMaxReqPerBatch]
    ServeOut* d_out = nullptr;
    int n_reqs = 0;                   // requests of the Delta V batch in flight
    unsigned long long* h_done = nullptr;   // mapped completion word: the last finished batch
    unsigned long long* d_done = nullptr;
    unsigned long long seq = 0;             // the batch in flight
    Clock::time_point t_fire{};
    std::vector<uint64_t> reqs;
    size_t desc_bytes = 0;
    int kind = 0;                     // batch in flight: 0 Delta V requests, 1 one S(q) request
    double* d_rho = nullptr;          // S(q): rho_q(m) of one run [M][run][2]
    double* h_sq = nullptr;           // S(q): mapped results [kMaxSsfQ]
    double* m_sq = nullptr;           // the device's address of h_sq
    // the batch's requests and chains, picked under the server lock and read by
    // build_batch after it (requests are immutable until retired, and the chain
    // deque keeps its elements in place)
    std::vector<std::pair<const Request*, const Chain*>> items;
    DirectRing ring;                  // kernel arguments and the slot's own AQL queue
    char* v_desc = nullptr;           // descriptor block in device memory, written through the BAR
    bool packet = false;              // the batch in flight went as packets
};

// Chain extents readable without the server lock: an append-only table,
// replaced by a copy of twice the capacity when full (the old tables stay alive
// until the server goes, so a reader holding one is never left dangling).
struct ChainExtents {
    int num_slices, num_alloc;
};
struct ExtentsTable {
    std::unique_ptr<ChainExtents[]> v;
    int cap = 0;
    std::atomic<int> n{0};
};

// One packed batch, handed from pack_batch (under the lock) to fire_batch (without).
struct BatchLaunch {
    int slot = 0;
    size_t o_bd = 0, o_gr = 0, bytes = 0;
    int n_groups = 0, n_reqs = 0;
    AzizDev az{};
    const ServeChainDev* chains = nullptr;
    // S(q) batch: the chain, its q count and inverse particle count (qn staged in h_desc)
    int kind = 0;
    ServeChainDev chain{};
    int nq = 0;
    double inv_n = 0.0, invL = 0.0;
    bool packet = false;              // descriptors in the slot's device block, kernels as packets
};

// k_serve_vnn's and k_serve_finalize's argument blocks as the AMDGPU ABI lays
// them out (each argument at its natural alignment)
struct ServeVnnArgs {
    const ServeGroupDev* groups;
    const ServeReqDev* reqs;
    const VnnBead* beads;
    const ServeChainDev* chains;
    AzizDev az;
    int chunk;
    double* partials;
};
struct ServeFinArgs {
    const ServeReqDev* reqs;
    const VnnBead* beads;
    const ServeChainDev* chains;
    const double* partials;
    ServeOut* out;
    unsigned long long seq;
};

}  // namespace

struct pimc_server {
    pimc_server_desc desc{};
    AzizParams params{};
    AzizDev az{};
    Arena arena;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    // a deque: add_chain's push_back keeps references to the other chains valid
    // (remove_chain and download hold one across a wait that drops the lock)
    std::deque<Chain> chains;
    std::deque<int> waiting_chains;
    std::atomic<ExtentsTable*> extents{nullptr};
    std::vector<std::unique_ptr<ExtentsTable>> extents_all;   // every table ever published
    std::vector<ServeChainDev> chain_table;
    ServeChainDev* d_chains = nullptr;
    int chain_cap = 0;
    bool chains_dirty = false;
    std::unordered_map<uint64_t, Request> reqs;
    std::deque<uint64_t> ready;       // chains with queued work, by request id order
    uint64_t next_id = 1;
    Slot slots[kSlots];
    int n_slots = 2;
    std::thread worker;
    std::atomic<bool> stop{false};
    std::atomic<uint64_t> work_events{0};   // bumped on every submit / state change the worker must see
    std::string err;
    int64_t launched_batches = 0, launched_reqs = 0;
    int64_t pairs_done = 0;
    double max_table_sep = 0;
    int kernel_error = 0;
    std::string kernel_error_msg;
    bool track_done = false;
    std::vector<std::vector<uint64_t>> done_lists;   // completed request ids per load-generator thread
    // completions per load-generator thread, read without the lock
    std::unique_ptr<std::atomic<uint64_t>[]> done_counts{new std::atomic<uint64_t>[64]()};
    uint64_t packed_events = 0;       // work_events as of the last packing pass

    std::vector<ServeReqDev> pack_rq;
    std::vector<VnnBead> pack_bd;
    std::vector<ServeGroupDev> pack_gr;

    int64_t launched_ssf = 0;

    // Delta V batches as AQL packets (PIMC_DIRECT_DISPATCH=0: HIP launches only)
    int device = 0;
    bool direct = false;
    DirectKernel dk_vnn[3], dk_fin;
    std::atomic<int64_t> packet_batches{0}, hip_batches{0};   // Delta V batches by how they went out
    bool setup_direct();

    void run();
    int pack_batch(int slot, BatchLaunch* L);
    void build_batch(BatchLaunch* L);
    hipError_t fire_batch(const BatchLaunch& L);
    hipError_t fire_ssf(const BatchLaunch& L);
    void complete(Slot& s);
    void admit_waiting();
};

namespace {

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// Arena block of a chain: positions, cell numbers, bead mask, beads per slice.
size_t chain_bytes(const pimc_desc& d) {
    const size_t MN = (size_t)d.num_slices * d.num_alloc;
    return align256(MN * 3 * sizeof(double)) + align256(MN * sizeof(int)) + align256(MN) +
           align256(sizeof(int) * (size_t)d.num_slices);
}

int sfail(pimc_server* s, int code, const std::string& m) {
    if (s) s->err = m;
    return code;
}

// Place a chain into the arena: positions, bead mask, cell numbers.
int place_chain(pimc_server* s, int id, const double* beads) {
    Chain& c = s->chains[id];
    const pimc_desc& d = c.d;
    const int M = d.num_slices, N = d.num_alloc;
    const size_t MN = (size_t)M * N;
    size_t off;
    if (!s->arena.alloc(c.bytes, &off)) return 1;   // no room yet
    c.off = off;
    char* base = s->arena.base + off;
    double* pos = (double*)base;
    int* cell = (int*)(base + align256(MN * 3 * sizeof(double)));
    uint8_t* on = (uint8_t*)((char*)cell + align256(MN * sizeof(int)));
    int* nbs = (int*)((char*)on + align256(MN));
    std::vector<double> soa(3 * MN);
    for (size_t t = 0; t < MN; ++t) {
        soa[t] = beads[3 * t];
        soa[t + MN] = beads[3 * t + 1];
        soa[t + 2 * MN] = beads[3 * t + 2];
    }
    hipStream_t st = s->slots[0].stream;
    if (hipMemcpyAsync(pos, soa.data(), sizeof(double) * 3 * MN, hipMemcpyHostToDevice, st) != hipSuccess) return -1;
    std::vector<uint8_t> ones(MN, 1);
    if (hipMemcpyAsync(on, ones.data(), MN, hipMemcpyHostToDevice, st) != hipSuccess) return -1;
    std::vector<int> counts(M, N);
    if (hipMemcpyAsync(nbs, counts.data(), sizeof(int) * M, hipMemcpyHostToDevice, st) != hipSuccess) return -1;
    ServeChainDev& cd = c.dev;
    cd.pos = pos;
    cd.plane = (int64_t)MN;
    cd.on = on;
    cd.cell = cell;
    cd.nbs = nbs;
    cd.nalloc = N;
    cd.num_slices = M;
    BoxDev& b = cd.box;
    b.L = d.side;
    b.invL = 1.0 / d.side;
    b.rc2 = d.rc * d.rc;
    int ng = (int)std::floor((d.side / d.rc) + 1.0e-7);
    b.ngrid = ng < 1 ? 1 : ng;
    b.grid_size = d.side / (1.0 * b.ngrid);
    const double h = d.side / 2.0;
    double ms = 0.0;
    ms += h * h;
    ms += h * h;
    ms += h * h;
    cd.table_length = int(std::sqrt(ms) / ((1.0E-6) * s->params.rm));
    cd.tau = d.tau;
    hipLaunchKernelGGL(k_serve_cells, dim3((unsigned)((MN + 255) / 256)), dim3(256), 0, st, pos, (int64_t)MN,
                       (int)MN, cell, b);
    if (hipStreamSynchronize(st) != hipSuccess) return -1;
    // chain table entry
    int idx = -1;
    for (int k = 0; k < (int)s->chain_table.size(); ++k)
        if (s->chain_table[k].pos == nullptr) {
            idx = k;
            break;
        }
    if (idx < 0) {
        idx = (int)s->chain_table.size();
        s->chain_table.push_back(cd);
    } else {
        s->chain_table[idx] = cd;
    }
    c.dev_index = idx;
    c.state = 1;
    s->chains_dirty = true;
    return 0;
}

// A batch has finished when its records (Delta V) or its completion word (S(q))
// carry its sequence number; its kernels have then read everything they read.
bool slot_finished(const Slot& S) {
    if (S.kind == 1) return __atomic_load_n(S.h_done, __ATOMIC_ACQUIRE) == S.seq;
    for (int k = S.n_reqs - 1; k >= 0; --k)
        if (__atomic_load_n(&S.h_out[k].seq, __ATOMIC_ACQUIRE) != S.seq) return false;
    return true;
}

// Every batch in flight, launched through HIP or sent as packets, is done with
// the chain table (a packet batch is waited for through its records, for at
// most 10 s).
void quiesce_slots(pimc_server* s) {
    for (auto& sl : s->slots) {
        if (sl.stream) (void)hipStreamSynchronize(sl.stream);
        if (sl.busy && sl.packet) {
            const auto t0 = Clock::now();
            while (!slot_finished(sl) && Clock::now() - t0 < std::chrono::seconds(10)) __builtin_ia32_pause();
        }
    }
}

int sync_chain_table(pimc_server* s, hipStream_t st) {
    if (!s->chains_dirty) return 0;
    const int n = (int)s->chain_table.size();
    if (n > s->chain_cap) {
        if (s->d_chains) {
            quiesce_slots(s);
            (void)hipFree(s->d_chains);
        }
        s->chain_cap = std::max(64, 2 * n);
        if (hipMalloc(&s->d_chains, sizeof(ServeChainDev) * s->chain_cap) != hipSuccess) return -1;
    }
    // every slot must see the new table before its next batch
    quiesce_slots(s);
    if (hipMemcpy(s->d_chains, s->chain_table.data(), sizeof(ServeChainDev) * n, hipMemcpyHostToDevice) != hipSuccess)
        return -1;
    s->chains_dirty = false;
    (void)st;
    return 0;
}

}  // namespace

void pimc_server::admit_waiting() {
    while (!waiting_chains.empty()) {
        int id = waiting_chains.front();
        Chain& c = chains[id];
        if (c.state != 0) {
            waiting_chains.pop_front();
            continue;
        }
        int r = place_chain(this, id, c.pending_beads.data());
        if (r != 0) break;   // still no room: keep arrival order
        c.pending_beads.clear();
        c.pending_beads.shrink_to_fit();
        waiting_chains.pop_front();
        if (!c.queue.empty()) ready.push_back(c.queue.front());
    }
}

// Packs every ready request that fits (at most one in flight per chain) into
// slot S's pinned descriptor block and marks the slot busy.  Runs under the
// server lock; returns 1 when a batch was packed, 0 when nothing is ready.
// Picks a slot's batch from the ready requests (under the server lock): an S(q)
// request alone, or Delta V requests up to the batch limits.
int pimc_server::pack_batch(int si, BatchLaunch* L) {
    Slot& S = slots[si];
    if (ready.empty()) return 0;
    if (sync_chain_table(this, S.stream) != 0) {
        kernel_error = PIMC_E_HIP;
        kernel_error_msg = "chain table upload failed";
        return -1;
    }
    S.reqs.clear();
    S.items.clear();
    if (reqs[ready.front()].kind == 1) {
        // an S(q) request takes the slot alone
        const uint64_t id = ready.front();
        ready.pop_front();
        Request& R = reqs[id];
        Chain& C = chains[R.chain];
        L->kind = 1;
        L->slot = si;
        L->chain = C.dev;
        L->nq = (int)R.qn.size() / 3;
        L->inv_n = 1.0 / C.d.num_alloc;
        L->invL = 1.0 / C.d.side;
        S.items.emplace_back(&R, &C);
        S.reqs.push_back(id);
        S.kind = 1;
        C.busy = true;
        S.busy = true;
        S.packet = false;   // not fired yet
        S.seq += 1;
        S.t_fire = Clock::now();
        ++launched_batches;
        ++launched_reqs;
        ++launched_ssf;
        return 1;
    }
    S.kind = 0;
    L->kind = 0;
    int n_bd = 0, n_gr = 0;
    while (!ready.empty() && (int)S.items.size() < kMaxReqPerBatch) {
        const uint64_t id = ready.front();
        Request& R = reqs[id];
        if (R.kind == 1) break;       // S(q) goes in the next free slot
        Chain& C = chains[R.chain];
        const int chunks = (C.d.num_alloc + kVnnChunk - 1) / kVnnChunk;
        const int nb = (int)R.beads.size();
        if (!S.items.empty() && (n_bd + nb > kMaxBeadsPerBatch || n_gr + nb * chunks > kMaxGroupsPerBatch)) break;
        ready.pop_front();
        S.items.emplace_back(&R, &C);
        n_bd += nb;
        n_gr += nb * chunks;
        S.reqs.push_back(id);
        C.busy = true;
        pairs_done += 2LL * nb * (C.d.num_alloc - 1);
    }
    if (S.items.empty()) return 0;
    L->n_reqs = (int)S.items.size();
    S.n_reqs = L->n_reqs;
    L->slot = si;
    L->chains = d_chains;
    S.busy = true;
    S.packet = false;   // not fired yet
    S.seq += 1;   // the sequence number this batch publishes when it completes
    S.t_fire = Clock::now();
    ++launched_batches;
    launched_reqs += (int64_t)S.items.size();
    return 1;
}

// A picked batch's descriptors into the slot's staging block (worker thread,
// without the server lock).
void pimc_server::build_batch(BatchLaunch* L) {
    Slot& S = slots[L->slot];
    if (L->kind == 1) {
        const Request& R = *S.items[0].first;
        std::memcpy(S.h_desc, R.qn.data(), sizeof(int) * R.qn.size());
        L->bytes = sizeof(int) * R.qn.size();
        return;
    }
    std::vector<ServeReqDev>& rq = pack_rq;
    std::vector<VnnBead>& bd = pack_bd;
    std::vector<ServeGroupDev>& gr = pack_gr;
    rq.clear();
    bd.clear();
    gr.clear();
    for (const auto& it : S.items) {
        const Request& R = *it.first;
        const Chain& C = *it.second;
        const int chunks = (C.d.num_alloc + kVnnChunk - 1) / kVnnChunk;
        const int nb = (int)R.beads.size();
        ServeReqDev d{};
        d.chain = C.dev_index;
        d.n_beads = nb;
        d.first_bead = (int)bd.size();
        d.chunks = chunks;
        d.first_group = (int)gr.size();
        d.commit_mode = R.commit_mode;
        d.shift = R.shift;
        d.out_slot = (int)rq.size();
        d.u = R.u;
        d.vfactor0 = R.vf[0];
        d.vfactor1 = R.vf[1];
        for (int k = 0; k < nb; ++k) {
            bd.push_back(R.beads[k]);
            for (int c = 0; c < chunks; ++c) gr.push_back(ServeGroupDev{(int)rq.size(), d.first_bead + k, c});
        }
        rq.push_back(d);
    }
    const size_t b_rq = sizeof(ServeReqDev) * rq.size();
    const size_t b_bd = sizeof(VnnBead) * bd.size();
    const size_t b_gr = sizeof(ServeGroupDev) * gr.size();
    L->o_bd = (b_rq + 255) & ~size_t(255);
    L->o_gr = (L->o_bd + b_bd + 255) & ~size_t(255);
    L->bytes = L->o_gr + b_gr;
    L->n_groups = (int)gr.size();
    L->packet = direct && S.v_desc && L->bytes <= kPacketDescBytes;
    char* dst = L->packet ? S.v_desc : S.h_desc;
    std::memcpy(dst, rq.data(), b_rq);
    std::memcpy(dst + L->o_bd, bd.data(), b_bd);
    std::memcpy(dst + L->o_gr, gr.data(), b_gr);
    L->az = az;
    if (desc.v_mode == PIMC_V_TABLE) {
        std::string e;
        L->az.table = cached_table(desc.aziz_year, 0, (int)(max_table_sep / ((1.0E-6) * params.rm)) + 1, &e);
    }
}

// Kernel objects of k_serve_vnn / k_serve_finalize and, for each slot in use, an
// argument ring with a queue of its own (the slot's two packets run in order)
// and a descriptor block in device memory; false (nothing kept) when any part
// is unavailable, and every batch then launches through HIP.
bool pimc_server::setup_direct() {
    if (hipGetDevice(&device) != hipSuccess) return false;
    static_assert(sizeof(ServeVnnArgs) == 272 && offsetof(ServeVnnArgs, az) == 32 &&
                      offsetof(ServeVnnArgs, chunk) == 256 && offsetof(ServeVnnArgs, partials) == 264 &&
                      sizeof(ServeFinArgs) == 48,
                  "serve argument blocks");
    const void* vnn[3] = {reinterpret_cast<const void*>(&k_serve_vnn<0>),
                          reinterpret_cast<const void*>(&k_serve_vnn<1>),
                          reinterpret_cast<const void*>(&k_serve_vnn<2>)};
    bool ok = true;
    for (int m = 0; m < 3 && ok; ++m)
        ok = direct_kernel(device, vnn[m], slots[0].stream, (uint32_t)sizeof(ServeVnnArgs), &dk_vnn[m]);
    ok = ok && direct_kernel(device, reinterpret_cast<const void*>(&k_serve_finalize), slots[0].stream,
                             (uint32_t)sizeof(ServeFinArgs), &dk_fin);
    for (int k = 0; k < n_slots && ok; ++k) {
        ok = direct_ring_create(device, &slots[k].ring, true);
        if (ok) {
            slots[k].v_desc = static_cast<char*>(direct_alloc_visible(device, kPacketDescBytes));
            ok = slots[k].v_desc != nullptr;
        }
    }
    if (!ok) {
        for (auto& S : slots) {
            direct_ring_release(&S.ring);
            direct_free_visible(S.v_desc);
            S.v_desc = nullptr;
        }
        return false;
    }
    direct = true;
    return true;
}

// Copies a packed batch to the device and launches it on its slot's stream.
// Runs without the server lock (HIP calls are thread safe; the slot belongs to
// the worker thread while busy), so submitters are never held up by a launch.
hipError_t pimc_server::fire_batch(const BatchLaunch& L) {
    Slot& S = slots[L.slot];
    (void)hipGetLastError();
    S.packet = L.kind == 0 && L.packet;
    if (L.kind == 1) return fire_ssf(L);
    if (L.packet) {
        // both kernels' arguments into the slot's ring, one flush for them and
        // the descriptors, two packets in order on the slot's queue
        const ServeReqDev* d_rq = (const ServeReqDev*)S.v_desc;
        const VnnBead* d_bd = (const VnnBead*)(S.v_desc + L.o_bd);
        const ServeGroupDev* d_gr = (const ServeGroupDev*)(S.v_desc + L.o_gr);
        const ServeVnnArgs va{d_gr, d_rq, d_bd, L.chains, L.az, kVnnChunk, S.d_partials};
        const ServeFinArgs fa{d_rq, d_bd, L.chains, S.d_partials, S.d_out, S.seq};
        const ArgPiece pv{&va, 0, (uint32_t)sizeof(va)}, pf{&fa, 0, (uint32_t)sizeof(fa)};
        const char* kv = direct_stage(&S.ring, &pv, 1);
        const char* kf = direct_stage(&S.ring, &pf, 1);
        direct_flush(device);
        direct_submit(&S.ring, dk_vnn[desc.v_mode], kv, (uint32_t)L.n_groups, (uint32_t)kVnnThreads);
        direct_submit(&S.ring, dk_fin, kf, (uint32_t)L.n_reqs, 64u);
        packet_batches.fetch_add(1, std::memory_order_relaxed);
        return hipSuccess;
    }
    hip_batches.fetch_add(1, std::memory_order_relaxed);
    // A small batch's descriptors are read in place from mapped host memory: a
    // copy would put a DMA start in front of the kernels.  Larger batches go
    // over in one copy, since every workgroup re-reads its descriptors.
    hipError_t e = hipSuccess;
    const char* base = S.m_desc;
    if (L.bytes > kZeroCopyBytes || !base) {
        e = hipMemcpyAsync(S.d_desc, S.h_desc, L.bytes, hipMemcpyHostToDevice, S.stream);
        base = S.d_desc;
    }
    const ServeReqDev* d_rq = (const ServeReqDev*)base;
    const VnnBead* d_bd = (const VnnBead*)(base + L.o_bd);
    const ServeGroupDev* d_gr = (const ServeGroupDev*)(base + L.o_gr);
    if (e == hipSuccess) {
        dim3 g((unsigned)L.n_groups), b(kVnnThreads);
        if (desc.v_mode == 0)
            hipLaunchKernelGGL(k_serve_vnn<0>, g, b, 0, S.stream, d_gr, d_rq, d_bd, L.chains, L.az, kVnnChunk,
                               S.d_partials);
        else if (desc.v_mode == 1)
            hipLaunchKernelGGL(k_serve_vnn<1>, g, b, 0, S.stream, d_gr, d_rq, d_bd, L.chains, L.az, kVnnChunk,
                               S.d_partials);
        else
            hipLaunchKernelGGL(k_serve_vnn<2>, g, b, 0, S.stream, d_gr, d_rq, d_bd, L.chains, L.az, kVnnChunk,
                               S.d_partials);
        hipLaunchKernelGGL(k_serve_finalize, dim3((unsigned)L.n_reqs), dim3(64), 0, S.stream, d_rq, d_bd, L.chains,
                           S.d_partials, S.d_out, S.seq);
        e = hipGetLastError();
    }
    return e;
}

// An S(q) request: its q list goes to the slot's device block, then per run of
// q that fits the slot's rho buffer the density modes (k_rho_lattice) and
// S(q) = sum_m |rho_q(m)|^2 / N (k_ssf_from_rho) straight into the slot's mapped
// result block -- the kernels and the per-q order of pimc_ssf.
hipError_t pimc_server::fire_ssf(const BatchLaunch& L) {
    Slot& S = slots[L.slot];
    const ServeChainDev& C = L.chain;
    const int M = C.num_slices;
    hipError_t e = hipMemcpyAsync(S.d_desc, S.h_desc, L.bytes, hipMemcpyHostToDevice, S.stream);
    const int* d_qn = reinterpret_cast<const int*>(S.d_desc);
    const int* h_qn = reinterpret_cast<const int*>(S.h_desc);
    const int run = (int)std::max<size_t>(1, kSsfRhoBytes / (sizeof(double) * 2 * (size_t)M));
    for (int k0 = 0; e == hipSuccess && k0 < L.nq; k0 += run) {
        const int n = std::min(run, L.nq - k0);
        RhoArgs A{};
        A.pos = C.pos;
        A.plane = C.plane;
        A.nalloc = C.nalloc;
        A.num_slices = M;
        A.nbs = C.nbs;
        A.invL = L.invL;
        A.nq = n;
        A.qn = d_qn + 3 * (size_t)k0;
        A.rho = S.d_rho;
        for (int d = 0; d < 3; ++d) A.nmax[d] = 0;
        for (int k = k0; k < k0 + n; ++k)
            for (int d = 0; d < 3; ++d) A.nmax[d] = std::max(A.nmax[d], std::abs(h_qn[3 * k + d]));
        e = launch_rho(true, S.stream, A);
        if (e == hipSuccess) e = launch_isf(S.stream, S.d_rho, n, M, L.inv_n, S.m_sq + k0, 1);
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_serve_publish, dim3(1), dim3(64), 0, S.stream, S.d_done, S.seq);
        e = hipGetLastError();
    }
    return e;
}

void pimc_server::complete(Slot& S) {
    const auto now = Clock::now();
    for (size_t k = 0; k < S.reqs.size(); ++k) {
        Request& R = reqs[S.reqs[k]];
        if (R.kind == 1) {
            const int nq = (int)R.qn.size() / 3;
            R.sq.assign(S.h_sq, S.h_sq + nq);
            R.out[0] = nq;
            R.out[1] = R.out[2] = R.out[3] = 0.0;
        } else {
            for (int t = 0; t < 4; ++t) R.out[t] = S.h_out[k].v[t];
        }
        R.done = true;
        R.t_done = now;
        if (track_done && R.tag >= 0 && R.tag < (int)done_lists.size()) {
            done_lists[R.tag].push_back(R.id);
            done_counts[R.tag].fetch_add(1, std::memory_order_release);
        }
        Chain& C = chains[R.chain];
        C.busy = false;
        if (!C.queue.empty() && C.queue.front() == R.id) C.queue.pop_front();
        if (!C.queue.empty() && C.state == 1) ready.push_back(C.queue.front());
    }
    S.reqs.clear();
    S.items.clear();
    S.busy = false;
}

// The worker: polls the in-flight batches' events and fires new batches with the
// lock released; takes the lock only to retire results and to pack requests.
void pimc_server::run() {
    std::unique_lock<std::mutex> lk(mu, std::defer_lock);
    auto finished = [](const Slot& S) { return slot_finished(S); };
    while (!stop.load()) {
        int done_k[kSlots];
        hipError_t done_e[kSlots];
        int nd = 0;
        bool any_busy = false, any_free = false;
        for (int k = 0; k < n_slots; ++k) {
            Slot& S = slots[k];
            if (!S.busy) {
                any_free = true;
                continue;
            }
            any_busy = true;
            if (finished(S)) {
                done_k[nd] = k;
                done_e[nd++] = hipSuccess;
            } else if (S.packet) {
                // packets carry no stream: a batch 10 s late has failed
                if (Clock::now() - S.t_fire > std::chrono::seconds(10) && !finished(S)) {
                    done_k[nd] = k;
                    done_e[nd++] = hipErrorLaunchTimeOut;
                }
            } else if (Clock::now() - S.t_fire > std::chrono::milliseconds(200)) {
                // a batch this late may have failed: ask the stream
                const hipError_t q = hipStreamQuery(S.stream);
                if (q != hipErrorNotReady && !finished(S)) {
                    done_k[nd] = k;
                    done_e[nd++] = q == hipSuccess ? hipErrorUnknown : q;
                }
            }
        }
        // nothing finished and nothing submitted since the last packing pass (or
        // no slot free for it): no reason to take the lock
        const uint64_t ev = work_events.load(std::memory_order_acquire);
        if (nd == 0 && (ev == packed_events || !any_free)) {
            if (any_busy) {
                __builtin_ia32_pause();
                continue;
            }
            // idle: spin on the submit counter for up to 2 ms (a sleeping
            // worker costs every request a thread wake-up), then block; the
            // counter is read every 8 pauses (every 64 took ~0.4 us longer to
            // see a submit)
            const auto t0 = Clock::now();
            bool woke = false;
            while (!stop.load(std::memory_order_relaxed) && Clock::now() - t0 < std::chrono::milliseconds(2)) {
                if (work_events.load(std::memory_order_acquire) != packed_events) {
                    woke = true;
                    break;
                }
                for (int p = 0; p < 8; ++p) __builtin_ia32_pause();
            }
            if (!woke) {
                lk.lock();
                if (work_events.load(std::memory_order_acquire) == packed_events && !stop.load())
                    cv_work.wait_for(lk, std::chrono::milliseconds(1));
                lk.unlock();
            }
            continue;
        }
        BatchLaunch L[kSlots];
        int nl = 0;
        lk.lock();
        packed_events = work_events.load(std::memory_order_acquire);
        for (int t = 0; t < nd; ++t) {
            if (done_e[t] != hipSuccess) {
                kernel_error = PIMC_E_HIP;
                kernel_error_msg = std::string("batch failed: ") + hipGetErrorString(done_e[t]);
            }
            complete(slots[done_k[t]]);
        }
        if (nd) cv_done.notify_all();
        if (!kernel_error)
            for (int k = 0; k < n_slots; ++k) {
                if (slots[k].busy) continue;
                const int r = pack_batch(k, &L[nl]);
                if (r <= 0) break;
                ++nl;
            }
        lk.unlock();
        for (int t = 0; t < nl; ++t) {
            build_batch(&L[t]);
            const hipError_t e = fire_batch(L[t]);
            if (e != hipSuccess) {
                lk.lock();
                kernel_error = PIMC_E_HIP;
                kernel_error_msg = std::string("serve launch: ") + hipGetErrorString(e);
                complete(slots[L[t].slot]);
                cv_done.notify_all();
                lk.unlock();
            }
        }
    }
}

extern "C" {

int pimc_server_create(const pimc_server_desc* d, pimc_server** out) {
    if (!d || !out) return PIMC_E_ARG;
    *out = nullptr;
    std::unique_ptr<pimc_server> s(new pimc_server());
    s->desc = *d;
    if (!aziz_params(d->aziz_year, &s->params)) return PIMC_E_ARG;
    if (d->v_mode < 0 || d->v_mode > 2 || d->hbm_budget < 4096) return PIMC_E_ARG;
    s->n_slots = std::max(1, std::min(kSlots, d->streams <= 0 ? 2 : d->streams));
    s->max_table_sep = d->max_side > 0 ? std::sqrt(3.0) * d->max_side / 2.0 : 200.0;
    s->az = make_aziz_dev(s->params, s->max_table_sep);
    if (hipMalloc(&s->arena.base, d->hbm_budget) != hipSuccess) return PIMC_E_NOMEM;
    s->arena.size = d->hbm_budget;
    s->arena.free_[0] = d->hbm_budget;
    for (int k = 0; k < kSlots; ++k) {
        Slot& S = s->slots[k];
        if (hipStreamCreateWithFlags(&S.stream, hipStreamNonBlocking) != hipSuccess) return PIMC_E_HIP;
        if (hipEventCreateWithFlags(&S.ev, hipEventDisableTiming) != hipSuccess) return PIMC_E_HIP;
        S.desc_bytes = sizeof(ServeReqDev) * kMaxReqPerBatch + sizeof(VnnBead) * kMaxBeadsPerBatch +
                       sizeof(ServeGroupDev) * kMaxGroupsPerBatch + 1024;
        if (hipHostMalloc((void**)&S.h_desc, S.desc_bytes, hipHostMallocMapped | hipHostMallocCoherent) !=
            hipSuccess)
            return PIMC_E_NOMEM;
        if (hipHostGetDevicePointer((void**)&S.m_desc, S.h_desc, 0) != hipSuccess) return PIMC_E_HIP;
        if (hipMalloc(&S.d_desc, S.desc_bytes) != hipSuccess) return PIMC_E_NOMEM;
        if (hipMalloc(&S.d_partials, sizeof(double) * 2 * kMaxGroupsPerBatch) != hipSuccess) return PIMC_E_NOMEM;
        if (hipHostMalloc((void**)&S.h_out, sizeof(ServeOut) * kMaxReqPerBatch,
                          hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess)
            return PIMC_E_NOMEM;
        std::memset(S.h_out, 0, sizeof(ServeOut) * kMaxReqPerBatch);   // no record carries a batch number yet
        if (hipHostGetDevicePointer((void**)&S.d_out, S.h_out, 0) != hipSuccess) return PIMC_E_HIP;
        if (hipHostMalloc((void**)&S.h_done, 64, hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess)
            return PIMC_E_NOMEM;
        *S.h_done = 0;
        if (hipHostGetDevicePointer((void**)&S.d_done, S.h_done, 0) != hipSuccess) return PIMC_E_HIP;
        if (hipMalloc(&S.d_rho, kSsfRhoBytes) != hipSuccess) return PIMC_E_NOMEM;
        if (hipHostMalloc((void**)&S.h_sq, sizeof(double) * kMaxSsfQ, hipHostMallocMapped | hipHostMallocCoherent) !=
            hipSuccess)
            return PIMC_E_NOMEM;
        if (hipHostGetDevicePointer((void**)&S.m_sq, S.h_sq, 0) != hipSuccess) return PIMC_E_HIP;
    }
    // PIMC_DIRECT_DISPATCH=0: every batch launches through HIP
    const char* dd = std::getenv("PIMC_DIRECT_DISPATCH");
    if (!(dd && std::strcmp(dd, "0") == 0)) (void)s->setup_direct();
    pimc_server* raw = s.get();
    raw->worker = std::thread([raw] { raw->run(); });
    *out = s.release();
    return PIMC_OK;
}

void pimc_server_destroy(pimc_server* s) {
    if (!s) return;
    s->stop.store(true);
    s->cv_work.notify_all();
    if (s->worker.joinable()) s->worker.join();
    for (auto& S : s->slots) {
        direct_ring_release(&S.ring);   // waits for the slot's packets
        direct_free_visible(S.v_desc);
        if (S.stream) (void)hipStreamSynchronize(S.stream);
        if (S.ev) (void)hipEventDestroy(S.ev);
        if (S.h_desc) (void)hipHostFree(S.h_desc);
        if (S.d_desc) (void)hipFree(S.d_desc);
        if (S.d_partials) (void)hipFree(S.d_partials);
        if (S.h_out) (void)hipHostFree(S.h_out);
        if (S.h_done) (void)hipHostFree(S.h_done);
        if (S.d_rho) (void)hipFree(S.d_rho);
        if (S.h_sq) (void)hipHostFree(S.h_sq);
        if (S.stream) (void)hipStreamDestroy(S.stream);
    }
    if (s->d_chains) (void)hipFree(s->d_chains);
    if (s->arena.base) (void)hipFree(s->arena.base);
    delete s;
}

const char* pimc_server_last_error(const pimc_server* s) { return s ? s->err.c_str() : ""; }

int pimc_server_add_chain(pimc_server* s, const pimc_desc* d, const double* beads, int* chain_id) {
    if (!s || !d || !beads || !chain_id) return PIMC_E_ARG;
    if (d->num_slices < 1 || d->num_alloc < 2 || !(d->side > 0) || !(d->tau > 0) || !(d->rc > 0))
        return sfail(s, PIMC_E_ARG, "invalid chain description");
    if (d->aziz_year != s->desc.aziz_year) return sfail(s, PIMC_E_ARG, "chain Aziz year differs from the server's");
    if (std::sqrt(3.0) * d->side / 2.0 > s->max_table_sep) return sfail(s, PIMC_E_ARG, "box larger than max_side");
    std::lock_guard<std::mutex> g(s->mu);
    Chain c;
    c.d = *d;
    c.bytes = chain_bytes(*d);
    if (c.bytes > s->arena.size) return sfail(s, PIMC_E_NOMEM, "chain larger than the HBM budget");
    const int id = (int)s->chains.size();
    s->chains.push_back(c);
    // publish the chain's extents to the lock-free submit path
    ExtentsTable* t = s->extents.load(std::memory_order_relaxed);
    if (!t || t->n.load(std::memory_order_relaxed) == t->cap) {
        auto nt = std::make_unique<ExtentsTable>();
        nt->cap = t ? 2 * t->cap : 256;
        nt->v.reset(new ChainExtents[nt->cap]);
        const int n = t ? t->n.load(std::memory_order_relaxed) : 0;
        for (int k = 0; k < n; ++k) nt->v[k] = t->v[k];
        nt->n.store(n, std::memory_order_relaxed);
        t = nt.get();
        s->extents_all.push_back(std::move(nt));
        s->extents.store(t, std::memory_order_release);
    }
    const int n_ext = t->n.load(std::memory_order_relaxed);
    t->v[n_ext] = ChainExtents{d->num_slices, d->num_alloc};
    t->n.store(n_ext + 1, std::memory_order_release);
    const bool queue_empty = s->waiting_chains.empty();
    int r = queue_empty ? place_chain(s, id, beads) : 1;
    if (r < 0) return sfail(s, PIMC_E_HIP, "chain upload failed");
    if (r > 0) {
        Chain& cc = s->chains[id];
        cc.pending_beads.assign(beads, beads + 3 * (size_t)d->num_slices * d->num_alloc);
        s->waiting_chains.push_back(id);
    }
    *chain_id = id;
    return PIMC_OK;
}

int pimc_server_chain_state(pimc_server* s, int chain, int* state) {
    if (!s || !state) return PIMC_E_ARG;
    std::lock_guard<std::mutex> g(s->mu);
    if (chain < 0 || chain >= (int)s->chains.size()) return PIMC_E_ARG;
    *state = s->chains[chain].state;
    return PIMC_OK;
}

int pimc_server_remove_chain(pimc_server* s, int chain) {
    if (!s) return PIMC_E_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    if (chain < 0 || chain >= (int)s->chains.size()) return PIMC_E_ARG;
    Chain& c = s->chains[chain];
    s->cv_done.wait(lk, [&] { return !c.busy && c.queue.empty(); });
    if (c.state == 1) {
        s->arena.release(c.off, c.bytes);
        s->chain_table[c.dev_index].pos = nullptr;
        s->chains_dirty = true;
    }
    c.state = 2;
    s->admit_waiting();
    s->work_events.fetch_add(1, std::memory_order_release);
    s->cv_work.notify_all();
    return PIMC_OK;
}

int pimc_server_download(pimc_server* s, int chain, double* beads) {
    if (!s || !beads) return PIMC_E_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    if (chain < 0 || chain >= (int)s->chains.size()) return PIMC_E_ARG;
    Chain& c = s->chains[chain];
    s->cv_done.wait(lk, [&] { return !c.busy && c.queue.empty(); });
    const size_t MN = (size_t)c.d.num_slices * c.d.num_alloc;
    if (c.state == 0) {
        std::memcpy(beads, c.pending_beads.data(), sizeof(double) * 3 * MN);
        return PIMC_OK;
    }
    if (c.state != 1) return sfail(s, PIMC_E_STATE, "chain removed");
    std::vector<double> soa(3 * MN);
    if (hipMemcpy(soa.data(), c.dev.pos, sizeof(double) * 3 * MN, hipMemcpyDeviceToHost) != hipSuccess)
        return PIMC_E_HIP;
    for (size_t t = 0; t < MN; ++t) {
        beads[3 * t] = soa[t];
        beads[3 * t + 1] = soa[t + MN];
        beads[3 * t + 2] = soa[t + 2 * MN];
    }
    return PIMC_OK;
}

}  // extern "C"

// Queues a built request behind its chain's earlier ones (the chain must not
// have been removed); the worker sees it at once when the chain is idle.  The
// server lock is held by the caller.
static int enqueue_locked(pimc_server* s, Request&& R, uint64_t* req_id) {
    if (s->kernel_error) return sfail(s, s->kernel_error, s->kernel_error_msg);
    const int chain = R.chain;
    if (chain < 0 || chain >= (int)s->chains.size() || s->chains[chain].state == 2)
        return sfail(s, PIMC_E_ARG, "unknown chain");
    Chain& c = s->chains[chain];
    R.id = s->next_id++;
    R.t_submit = Clock::now();
    const uint64_t id = R.id;
    s->reqs.emplace(id, std::move(R));
    c.queue.push_back(id);
    if (c.queue.size() == 1 && c.state == 1 && !c.busy) s->ready.push_back(id);
    *req_id = id;
    return PIMC_OK;
}

// The submit counter moves under the lock: the worker reads it under the lock
// before it blocks, so a bump cannot fall between its check and its wait.
static int server_enqueue(pimc_server* s, Request&& R, uint64_t* req_id) {
    std::lock_guard<std::mutex> g(s->mu);
    const int r = enqueue_locked(s, std::move(R), req_id);
    if (r) return r;
    s->work_events.fetch_add(1, std::memory_order_release);
    s->cv_work.notify_one();
    return PIMC_OK;
}

// Several built requests under one acquisition of the server lock (a client
// holding many chains resubmits after each completion list); ids[k] receives
// request k's id.  Stops at the first refused request and returns its code.
static int server_enqueue_many(pimc_server* s, std::vector<Request>& rs, uint64_t* ids) {
    int r = PIMC_OK;
    std::lock_guard<std::mutex> g(s->mu);
    for (size_t k = 0; k < rs.size() && !r; ++k) r = enqueue_locked(s, std::move(rs[k]), &ids[k]);
    s->work_events.fetch_add(1, std::memory_order_release);
    s->cv_work.notify_one();
    return r;
}

// Checks and builds a Delta V request (without the server lock).
static int server_build(pimc_server* s, int chain, const int32_t* slice_ptcl, const double* new_pos, int n,
                        int shift, const double vfactor[2], int commit_mode, double u, int tag, Request* out) {
    if (!s || !slice_ptcl || !new_pos || n < 1 || !out || shift < 1 || commit_mode < 0 || commit_mode > 2)
        return PIMC_E_ARG;
    if (n > kMaxBeadsPerBatch / 4) return sfail(s, PIMC_E_ARG, "too many beads in one request");
    // the chain's extents from the lock-free table (the state and a kernel error
    // are checked when the request is queued); the request is built outside the
    // lock, so the server lock is held only to queue it
    int num_slices = 0, num_alloc = 0;
    {
        const ExtentsTable* t = s->extents.load(std::memory_order_acquire);
        if (chain < 0 || !t || chain >= t->n.load(std::memory_order_acquire)) {
            std::lock_guard<std::mutex> g(s->mu);
            return sfail(s, PIMC_E_ARG, "unknown chain");
        }
        num_slices = t->v[chain].num_slices;
        num_alloc = t->v[chain].num_alloc;
    }
    std::vector<char> seen(num_slices, 0);
    Request R;
    R.chain = chain;
    R.shift = shift;
    R.vf[0] = vfactor ? vfactor[0] : 1.0;
    R.vf[1] = vfactor ? vfactor[1] : 1.0;
    R.commit_mode = commit_mode;
    R.u = u;
    R.beads.resize(n);
    for (int k = 0; k < n; ++k) {
        const int m = slice_ptcl[2 * k], i = slice_ptcl[2 * k + 1];
        if (m < 0 || m >= num_slices || i < 0 || i >= num_alloc) {
            std::lock_guard<std::mutex> g(s->mu);
            return sfail(s, PIMC_E_ARG, "bead locator out of range");
        }
        if (seen[m]) {
            std::lock_guard<std::mutex> g(s->mu);
            return sfail(s, PIMC_E_ARG, "request beads must sit on distinct slices");
        }
        seen[m] = 1;
        VnnBead& b = R.beads[k];
        b.walker = chain;
        b.slice = m;
        b.ptcl = i;
        b.flags = 2;
        for (int t = 0; t < 3; ++t) {
            b.old_pos[t] = 0.0;
            b.new_pos[t] = new_pos[3 * k + t];
        }
    }
    R.tag = tag;
    *out = std::move(R);
    return PIMC_OK;
}

static int server_submit(pimc_server* s, int chain, const int32_t* slice_ptcl, const double* new_pos, int n,
                         int shift, const double vfactor[2], int commit_mode, double u, int tag, uint64_t* req_id) {
    if (!req_id) return PIMC_E_ARG;
    Request R;
    const int r = server_build(s, chain, slice_ptcl, new_pos, n, shift, vfactor, commit_mode, u, tag, &R);
    if (r) return r;
    return server_enqueue(s, std::move(R), req_id);
}

// S(q) of a chain's current worldlines for the lattice vectors qn [nq][3]
// (q = 2 pi n / L), queued behind the chain's earlier requests.
static int server_build_ssf(pimc_server* s, int chain, std::vector<int>&& qn, int tag, Request* out) {
    const int nq = (int)qn.size() / 3;
    if (nq < 1 || nq > kMaxSsfQ) return sfail(s, PIMC_E_ARG, "S(q) requests take 1..65536 q-vectors");
    for (int v : qn)
        if (std::abs(v) > rho_max_nmax()) return sfail(s, PIMC_E_UNSUPPORTED, "lattice component beyond 40");
    Request R;
    R.kind = 1;
    R.chain = chain;
    R.qn = std::move(qn);
    R.tag = tag;
    *out = std::move(R);
    return PIMC_OK;
}

static int server_submit_ssf(pimc_server* s, int chain, std::vector<int>&& qn, int tag, uint64_t* req_id) {
    Request R;
    const int r = server_build_ssf(s, chain, std::move(qn), tag, &R);
    if (r) return r;
    return server_enqueue(s, std::move(R), req_id);
}

// The results of the ids in `ids` (all completed) in one critical section;
// the requests' storage is released after the lock.
static int server_take(pimc_server* s, const std::vector<uint64_t>& ids, std::vector<double>& out5) {
    std::vector<Request> gone;
    gone.reserve(ids.size());
    out5.resize(5 * ids.size());
    {
        std::lock_guard<std::mutex> g(s->mu);
        for (size_t k = 0; k < ids.size(); ++k) {
            auto it = s->reqs.find(ids[k]);
            if (it == s->reqs.end() || !it->second.done) return sfail(s, PIMC_E_STATE, "result not ready");
            for (int t = 0; t < 4; ++t) out5[5 * k + t] = it->second.out[t];
            out5[5 * k + 4] = std::chrono::duration<double, std::micro>(it->second.t_done - it->second.t_submit).count();
            gone.push_back(std::move(it->second));
            s->reqs.erase(it);
        }
    }
    return PIMC_OK;
}

extern "C" {

int pimc_server_submit(pimc_server* s, int chain, const int32_t* slice_ptcl, const double* new_pos, int n,
                       int shift, const double vfactor[2], int commit_mode, double u, uint64_t* req_id) {
    return server_submit(s, chain, slice_ptcl, new_pos, n, shift, vfactor, commit_mode, u, -1, req_id);
}

int pimc_server_submit_ssf(pimc_server* s, int chain, const double* q, int nq, uint64_t* req_id) {
    if (!s || !q || !req_id || nq < 1) return PIMC_E_ARG;
    double side = 0.0;
    {
        std::lock_guard<std::mutex> g(s->mu);
        if (s->kernel_error) return sfail(s, s->kernel_error, s->kernel_error_msg);
        if (chain < 0 || chain >= (int)s->chains.size() || s->chains[chain].state == 2)
            return sfail(s, PIMC_E_ARG, "unknown chain");
        side = s->chains[chain].d.side;
    }
    if (nq > kMaxSsfQ) return sfail(s, PIMC_E_ARG, "S(q) requests take 1..65536 q-vectors");
    std::vector<int> qn(3 * (size_t)nq);
    for (int k = 0; k < nq; ++k)
        if (!lattice_q(q + 3 * k, side, &qn[3 * (size_t)k]))
            return sfail(s, PIMC_E_UNSUPPORTED, "S(q) requests take lattice q = 2 pi n / L, |n_d| <= 40");
    return server_submit_ssf(s, chain, std::move(qn), -1, req_id);
}

int pimc_server_wait_ssf(pimc_server* s, uint64_t req_id, double* sq, double* latency_us) {
    if (!s || !sq) return PIMC_E_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(req_id);
    if (it == s->reqs.end()) return sfail(s, PIMC_E_ARG, "unknown request");
    Request* R = &it->second;
    if (R->kind != 1) return sfail(s, PIMC_E_ARG, "not an S(q) request: use pimc_server_wait");
    s->cv_done.wait(lk, [&] { return R->done || s->kernel_error; });
    if (!R->done) return sfail(s, s->kernel_error, s->kernel_error_msg);
    std::copy(R->sq.begin(), R->sq.end(), sq);
    if (latency_us) *latency_us = std::chrono::duration<double, std::micro>(R->t_done - R->t_submit).count();
    Request gone = std::move(*R);
    s->reqs.erase(req_id);
    lk.unlock();
    return PIMC_OK;
}

// out5 = {A_old, A_new, dU, accepted, latency_us}; releases the request.
int pimc_server_wait(pimc_server* s, uint64_t req_id, double* out5) {
    if (!s) return PIMC_E_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(req_id);
    if (it == s->reqs.end()) return sfail(s, PIMC_E_ARG, "unknown request");
    if (it->second.kind != 0) return sfail(s, PIMC_E_ARG, "an S(q) request: use pimc_server_wait_ssf");
    // an element pointer survives rehashes by concurrent submits; an iterator does not
    Request* R = &it->second;
    s->cv_done.wait(lk, [&] { return R->done || s->kernel_error; });
    if (!R->done) return sfail(s, s->kernel_error, s->kernel_error_msg);
    if (out5) {
        for (int t = 0; t < 4; ++t) out5[t] = R->out[t];
        out5[4] = std::chrono::duration<double, std::micro>(R->t_done - R->t_submit).count();
    }
    Request gone = std::move(*R);   // its storage is released after the lock
    s->reqs.erase(req_id);
    lk.unlock();
    return PIMC_OK;
}

int pimc_server_poll(pimc_server* s, uint64_t req_id, int* done) {
    if (!s || !done) return PIMC_E_ARG;
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->reqs.find(req_id);
    if (it == s->reqs.end()) return sfail(s, PIMC_E_ARG, "unknown request");
    *done = it->second.done ? 1 : 0;
    return PIMC_OK;
}

int pimc_server_stats(pimc_server* s, int64_t* batches, int64_t* requests, int64_t* pairs, size_t* free_bytes) {
    if (!s) return PIMC_E_ARG;
    std::lock_guard<std::mutex> g(s->mu);
    if (batches) *batches = s->launched_batches;
    if (requests) *requests = s->launched_reqs;
    if (pairs) *pairs = s->pairs_done;
    if (free_bytes) *free_bytes = s->arena.free_bytes();
    return PIMC_OK;
}

int pimc_server_dispatch_stats(pimc_server* s, int64_t* packet_batches, int64_t* hip_batches) {
    if (!s) return PIMC_E_ARG;
    if (packet_batches) *packet_batches = s->packet_batches.load(std::memory_order_relaxed);
    if (hip_batches) *hip_batches = s->hip_batches.load(std::memory_order_relaxed);
    return PIMC_OK;
}

}  // extern "C"

// ===========================================================================
// Synthetic load: Poisson arrivals of Delta V requests over the resident chains.
// T generator threads, thread t owning the chains c = t (mod T), each with its
// own random stream, arrival process (rate / T) and completion list, so
// submissions are not limited by one host thread.
// ===========================================================================
namespace {

struct GenChain {
    pimc_desc d{};                     // the chain's description, copied under the server lock
    std::vector<double> mirror;        // [M][N][3], kept equal to the device copy
    bool inflight_ssf = false;         // the request in flight is an S(q)
    std::deque<Clock::time_point> arrivals;
    uint64_t inflight = 0;
    Clock::time_point t_arrival;
    std::vector<int32_t> sp;
    std::vector<double> prop;
};

struct GenStats {
    int64_t completed = 0, accepted = 0, pairs = 0;
    std::vector<double> lat;
    std::vector<double> svc;   // submit -> result retired by the worker
    int64_t ssf_completed = 0;
    double ssf_phasors = 0.0;  // (q, bead) density-mode terms of the completed S(q) requests
    std::vector<double> ssf_lat;
    int err = 0;
};

void gen_thread(pimc_server* s, const pimc_load_desc* L, int t, int T, std::vector<GenChain>& gc,
                const std::vector<int>& ssf_qn, Clock::time_point t0, Clock::time_point t_end, GenStats& st) {
    std::mt19937_64 rng(L->seed + 7919ull * (uint64_t)t);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    std::normal_distribution<double> G(0.0, 1.0);
    const bool closed = L->rate <= 0;
    std::exponential_distribution<double> Exp(closed ? 1.0 : L->rate / T);
    std::vector<int> mine;
    for (int c = t; c < (int)gc.size(); c += T) mine.push_back(c);
    if (mine.empty()) return;
    st.lat.reserve(1 << 18);
    auto next_arrival = t0;
    int inflight = 0;
    bool drained = false;
    std::unordered_map<uint64_t, int> by_id;
    std::deque<int> runnable;   // own chains with an arrival waiting and nothing in flight
    auto propose = [&](int c) {
        GenChain& g = gc[c];
        const pimc_desc& d = g.d;
        const int M = d.num_slices, N = d.num_alloc;
        g.sp.clear();
        g.prop.clear();
        const int p = (int)(U(rng) * N) % N;
        if (U(rng) < L->p_com) {
            // CenterOfMass: the whole worldline of p shifted by one Gaussian vector
            double sh[3];
            for (double& x : sh) x = L->com_sigma * G(rng);
            for (int m = 0; m < M; ++m) {
                g.sp.push_back(m);
                g.sp.push_back(p);
                for (int k = 0; k < 3; ++k) {
                    double x = g.mirror[3 * ((size_t)m * N + p) + k] + sh[k];
                    x -= d.side * std::floor(x * (1.0 / d.side) + 0.5);
                    g.prop.push_back(x);
                }
            }
        } else {
            int seg = L->seg_min + (int)(U(rng) * (L->seg_max - L->seg_min + 1));
            seg = std::min(seg, M);
            const int m0 = (int)(U(rng) * M) % M;
            const double sd = std::sqrt(L->lambda * d.tau);
            for (int k = 1; k < seg; ++k) {
                const int m = (m0 + k) % M;
                g.sp.push_back(m);
                g.sp.push_back(p);
                for (int a = 0; a < 3; ++a) {
                    double x = g.mirror[3 * ((size_t)m * N + p) + a] + sd * G(rng);
                    x -= d.side * std::floor(x * (1.0 / d.side) + 0.5);
                    g.prop.push_back(x);
                }
            }
        }
    };
    // a pass's new requests are built without the lock and queued under one
    // acquisition of it
    std::vector<std::pair<int, Clock::time_point>> batch;
    std::vector<Request> built;
    std::vector<uint64_t> ids;
    auto submit_batch = [&]() -> int {
        built.resize(batch.size());
        for (size_t k = 0; k < batch.size(); ++k) {
            const int c = batch[k].first;
            GenChain& g = gc[c];
            int r;
            g.inflight_ssf = !ssf_qn.empty() && U(rng) < L->p_ssf;
            if (g.inflight_ssf) {
                r = server_build_ssf(s, c, std::vector<int>(ssf_qn), t, &built[k]);
            } else {
                propose(c);
                const double vf[2] = {1.0, 1.0};
                r = server_build(s, c, g.sp.data(), g.prop.data(), (int)g.sp.size() / 2, 1, vf, 1, U(rng), t,
                                 &built[k]);
            }
            if (r) return r;
        }
        ids.assign(batch.size(), 0);
        const int r = server_enqueue_many(s, built, ids.data());
        if (r) return r;
        for (size_t k = 0; k < batch.size(); ++k) {
            GenChain& g = gc[batch[k].first];
            g.inflight = ids[k];
            g.t_arrival = batch[k].second;
            by_id[ids[k]] = batch[k].first;
            ++inflight;
        }
        batch.clear();
        built.clear();
        return 0;
    };
    if (closed) {
        for (int c : mine) batch.emplace_back(c, Clock::now());
        if ((st.err = submit_batch())) return;
    }
    std::vector<uint64_t> finished;
    std::vector<double> outs;
    uint64_t seen_done = 0;
    unsigned polls = 0;
    while (true) {
        const auto now = Clock::now();
        const bool open = now < t_end;
        if (!closed && open) {
            while (next_arrival <= now) {
                const int c = mine[(size_t)(U(rng) * mine.size()) % mine.size()];
                gc[c].arrivals.push_back(next_arrival);
                if (!gc[c].inflight && gc[c].arrivals.size() == 1) runnable.push_back(c);
                next_arrival += std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(Exp(rng)));
            }
        }
        // completions of this thread's requests, in the order the server finished them
        finished.clear();
        // the completion list is taken under the lock only when the worker has
        // added to it (or every 1024th pass, to see a kernel error)
        const uint64_t dc = s->done_counts[t].load(std::memory_order_acquire);
        if (dc != seen_done || (++polls & 1023u) == 0) {
            seen_done = dc;
            std::lock_guard<std::mutex> g(s->mu);
            finished.swap(s->done_lists[t]);
            if (s->kernel_error) {
                st.err = sfail(s, s->kernel_error, s->kernel_error_msg);
                return;
            }
        }
        if (!finished.empty() && (st.err = server_take(s, finished, outs))) return;
        for (size_t f = 0; f < finished.size(); ++f) {
            const uint64_t id = finished[f];
            auto it = by_id.find(id);
            if (it == by_id.end()) continue;
            const int c = it->second;
            by_id.erase(it);
            GenChain& g = gc[c];
            const double* o = outs.data() + 5 * f;
            const auto t_done = Clock::now();
            const pimc_desc& d = g.d;
            if (g.inflight_ssf) {
                st.ssf_lat.push_back(std::chrono::duration<double, std::micro>(t_done - g.t_arrival).count());
                ++st.ssf_completed;
                st.ssf_phasors += (double)(ssf_qn.size() / 3) * d.num_slices * d.num_alloc;
            } else {
                st.lat.push_back(std::chrono::duration<double, std::micro>(t_done - g.t_arrival).count());
                st.svc.push_back(o[4]);
                ++st.completed;
                st.pairs += 2LL * ((int64_t)g.sp.size() / 2) * (d.num_alloc - 1);
            }
            if (!g.inflight_ssf && o[3] > 0.5) {
                ++st.accepted;
                for (size_t k = 0; k < g.sp.size() / 2; ++k) {
                    const size_t bi = (size_t)g.sp[2 * k] * d.num_alloc + g.sp[2 * k + 1];
                    for (int a = 0; a < 3; ++a) g.mirror[3 * bi + a] = g.prop[3 * k + a];
                }
            }
            g.inflight = 0;
            --inflight;
            if (!closed && !g.arrivals.empty()) runnable.push_back(c);
            if (closed && open) batch.emplace_back(c, Clock::now());
        }
        while (!runnable.empty()) {
            const int c = runnable.front();
            runnable.pop_front();
            GenChain& g = gc[c];
            if (!g.inflight && !g.arrivals.empty()) {
                const auto ta = g.arrivals.front();
                g.arrivals.pop_front();
                batch.emplace_back(c, ta);
                g.inflight = ~0ull;   // claimed by this pass's batch
            }
        }
        if (!batch.empty() && (st.err = submit_batch())) return;
        if (finished.empty()) std::this_thread::yield();
        if (!open && inflight == 0) break;
        if (!open && now - t_end > std::chrono::seconds(30)) {
            st.err = sfail(s, PIMC_E_STATE, "synthetic load: requests still pending 30 s after the load window");
            return;
        }
        if (!open && !drained) {
            // stop accepting arrivals; drop the ones not yet submitted
            for (int c : mine) gc[c].arrivals.clear();
            runnable.clear();
            drained = true;
        }
    }
}

}  // namespace

extern "C" int pimc_server_run_synthetic(pimc_server* s, const pimc_load_desc* L, double* stats) {
    if (!s || !L || !stats || L->n_chains < 1 || L->duration_s <= 0 || L->seg_min < 2 || L->seg_max < L->seg_min ||
        L->threads < 0 || L->threads > 64 || !(L->p_ssf >= 0.0 && L->p_ssf <= 1.0) ||
        (L->p_ssf > 0.0 && (L->ssf_nmax < 0 || L->ssf_nmax > 20)))
        return PIMC_E_ARG;
    const int nc = L->n_chains;
    const int T = std::max(1, std::min(L->threads, nc));
    std::vector<GenChain> gc(nc);
    {
        std::lock_guard<std::mutex> g(s->mu);
        if (nc > (int)s->chains.size()) return sfail(s, PIMC_E_ARG, "more load chains than server chains");
        s->done_lists.assign(T, {});
        for (int t = 0; t < T; ++t) s->done_counts[t].store(0, std::memory_order_relaxed);
        s->track_done = true;
    }
    struct TrackOff {
        pimc_server* s;
        ~TrackOff() {
            std::lock_guard<std::mutex> g(s->mu);
            s->track_done = false;
            s->done_lists.clear();
        }
    } track_off{s};
    {
        std::lock_guard<std::mutex> g(s->mu);
        for (int c = 0; c < nc; ++c) gc[c].d = s->chains[c].d;
    }
    for (int c = 0; c < nc; ++c) {
        const pimc_desc& d = gc[c].d;
        gc[c].mirror.resize(3 * (size_t)d.num_slices * d.num_alloc);
        int r = pimc_server_download(s, c, gc[c].mirror.data());
        if (r) return r;
    }
    int64_t b0 = 0, r0 = 0;
    pimc_server_stats(s, &b0, &r0, nullptr, nullptr);
    const auto t0 = Clock::now();
    const auto t_end = t0 + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(L->duration_s));
    // the S(q) requests' lattice vectors: every n with |n| <= ssf_nmax, in the
    // reference's max_int order (lexicographic from -n, src/estimator.cpp:487-541)
    std::vector<int> ssf_qn;
    if (L->p_ssf > 0.0)
        for (int a = -L->ssf_nmax; a <= L->ssf_nmax; ++a)
            for (int b = -L->ssf_nmax; b <= L->ssf_nmax; ++b)
                for (int c = -L->ssf_nmax; c <= L->ssf_nmax; ++c)
                    if (a * a + b * b + c * c <= L->ssf_nmax * L->ssf_nmax) {
                        ssf_qn.push_back(a);
                        ssf_qn.push_back(b);
                        ssf_qn.push_back(c);
                    }
    std::vector<GenStats> st(T);
    std::vector<std::thread> th;
    for (int t = 0; t < T; ++t)
        th.emplace_back(gen_thread, s, L, t, T, std::ref(gc), std::cref(ssf_qn), t0, t_end, std::ref(st[t]));
    for (auto& x : th) x.join();
    for (auto& x : st)
        if (x.err) return x.err;
    const double secs = std::chrono::duration<double>(Clock::now() - t0).count();
    int64_t b1 = 0, r1 = 0;
    pimc_server_stats(s, &b1, &r1, nullptr, nullptr);
    std::vector<double> lat, svc;
    int64_t completed = 0, accepted = 0, pairs = 0;
    for (auto& x : st) {
        lat.insert(lat.end(), x.lat.begin(), x.lat.end());
        svc.insert(svc.end(), x.svc.begin(), x.svc.end());
        completed += x.completed;
        accepted += x.accepted;
        pairs += x.pairs;
    }
    std::sort(lat.begin(), lat.end());
    auto pct = [&](double q) { return lat.empty() ? 0.0 : lat[std::min(lat.size() - 1, (size_t)(q * (lat.size() - 1)))]; };
    double mean = 0;
    for (double x : lat) mean += x;
    mean = lat.empty() ? 0 : mean / lat.size();
    stats[0] = (double)completed;
    stats[1] = completed / secs;
    stats[2] = pairs / secs;
    stats[3] = pct(0.50);
    stats[4] = pct(0.99);
    stats[5] = mean;
    stats[6] = completed ? (double)accepted / completed : 0.0;
    stats[7] = (double)(b1 - b0);
    stats[8] = (b1 > b0) ? (double)(r1 - r0) / (b1 - b0) : 0.0;
    stats[9] = secs;
    std::sort(svc.begin(), svc.end());
    auto spct = [&](double q) { return svc.empty() ? 0.0 : svc[std::min(svc.size() - 1, (size_t)(q * (svc.size() - 1)))]; };
    stats[10] = spct(0.50);
    stats[11] = spct(0.99);
    std::vector<double> sl;
    int64_t ssf_done = 0;
    double phasors = 0.0;
    for (auto& x : st) {
        sl.insert(sl.end(), x.ssf_lat.begin(), x.ssf_lat.end());
        ssf_done += x.ssf_completed;
        phasors += x.ssf_phasors;
    }
    std::sort(sl.begin(), sl.end());
    auto qpct = [&](double q) { return sl.empty() ? 0.0 : sl[std::min(sl.size() - 1, (size_t)(q * (sl.size() - 1)))]; };
    stats[12] = (double)ssf_done;
    stats[13] = qpct(0.50);
    stats[14] = qpct(0.99);
    stats[15] = phasors / secs;
    return PIMC_OK;
}
