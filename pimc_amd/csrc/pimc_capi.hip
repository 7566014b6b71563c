// pimc_capi.hip -- host side of the C-ABI declared in include/pimc_gpu.h.
//
// A pimc_ctx owns one chain's worldlines in HBM (FP64 SoA planes x|y|z of
// [M][N_alloc]), the bead mask, beads-per-slice, the bit-exact cell lists and
// the pair work list, plus one HIP stream (default priority) and pinned
// staging buffers.  Every entry point validates its arguments on the host,
// launches on the context's stream and synchronises before returning.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/pimc_gpu.h"
#include "pimc_kernels.hpp"


#include "pimc_host.hpp"

using namespace pimc;

// ===========================================================================
// errors
// ===========================================================================
static thread_local std::string g_last_error;

static int fail(pimc_ctx* ctx, int code, const std::string& msg) {
    g_last_error = msg;
    if (ctx) {
        ctx->err = msg;
        ++ctx->fail_count;
    }
    return code;
}

#define HIP_OR_FAIL(ctx, expr)                                                              \
    do {                                                                                    \
        hipError_t _e = (expr);                                                             \
        if (_e != hipSuccess)                                                               \
            return fail((ctx), PIMC_E_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

// ===========================================================================
// Aziz constants and tables (src/potential.cpp:1741-1909)
// ===========================================================================
namespace pimc {

bool aziz_params(int year, AzizParams* p) {
    if (year == 1979) {
        *p = {10.8, 2.9673, 1.241314, 13.353384, 0.0, 1.3732412, 0.4253785, 0.1781, 0.5448504E6};
    } else if (year == 1987) {
        *p = {10.948, 2.9673, 1.4826, 10.43329537, -2.27965105, 1.36745214, 0.42123807, 0.17473318, 1.8443101E5};
    } else if (year == 1995) {
        *p = {10.956, 2.9683, 1.438, 10.5717543, -2.07758779, 1.35186623, 0.4149514, 0.17151143, 1.86924404E5};
    } else {
        return false;
    }
    return true;
}

// valueV / valuedVdr / valued2Vdr2 of the reference, FP64, in its expression order.
#pragma STDC FP_CONTRACT OFF
static double aziz_F(const AzizParams& p, double x) {
    return (x < p.D ? std::exp(-(p.D / x - 1.0) * (p.D / x - 1.0)) : 1.0);
}
static double aziz_dF(const AzizParams& p, double x) {
    double ix = 1.0 / x;
    double r = 2.0 * p.D * ix * ix * (p.D * ix - 1.0) * std::exp(-(p.D * ix - 1.0) * (p.D * ix - 1.0));
    return (x < p.D ? r : 0.0);
}
static double aziz_d2F(const AzizParams& p, double x) {
    double ix = 1.0 / x;
    const double D = p.D;
    double r = 2.0 * D * ix * ix * ix * (2.0 * D * D * D * ix * ix * ix - 4.0 * D * D * ix * ix - D * ix + 2.0) *
               std::exp(-(D * ix - 1.0) * (D * ix - 1.0));
    return (x < D ? r : 0.0);
}
double aziz_value(const AzizParams& p, double r, int what) {
    double x = r / p.rm;
    if (what == 0) {
        double Urep = p.A * std::exp(-p.alpha * x + p.beta * x * x);
        if (x < 1.0e-7) return 0.0;
        if (x < 0.01) return (p.eps * Urep);
        double ix2 = 1.0 / (x * x);
        double ix6 = ix2 * ix2 * ix2;
        double ix8 = ix6 * ix2;
        double ix10 = ix8 * ix2;
        double Uatt = -(p.C6 * ix6 + p.C8 * ix8 + p.C10 * ix10) * aziz_F(p, x);
        return (p.eps * (Urep + Uatt));
    }
    if (what == 1) {
        double T1 = p.A * (-p.alpha + 2.0 * p.beta * x) * std::exp(-p.alpha * x + p.beta * x * x);
        if (x < 1.0e-7) return 0.0;
        if (x < 0.01) return ((p.eps / p.rm) * T1);
        double ix = 1.0 / x;
        double ix2 = ix * ix;
        double ix6 = ix2 * ix2 * ix2;
        double ix7 = ix6 * ix;
        double ix8 = ix6 * ix2;
        double ix9 = ix8 * ix;
        double ix10 = ix8 * ix2;
        double ix11 = ix10 * ix;
        double T2 = (6.0 * p.C6 * ix7 + 8.0 * p.C8 * ix9 + 10.0 * p.C10 * ix11) * aziz_F(p, x);
        double T3 = -(p.C6 * ix6 + p.C8 * ix8 + p.C10 * ix10) * aziz_dF(p, x);
        return ((p.eps / p.rm) * (T1 + T2 + T3));
    }
    double ab = (p.alpha - 2.0 * p.beta * x) * (p.alpha - 2.0 * p.beta * x);
    double T1 = p.A * (2 * p.beta + ab) * std::exp(-p.alpha * x + p.beta * x * x);
    if (x < 1.0e-7) return 0.0;
    if (x < 0.01) return ((p.eps / p.rm) * T1);
    double ix = 1.0 / x;
    double ix2 = ix * ix;
    double ix6 = ix2 * ix2 * ix2;
    double ix7 = ix6 * ix;
    double ix8 = ix6 * ix2;
    double ix9 = ix8 * ix;
    double ix10 = ix8 * ix2;
    double ix11 = ix10 * ix;
    double ix12 = ix11 * ix;
    double T2 = -(42.0 * p.C6 * ix8 + 72.0 * p.C8 * ix10 + 110.0 * p.C10 * ix12) * aziz_F(p, x);
    double T3 = 2.0 * (6.0 * p.C6 * ix7 + 8.0 * p.C8 * ix9 + 10.0 * p.C10 * ix11) * aziz_dF(p, x);
    double T4 = -(p.C6 * ix6 + p.C8 * ix8 + p.C10 * ix10) * aziz_d2F(p, x);
    return ((p.eps / (p.rm * p.rm)) * (T1 + T2 + T3 + T4));
}

AzizDev make_aziz_dev(const AzizParams& p, double max_sep) {
    AzizDev a{};
    a.dr = (1.0E-6) * p.rm;
    a.table_length = int(max_sep / a.dr);
    a.s = a.dr / p.rm;
    // branch bins from the accumulated radius r_n (include/potential.h:176-183)
    a.k_eps = a.k_hc = a.k_d = a.table_length;
    double r = 0.0;
    bool f0 = false, f1 = false, f2 = false;
    for (int n = 0; n < a.table_length; ++n) {
        const double x = r / p.rm;
        if (!f0 && !(x < 1.0e-7)) { a.k_eps = n; f0 = true; }
        if (!f1 && !(x < 0.01)) { a.k_hc = n; f1 = true; }
        if (!f2 && !(x < p.D)) { a.k_d = n; f2 = true; }
        if (f0 && f1 && f2) break;
        r += a.dr;
    }
    // k_far: the first bin from which on eps A exp(-alpha x + beta x^2) is below
    // 2^-53 of the dispersion.  The ratio falls with x beyond x = 10/alpha < D,
    // so a bisection over [k_d, table_length] finds it (table_length: none).
    auto far = [&](int k) {
        const long double x = (long double)k * a.s, x2 = 1.0L / (x * x), x6 = x2 * x2 * x2;
        const long double rep = (long double)p.A * std::exp(-(long double)p.alpha * x + (long double)p.beta * x * x);
        const long double disp = x6 * ((long double)p.C6 + x2 * ((long double)p.C8 + x2 * (long double)p.C10));
        return rep < std::ldexp(disp, -53);
    };
    int lo = a.k_d, hi = a.table_length;
    while (lo < hi) {
        const int mid = lo + (hi - lo) / 2;
        if (far(mid))
            hi = mid;
        else
            lo = mid + 1;
    }
    a.k_far = lo;
    const double log2e = 1.4426950408889634074;
    a.rep_c0 = std::log2(p.eps * p.A);
    a.rep_c1 = -p.alpha * a.s * log2e;
    a.rep_c2 = p.beta * a.s * a.s * log2e;
    a.eC6d = p.eps * p.C6;
    a.eC8d = p.eps * p.C8;
    a.eC10d = p.eps * p.C10;
    const double s2 = a.s * a.s, s6 = s2 * s2 * s2;
    a.eC6k = a.eC6d / s6;
    a.eC8k = a.eC8d / (s6 * s2);
    a.eC10k = a.eC10d / (s6 * s2 * s2);
    a.Ds = p.D / a.s;
    a.eps = p.eps;
    a.A = p.A;
    a.alpha = p.alpha;
    a.beta = p.beta;
    a.D = p.D;
    a.C6 = p.C6;
    a.C8 = p.C8;
    a.C10 = p.C10;
    a.eps_rm = p.eps / p.rm;
    a.inv_rm = 1.0 / p.rm;
    a.table = nullptr;
    a.dtable = nullptr;
    a.d2table = nullptr;
    return a;
}

// initLookupTable: entry n = value at the n-fold accumulated radius.
void build_table(const AzizParams& p, int length, int what, std::vector<double>& out) {
    out.assign(length, 0.0);
    const double dr = (1.0E-6) * p.rm;
    double r = 0.0;
    for (int n = 0; n < length; ++n) {
        out[n] = aziz_value(p, r, what);
        r += dr;
    }
}

double aziz_tail(const AzizParams& p, double rc) {
    double rmorc = p.rm / rc;
    double t2 = p.C6 * std::pow(rmorc, 3.0) / 3.0;
    double t3 = p.C8 * std::pow(rmorc, 5.0) / 5.0;
    double t4 = p.C10 * std::pow(rmorc, 7.0) / 7.0;
    return 2.0 * M_PI * p.eps * (-p.rm * p.rm * p.rm * (t2 + t3 + t4));
}
#pragma STDC FP_CONTRACT ON

// Process-wide device tables: one per (year, derivative), grown on demand.
struct TableCache {
    std::mutex mu;
    struct Entry {
        double* d = nullptr;
        int length = 0;
    };
    std::map<std::pair<int, int>, Entry> tables;
    const double* get(int year, int what, int length, std::string* err) {
        std::lock_guard<std::mutex> g(mu);
        Entry& e = tables[{year, what}];
        if (e.d && e.length >= length) return e.d;
        AzizParams p;
        if (!aziz_params(year, &p)) {
            *err = "unknown Aziz year";
            return nullptr;
        }
        std::vector<double> h;
        build_table(p, length, what, h);
        double* d = nullptr;
        if (hipMalloc(&d, sizeof(double) * std::max(1, length)) != hipSuccess) {
            *err = "hipMalloc of the Aziz table failed";
            return nullptr;
        }
        if (hipMemcpy(d, h.data(), sizeof(double) * length, hipMemcpyHostToDevice) != hipSuccess) {
            *err = "upload of the Aziz table failed";
            (void)hipFree(d);
            return nullptr;
        }
        // Older, shorter tables stay allocated: contexts may still point at them.
        e.d = d;
        e.length = length;
        return e.d;
    }
};
TableCache& table_cache() {
    static TableCache* c = new TableCache();
    return *c;
}

const double* cached_table(int year, int what, int length, std::string* err) {
    return table_cache().get(year, what, length, err);
}

}  // namespace pimc

// ===========================================================================
// timing
// ===========================================================================
static const char* kKernelNames[] = {"pair_action", "delta_action", "rho", "ssf", "isf", "gradv2", "gradvnn",
                                     "virial", "chain_run"};

constexpr int kTimingSlots = 2048;

// The event pool is created when timing is first switched on (not inside the
// first timed call, where its ~0.5 ms would count against that call).
bool pimc_ctx::ensure_event_pool() {
    if (!ev_pool.empty()) return true;
    ev_pool.assign(2 * kTimingSlots, nullptr);
    for (auto& e : ev_pool)
        if (hipEventCreate(&e) != hipSuccess) return false;
    return true;
}

int pimc_ctx::timed_slot(int k) {
    if (!timing || !((timing_kinds >> k) & 1u)) return -1;
    if (!ensure_event_pool()) {
        timing = false;
        return -1;
    }
    if (ev_next == kTimingSlots) resolve_timing();
    const int slot = ev_next++;
    pending.push_back({k, slot});
    return slot;
}
void pimc_ctx::timed_begin(int k) {
    pending_end = timed_slot(k);
    if (pending_end >= 0) (void)hipEventRecord(ev_pool[2 * pending_end], stream);
}
void pimc_ctx::timed_end() {
    if (pending_end >= 0) (void)hipEventRecord(ev_pool[2 * pending_end + 1], stream);
    pending_end = -1;
}
pimc::KernelTimer pimc_ctx::kernel_timer(int k) {
    pimc::KernelTimer t;
    const int slot = timed_slot(k);
    if (slot < 0) return t;
    t.start = ev_pool[2 * slot];
    t.stop = ev_pool[2 * slot + 1];
    return t;
}
void pimc_ctx::resolve_timing() {
    for (auto& p : pending) {
        float ms = 0.f;
        hipEvent_t a = ev_pool[2 * p.ev], b = ev_pool[2 * p.ev + 1];
        if (hipEventSynchronize(b) == hipSuccess && hipEventElapsedTime(&ms, a, b) == hipSuccess) {
            kernel_ms[p.kind] += ms;
            kernel_count[p.kind] += 1;
        }
    }
    pending.clear();
    ev_next = 0;
}

// Waits until the call's last kernel has published seq s in mapped memory: a
// spin on a host cache line instead of a stream synchronisation (whose wake-up
// costs several microseconds).  After 0.2 s it falls back to the stream
// synchronisation, which also returns an error the stream may hold.
int pimc_ctx::wait_done(unsigned long long s) {
    const unsigned long long* f = reinterpret_cast<const unsigned long long*>(h_map + kDoneSlot);
    const auto t0 = std::chrono::steady_clock::now();
    for (unsigned it = 1;; ++it) {
        if (__atomic_load_n(f, __ATOMIC_ACQUIRE) == s) return 0;
        if ((it & 4095u) == 0 && std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(200)) break;
        __builtin_ia32_pause();
    }
    const hipError_t e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return (int)e;
    return __atomic_load_n(f, __ATOMIC_ACQUIRE) == s ? 0 : -1;
}

// packet: the kernel went through the AQL queue, which a stream
// synchronisation does not cover; the spin then ends only after 10 s.
int pimc_ctx::wait_direct(int n, unsigned long long s, bool packet) {
    const auto t0 = std::chrono::steady_clock::now();
    const auto limit = std::chrono::milliseconds(packet ? 10000 : 200);
    int k = 0;
    for (unsigned it = 1;; ++it) {
        while (k < n && __atomic_load_n(&h_vdir[k].seq, __ATOMIC_ACQUIRE) == s) ++k;
        if (k == n) return 0;
        if ((it & 4095u) == 0 && std::chrono::steady_clock::now() - t0 > limit) break;
        __builtin_ia32_pause();
    }
    if (packet) return -1;
    const hipError_t e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return (int)e;
    for (; k < n; ++k)
        if (__atomic_load_n(&h_vdir[k].seq, __ATOMIC_ACQUIRE) != s) return -1;
    return 0;
}

// ===========================================================================
// context management
// ===========================================================================
template <typename T>
static hipError_t dalloc(T** p, size_t n) {
    return hipMalloc((void**)p, sizeof(T) * std::max<size_t>(n, 1));
}

bool pimc_ctx::direct_ready(int mode) {
    if (direct_state < 0) {
        direct_state = 0;
        if (pimc::direct_ring_create(device, &dring)) {
            bool all = true;
            for (int m = 0; m < 3 && all; ++m)
                all = pimc::direct_kernel(device, pimc::vnn_direct_stub(m), stream, pimc::kVdirBytes, &dk_vdir[m]);
            if (all)
                direct_state = 1;
            else
                pimc::direct_ring_release(&dring);
        }
    }
    return direct_state == 1 && mode >= 0 && mode < 3 && dk_vdir[mode].ok();
}

bool pimc_ctx::stream_quiet() {
    if (fail_count != fail_seen) {
        if (hipStreamSynchronize(stream) != hipSuccess) return false;
        fail_seen = fail_count;
    }
    return true;
}

pimc_ctx::~pimc_ctx() {
    if (stream) (void)hipStreamSynchronize(stream);
    if (copy_stream) {
        (void)hipStreamSynchronize(copy_stream);
        (void)hipStreamDestroy(copy_stream);
    }
    for (auto& e : chunk_ev)
        if (e) (void)hipEventDestroy(e);
    pimc::direct_ring_release(&dring);
    resolve_timing();
    for (auto& e : ev_pool)
        if (e) (void)hipEventDestroy(e);
    for (void* p : dev_allocs) (void)hipFree(p);
    for (void* p : host_allocs) (void)hipHostFree(p);
    if (stream) (void)hipStreamDestroy(stream);
}

template <typename T>
static int ctx_dalloc(pimc_ctx* c, T** p, size_t n) {
    hipError_t e = dalloc(p, n);
    if (e != hipSuccess) return fail(c, PIMC_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    c->dev_allocs.push_back((void*)*p);
    return PIMC_OK;
}
template <typename T>
static int ctx_halloc(pimc_ctx* c, T** p, size_t n) {
    hipError_t e = hipHostMalloc((void**)p, sizeof(T) * std::max<size_t>(n, 1), hipHostMallocDefault);
    if (e != hipSuccess) return fail(c, PIMC_E_NOMEM, std::string("hipHostMalloc: ") + hipGetErrorString(e));
    c->host_allocs.push_back((void*)*p);
    return PIMC_OK;
}

static void set_box(pimc_ctx* c) {
    BoxDev& b = c->box;
    b.L = c->desc.side;
    b.invL = 1.0 / c->desc.side;
    b.rc2 = c->desc.rc * c->desc.rc;
    // LookupTable::setupNNGrid (src/lookuptable.cpp:73-90)
    int ng = (int)std::floor((c->desc.side / c->desc.rc) + 1.0e-7);
    if (ng < 1) ng = 1;
    b.ngrid = ng;
    b.grid_size = c->desc.side / (1.0 * ng);
}

static void build_grid_nn(pimc_ctx* c) {
    // 27 neighbour cells per cell, duplicates -1 keeping the last occurrence
    // (src/lookuptable.cpp:111-184)
    const int ng = c->box.ngrid;
    const int cells = ng * ng * ng;
    c->grid_nn.assign((size_t)cells * 27, -1);
    for (int n = 0; n < cells; ++n) {
        const int g[3] = {n / (ng * ng), (n / ng) % ng, n % ng};
        int row[27];
        for (int m = 0; m < 27; ++m) {
            const int sh[3] = {m / 9 - 1, (m / 3) % 3 - 1, m % 3 - 1};
            int t[3];
            for (int i = 0; i < 3; ++i) {
                t[i] = g[i] + sh[i];
                if (t[i] == ng) t[i] = 0;
                else if (t[i] == -1) t[i] = ng - 1;
            }
            row[m] = (t[0] * ng + t[1]) * ng + t[2];
        }
        for (int m = 26; m >= 0; --m) {
            if (row[m] < 0) continue;
            for (int p = m - 1; p >= 0; --p)
                if (row[p] == row[m]) row[p] = -1;
        }
        for (int m = 0; m < 27; ++m) c->grid_nn[(size_t)n * 27 + m] = row[m];
    }
}

// Pairs per lane of a tile: about two tiles per persistent workgroup (fine
// enough for the pair-count balance of the ranges, few enough partials), within
// [2, 1024].
// About two tiles per workgroup, but never tiles under 16 pairs per lane: below
// that a tile's fixed cost (descriptor, partial, barrier) outweighs its pairs.
static int pairs_per_lane_for(int64_t total_pairs, int G) {
    const int64_t per_lane = total_pairs / ((int64_t)G * kPairThreads * 2);
    return (int)std::max<int64_t>(16, std::min<int64_t>(1024, per_lane));
}

static void free_dev(pimc_ctx* c, void* p) {
    if (!p) return;
    (void)hipFree(p);
    c->dev_allocs.erase(std::remove(c->dev_allocs.begin(), c->dev_allocs.end(), p), c->dev_allocs.end());
}

static int rebuild_work(pimc_ctx* c) {
    const int M = c->desc.num_slices;
    std::vector<SliceDesc> sl(M);
    int64_t total = 0;
    int maxn = 0;
    for (int m = 0; m < M; ++m) {
        const int64_t n = c->h_nbs[m];
        total += n * (n - 1) / 2;
        maxn = std::max(maxn, c->h_nbs[m]);
    }
    c->lds_beads = maxn <= kLdsBeads ? std::max(maxn, 1) : 0;
    c->pair_lds = c->tune_lds == 0 ? 0 : c->lds_beads;
    // spatial order and the window form (mixed mode): forced (2), or automatic
    // for slices of more than 1024 beads, where most pairs lie beyond the
    // repulsion's reach (N = 2048 x 250: pair kernel 801 -> 676 us; N = 1024: 232
    // -> 214 us, about what the ordering pass costs after a path change)
    c->pair_spatial = c->desc.v_mode == PIMC_V_MIXED && c->pair_lds >= 512 &&
                      (c->tune_lds == 2 || (c->tune_lds == -1 && c->pair_lds > 1024));
    c->perm_version = ~0ull;
    const int G0 = c->tune_gmul > 0 ? c->tune_gmul * std::max(1, c->num_cus) : pair_grid_for(c->lds_beads, c->num_cus);
    const int ppl = c->tune_ppl > 0 ? c->tune_ppl : pairs_per_lane_for(total, G0);
    std::vector<RowTile> tiles;
    for (int m = 0; m < M; ++m) {
        const int n = c->h_nbs[m];
        sl[m].n = n;
        sl[m].slice = m;
        sl[m].walker = 0;
        sl[m].first_item = (int)tiles.size();
        sl[m].pos = (int64_t)m * c->desc.num_alloc;
        sl[m].plane = c->plane;
        append_row_tiles(m, sl[m], ppl, tiles, c->pair_spatial);
    }
    c->n_tiles = (int)tiles.size();
    c->n_groups = std::max(1, std::min(G0, c->n_tiles));
    std::vector<int> ranges;
    balance_row_tiles(tiles, c->n_groups, ranges);
    int r;
    if (c->n_tiles > c->cap_tiles) {
        free_dev(c, c->d_tiles);
        free_dev(c, c->d_partials);
        c->d_tiles = nullptr;
        c->d_partials = nullptr;
        if ((r = ctx_dalloc(c, &c->d_tiles, c->n_tiles))) return r;
        if ((r = ctx_dalloc(c, &c->d_partials, c->n_tiles))) return r;
        c->cap_tiles = c->n_tiles;
    }
    if (c->pair_spatial && !c->d_perm && (r = ctx_dalloc(c, &c->d_perm, (size_t)M * c->desc.num_alloc))) return r;
    if (c->n_groups + 1 > c->cap_groups) {
        free_dev(c, c->d_ranges);
        c->d_ranges = nullptr;
        if ((r = ctx_dalloc(c, &c->d_ranges, c->n_groups + 1))) return r;
        c->cap_groups = c->n_groups + 1;
    }
    HIP_OR_FAIL(c, hipMemcpyAsync(c->d_slices, sl.data(), sizeof(SliceDesc) * M, hipMemcpyHostToDevice, c->stream));
    if (c->n_tiles)
        HIP_OR_FAIL(c, hipMemcpyAsync(c->d_tiles, tiles.data(), sizeof(RowTile) * c->n_tiles, hipMemcpyHostToDevice,
                                      c->stream));
    HIP_OR_FAIL(c, hipMemcpyAsync(c->d_ranges, ranges.data(), sizeof(int) * ranges.size(), hipMemcpyHostToDevice,
                                  c->stream));
    HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
    return PIMC_OK;
}

static AzizDev ctx_aziz(pimc_ctx* c, int* rc) {
    AzizDev a = c->az;
    *rc = PIMC_OK;
    if (c->desc.v_mode == PIMC_V_TABLE) {
        std::string err;
        a.table = table_cache().get(c->desc.aziz_year, 0, a.table_length, &err);
        if (!a.table) *rc = fail(c, PIMC_E_NOMEM, err);
        if (c->correction() && !*rc) {
            // gradV gathers lookupdVdr (include/potential.h:996-1002)
            a.dtable = table_cache().get(c->desc.aziz_year, 1, a.table_length, &err);
            if (!a.dtable) *rc = fail(c, PIMC_E_NOMEM, err);
        }
    }
    return a;
}

// The context's Aziz constants with its tables resolved (for the chain driver).
namespace pimc {
AzizDev context_aziz(pimc_ctx* c, int* rc) { return ctx_aziz(c, rc); }
}  // namespace pimc

static void fill_walker(pimc_ctx* c, WalkerDesc* w, int shift, const double vf[2]) {
    w->box = c->box;
    w->table_length = c->az.table_length;
    w->num_slices = c->desc.num_slices;
    w->first_slice = 0;
    w->shift = shift;
    w->vfactor0 = vf ? vf[0] : c->vf_default[0];
    w->vfactor1 = vf ? vf[1] : c->vf_default[1];
    w->tau = c->desc.tau;
    w->gvf0 = c->gvf[0];
    w->gvf1 = c->gvf[1];
    w->lambda = c->desc.lambda;
    w->worm = c->worm;
}

// gradVSquared(slice) for every slice into c->d_g2part (partials per slice).
static int run_gradv2(pimc_ctx* c, const AzizDev& az) {
    const int M = c->desc.num_slices, N = c->desc.num_alloc;
    if (c->desc.v_mode == PIMC_V_TABLE && !az.dtable)
        return fail(c, PIMC_E_STATE, "gradVSquared in the table mode needs the dV/dr table");
    const int B = (N + 255) / 256;
    if (!c->d_g2part || c->g2_blocks < B) {
        int r;
        if ((r = ctx_dalloc(c, &c->d_g2part, (size_t)M * B))) return r;
        if (!c->d_g2slice && (r = ctx_dalloc(c, &c->d_g2slice, (size_t)M))) return r;
    }
    const int C = gradv2_chunks(N, M, c->num_cus);
    if (C > 1 && c->cap_g2f < (size_t)M * C * N * 3) {
        int r;
        free_dev(c, c->d_g2f);
        c->d_g2f = nullptr;
        if ((r = ctx_dalloc(c, &c->d_g2f, (size_t)M * C * N * 3))) return r;
        c->cap_g2f = (size_t)M * C * N * 3;
    }
    int blocks = 0;
    c->timed_begin(5);
    HIP_OR_FAIL(c, launch_gradv2(c->desc.v_mode, c->stream, c->d_pos, c->plane, N, M, c->d_nbs, c->box, az,
                                 c->lds_beads, c->d_g2part, &blocks, C, c->d_g2f));
    c->timed_end();
    c->g2_blocks = blocks;
    return PIMC_OK;
}

extern "C" {

const char* pimc_last_error(const pimc_ctx* ctx) {
    if (ctx) return ctx->err.c_str();
    return g_last_error.c_str();
}

int pimc_ctx_create(const pimc_desc* d, pimc_ctx** out) {
    if (!d || !out) return fail(nullptr, PIMC_E_ARG, "null argument");
    *out = nullptr;
    if (d->num_slices < 1 || d->num_alloc < 1 || !(d->side > 0) || !(d->tau > 0) || !(d->rc > 0))
        return fail(nullptr, PIMC_E_ARG, "invalid pimc_desc sizes");
    if (d->v_mode < PIMC_V_FP32 || d->v_mode > PIMC_V_TABLE) return fail(nullptr, PIMC_E_ARG, "invalid v_mode");
    AzizParams p;
    if (!aziz_params(d->aziz_year, &p)) return fail(nullptr, PIMC_E_ARG, "aziz_year must be 1979, 1987 or 1995");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fail(nullptr, PIMC_E_HIP, "no HIP device available");
    std::unique_ptr<pimc_ctx> c(new pimc_ctx());
    c->desc = *d;
    c->params = p;
    {
        int dev = 0, cus = 0;
        HIP_OR_FAIL(nullptr, hipGetDevice(&dev));
        c->device = dev;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && cus > 0)
            c->num_cus = cus;
        // PIMC_DIRECT_DISPATCH=0: short Delta V requests launch through HIP
        const char* dd = std::getenv("PIMC_DIRECT_DISPATCH");
        if (dd && std::strcmp(dd, "0") == 0) c->direct_state = 0;
    }
    const int M = d->num_slices, N = d->num_alloc;
    c->plane = (int64_t)M * N;
    set_box(c.get());
    // Prism::maxSep = |side/2| over the three axes (src/container.cpp:100)
    const double h = d->side / 2.0;
    double ms = 0.0;
    ms += h * h;
    ms += h * h;
    ms += h * h;
    c->max_sep = std::sqrt(ms);
    c->az = make_aziz_dev(p, c->max_sep);
    c->tailV = aziz_tail(p, d->rc);
    c->worm = WormDev{{-1, -1}, {-1, -1}, {-1, -1}, {-1, -1}};
    HIP_OR_FAIL(nullptr, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    const int cells = c->box.ngrid * c->box.ngrid * c->box.ngrid;
    c->cells = cells;
    int r;
    if ((r = ctx_dalloc(c.get(), &c->d_pos, 3 * (size_t)M * N))) return r;
    if ((r = ctx_dalloc(c.get(), &c->d_on, (size_t)M * N))) return r;
    if ((r = ctx_dalloc(c.get(), &c->d_nbs, M))) return r;
    if ((r = ctx_dalloc(c.get(), &c->d_cell, (size_t)M * N))) return r;
    if ((r = ctx_dalloc(c.get(), &c->d_label, (size_t)M * N))) return r;
    if ((r = ctx_dalloc(c.get(), &c->d_numlab, (size_t)cells * M))) return r;
    if ((r = ctx_dalloc(c.get(), &c->d_hash, (size_t)cells * M * N))) return r;
    if ((r = ctx_dalloc(c.get(), &c->d_grid_nn, (size_t)cells * 27))) return r;
    if ((r = ctx_dalloc(c.get(), &c->d_slices, M))) return r;
    if ((r = ctx_dalloc(c.get(), &c->d_walker, 1))) return r;
    if ((r = ctx_dalloc(c.get(), &c->d_vslice, M))) return r;
    if ((r = ctx_dalloc(c.get(), &c->d_out, 16))) return r;
    if ((r = ctx_dalloc(c.get(), &c->d_hist, 64))) return r;
    if ((r = ctx_dalloc(c.get(), &c->d_kin_partials, 1024))) return r;
    if ((r = ctx_halloc(c.get(), &c->h_stage, std::max<size_t>(3 * (size_t)M * N, 256)))) return r;
    if ((r = ctx_halloc(c.get(), &c->h_out, 64))) return r;
    {
        // small results are written by the kernels straight into host memory
        hipError_t e = hipHostMalloc((void**)&c->h_map, sizeof(double) * 64,
                                     hipHostMallocMapped | hipHostMallocCoherent);
        if (e != hipSuccess) return fail(nullptr, PIMC_E_NOMEM, std::string("hipHostMalloc: ") + hipGetErrorString(e));
        c->host_allocs.push_back((void*)c->h_map);
        std::memset(c->h_map, 0, sizeof(double) * 64);
        e = hipHostGetDevicePointer((void**)&c->d_map, c->h_map, 0);
        if (e != hipSuccess) return fail(nullptr, PIMC_E_HIP, std::string("hipHostGetDevicePointer: ") + hipGetErrorString(e));
        e = hipHostMalloc((void**)&c->h_vdir, sizeof(VnnDirectOut) * kVnnInline,
                          hipHostMallocMapped | hipHostMallocCoherent);
        if (e != hipSuccess) return fail(nullptr, PIMC_E_NOMEM, std::string("hipHostMalloc: ") + hipGetErrorString(e));
        c->host_allocs.push_back((void*)c->h_vdir);
        std::memset(c->h_vdir, 0, sizeof(VnnDirectOut) * kVnnInline);
        e = hipHostGetDevicePointer((void**)&c->d_vdir, c->h_vdir, 0);
        if (e != hipSuccess) return fail(nullptr, PIMC_E_HIP, std::string("hipHostGetDevicePointer: ") + hipGetErrorString(e));
    }
    HIP_OR_FAIL(c.get(), hipMemsetAsync(c->d_hash, 0xff, sizeof(int) * (size_t)cells * M * N, c->stream));
    HIP_OR_FAIL(c.get(), hipMemsetAsync(c->d_numlab, 0, sizeof(int) * (size_t)cells * M, c->stream));
    build_grid_nn(c.get());
    HIP_OR_FAIL(c.get(), hipMemcpyAsync(c->d_grid_nn, c->grid_nn.data(), sizeof(int) * c->grid_nn.size(),
                                        hipMemcpyHostToDevice, c->stream));
    c->h_nbs.assign(M, 0);
    c->h_on.assign((size_t)M * N, 0);
    HIP_OR_FAIL(c.get(), hipStreamSynchronize(c->stream));
    *out = c.release();
    return PIMC_OK;
}

void pimc_ctx_destroy(pimc_ctx* ctx) { delete ctx; }

int pimc_set_v_mode(pimc_ctx* c, int v_mode) {
    if (!c) return fail(nullptr, PIMC_E_ARG, "null ctx");
    if (v_mode < PIMC_V_FP32 || v_mode > PIMC_V_TABLE) return fail(c, PIMC_E_ARG, "invalid v_mode");
    c->desc.v_mode = v_mode;
    ++c->path_version;
    return c->have_path ? rebuild_work(c) : PIMC_OK;   // the tiles' shape follows the mode
}

int pimc_upload_path(pimc_ctx* c, const double* beads, const uint8_t* bead_on, const int32_t* num_beads) {
    if (!c || !beads) return fail(c, PIMC_E_ARG, "null argument");
    const int M = c->desc.num_slices, N = c->desc.num_alloc;
    const size_t MN = (size_t)M * N;
    for (int m = 0; m < M; ++m) {
        const int nb = num_beads ? num_beads[m] : N;
        if (nb < 0 || nb > N) return fail(c, PIMC_E_ARG, "num_beads out of range");
        c->h_nbs[m] = nb;
    }
    for (size_t t = 0; t < MN; ++t) c->h_on[t] = bead_on ? (bead_on[t] ? 1 : 0) : 1;
    // AoS [M][N][3] -> SoA planes
    for (size_t t = 0; t < MN; ++t) {
        c->h_stage[t] = beads[3 * t];
        c->h_stage[t + MN] = beads[3 * t + 1];
        c->h_stage[t + 2 * MN] = beads[3 * t + 2];
    }
    HIP_OR_FAIL(c, hipMemcpyAsync(c->d_pos, c->h_stage, sizeof(double) * 3 * MN, hipMemcpyHostToDevice, c->stream));
    HIP_OR_FAIL(c, hipMemcpyAsync(c->d_on, c->h_on.data(), MN, hipMemcpyHostToDevice, c->stream));
    HIP_OR_FAIL(c, hipMemcpyAsync(c->d_nbs, c->h_nbs.data(), sizeof(int) * M, hipMemcpyHostToDevice, c->stream));
    HIP_OR_FAIL(c, hipMemsetAsync(c->d_hash, 0xff, sizeof(int) * (size_t)c->cells * MN, c->stream));
    HIP_OR_FAIL(c, launch_cell_build(c->stream, c->d_pos, c->plane, N, M, c->d_on, c->d_cell, c->d_label,
                                     c->d_numlab, c->d_hash, c->cells, c->box));
    HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
    c->have_path = true;
    ++c->path_version;
    return rebuild_work(c);
}

int pimc_set_worm(pimc_ctx* c, const int32_t head[2], const int32_t tail[2], const int32_t special1[2],
                  const int32_t special2[2]) {
    if (!c) return fail(nullptr, PIMC_E_ARG, "null ctx");
    auto cp = [](int* d, const int32_t* s) {
        d[0] = s ? s[0] : -1;
        d[1] = s ? s[1] : -1;
    };
    cp(c->worm.head, head);
    cp(c->worm.tail, tail);
    cp(c->worm.sp1, special1);
    cp(c->worm.sp2, special2);
    ++c->path_version;
    return PIMC_OK;
}

int pimc_download_path(pimc_ctx* c, double* beads) {
    if (!c || !beads) return fail(c, PIMC_E_ARG, "null argument");
    if (!c->have_path) return fail(c, PIMC_E_STATE, "no path uploaded");
    const size_t MN = (size_t)c->desc.num_slices * c->desc.num_alloc;
    HIP_OR_FAIL(c, hipMemcpyAsync(c->h_stage, c->d_pos, sizeof(double) * 3 * MN, hipMemcpyDeviceToHost, c->stream));
    HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
    for (size_t t = 0; t < MN; ++t) {
        beads[3 * t] = c->h_stage[t];
        beads[3 * t + 1] = c->h_stage[t + MN];
        beads[3 * t + 2] = c->h_stage[t + 2 * MN];
    }
    return PIMC_OK;
}

static int check_beads(pimc_ctx* c, const int32_t* sp, int n) {
    if (n < 0 || (n > 0 && !sp)) return fail(c, PIMC_E_ARG, "bad bead list");
    for (int k = 0; k < n; ++k)
        if (sp[2 * k] < 0 || sp[2 * k] >= c->desc.num_slices || sp[2 * k + 1] < 0 ||
            sp[2 * k + 1] >= c->desc.num_alloc)
            return fail(c, PIMC_E_ARG, "bead locator out of range");
    return PIMC_OK;
}

static int ensure_req(pimc_ctx* c, int n) {
    if (n <= c->cap_req) return PIMC_OK;
    int cap = std::max(64, n);
    int r;
    if ((r = ctx_dalloc(c, &c->d_req_beads, cap))) return r;
    if ((r = ctx_halloc(c, &c->h_req_beads, cap))) return r;
    if ((r = ctx_dalloc(c, &c->d_req_sp, 2 * (size_t)cap))) return r;
    if ((r = ctx_dalloc(c, &c->d_req_pos, 3 * (size_t)cap))) return r;
    if ((r = ctx_halloc(c, &c->h_req_sp, 2 * (size_t)cap))) return r;
    if ((r = ctx_halloc(c, &c->h_req_pos, 3 * (size_t)cap))) return r;
    if ((r = ctx_dalloc(c, &c->d_per_bead, 2 * (size_t)cap))) return r;
    if ((r = ctx_halloc(c, &c->h_per_bead, 2 * (size_t)cap))) return r;
    if ((r = ctx_dalloc(c, &c->d_per_bead_cor, 2 * (size_t)cap))) return r;
    if ((r = ctx_halloc(c, &c->h_per_bead_cor, 2 * (size_t)cap))) return r;
    const int chunks = (c->desc.num_alloc + kVnnChunk - 1) / kVnnChunk;
    if ((r = ctx_dalloc(c, &c->d_vnn_partials, 2 * (size_t)cap * chunks))) return r;
    if (!c->d_vnn_ticket) {
        if ((r = ctx_dalloc(c, &c->d_vnn_ticket, 1))) return r;
        HIP_OR_FAIL(c, hipMemsetAsync(c->d_vnn_ticket, 0, sizeof(unsigned), c->stream));
    }
    c->cap_req = cap;
    return PIMC_OK;
}

int pimc_update_beads(pimc_ctx* c, const int32_t* sp, const double* pos, int n) {
    if (!c) return fail(nullptr, PIMC_E_ARG, "null ctx");
    if (!c->have_path) return fail(c, PIMC_E_STATE, "no path uploaded");
    int r;
    if ((r = check_beads(c, sp, n))) return r;
    if (n == 0) return PIMC_OK;
    if (!pos) return fail(c, PIMC_E_ARG, "null positions");
    if ((r = ensure_req(c, n))) return r;
    std::memcpy(c->h_req_sp, sp, sizeof(int) * 2 * n);
    std::memcpy(c->h_req_pos, pos, sizeof(double) * 3 * n);
    HIP_OR_FAIL(c, hipMemcpyAsync(c->d_req_sp, c->h_req_sp, sizeof(int) * 2 * n, hipMemcpyHostToDevice, c->stream));
    HIP_OR_FAIL(c, hipMemcpyAsync(c->d_req_pos, c->h_req_pos, sizeof(double) * 3 * n, hipMemcpyHostToDevice,
                                  c->stream));
    // beads on distinct slices touch disjoint cell tables: one lane each
    std::vector<char> seen(c->desc.num_slices, 0);
    bool distinct = true;
    for (int k = 0; k < n && distinct; ++k) {
        if (seen[sp[2 * k]]) distinct = false;
        seen[sp[2 * k]] = 1;
    }
    HIP_OR_FAIL(c, launch_update_beads(c->stream, c->d_req_sp, c->d_req_pos, n, c->d_pos, c->plane,
                                       c->desc.num_alloc, c->desc.num_slices, c->d_cell, c->d_label, c->d_numlab,
                                       c->d_hash, c->box, distinct));
    ++c->path_version;
    HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
    return PIMC_OK;
}

int pimc_path_version(const pimc_ctx* c, uint64_t* version) {
    if (!c || !version) return fail(nullptr, PIMC_E_ARG, "null argument");
    *version = c->path_version;
    return PIMC_OK;
}

int pimc_gradv_squared(pimc_ctx* c, double* g2) {
    if (!c || !g2) return fail(c, PIMC_E_ARG, "null argument");
    if (!c->have_path) return fail(c, PIMC_E_STATE, "no path uploaded");
    const int M = c->desc.num_slices;
    int r;
    AzizDev az = ctx_aziz(c, &r);
    if (r) return r;
    if (c->desc.v_mode == PIMC_V_TABLE && !az.dtable) {
        // gradV gathers lookupdVdr whether or not the context's action carries the correction
        std::string err;
        az.dtable = table_cache().get(c->desc.aziz_year, 1, az.table_length, &err);
        if (!az.dtable) return fail(c, PIMC_E_NOMEM, err);
    }
    if ((r = run_gradv2(c, az))) return r;
    // the per-block partials of every slice, summed in block order as the
    // pair kernel's finalisation sums them into the corrected action
    const int B = c->g2_blocks;
    std::vector<double> part((size_t)M * B);
    HIP_OR_FAIL(c, hipMemcpyAsync(part.data(), c->d_g2part, sizeof(double) * part.size(), hipMemcpyDeviceToHost,
                                  c->stream));
    HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
    for (int m = 0; m < M; ++m) {
        double t = 0.0;
        for (int b = 0; b < B; ++b) t += part[(size_t)m * B + b];
        g2[m] = t;
    }
    return PIMC_OK;
}

int pimc_potential_slices(pimc_ctx* c, double* v_slices, int32_t* sep_hist) {
    const double vf[2] = {1.0, 1.0};
    double act = 0.0;
    int r = pimc_total_action(c, 1, vf, &act, v_slices);
    if (r) return r;
    if (sep_hist) {
        const int m = c->desc.num_slices - 1;
        const double dsep = 0.5 * std::sqrt(3.0) * c->desc.side / (1.0 * 50);
        HIP_OR_FAIL(c, hipMemsetAsync(c->d_hist, 0, sizeof(int) * 64, c->stream));
        HIP_OR_FAIL(c, launch_sep_hist(c->stream, c->d_pos, c->plane, m, c->h_nbs[m], c->desc.num_alloc, c->box,
                                       dsep, c->d_hist));
        HIP_OR_FAIL(c, hipMemcpyAsync(c->h_stage, c->d_hist, sizeof(int) * 50, hipMemcpyDeviceToHost, c->stream));
        HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
        std::memcpy(sep_hist, c->h_stage, sizeof(int) * 50);
    }
    return PIMC_OK;
}

int pimc_total_action(pimc_ctx* c, int shift, const double vfactor[2], double* action, double* v_slices) {
    if (!c || !action) return fail(c, PIMC_E_ARG, "null argument");
    if (!c->have_path) return fail(c, PIMC_E_STATE, "no path uploaded");
    if (shift < 1) return fail(c, PIMC_E_ARG, "shift must be >= 1");
    const int M = c->desc.num_slices;
    int r;
    AzizDev az = ctx_aziz(c, &r);
    if (r) return r;
    WalkerDesc w;
    std::memset(&w, 0, sizeof(w));
    fill_walker(c, &w, shift, vfactor);
    if (!c->walker_valid || std::memcmp(&w, &c->walker_cache, sizeof(w)) != 0) {
        HIP_OR_FAIL(c, hipMemcpyAsync(c->d_walker, &w, sizeof(w), hipMemcpyHostToDevice, c->stream));
        c->walker_cache = w;
        c->walker_valid = true;
    }
    if (c->n_tiles > 0) {
        const bool cor = c->correction();
        if (cor && (r = run_gradv2(c, az))) return r;
        const bool spatial = c->pair_spatial && c->desc.v_mode == PIMC_V_MIXED;
        if (spatial && c->perm_version != c->path_version) {
            HIP_OR_FAIL(c, launch_slice_order(c->stream, c->d_slices, M, c->d_pos, c->box.invL, c->d_perm));
            c->perm_version = c->path_version;
        }
        const KernelTimer kt = c->kernel_timer(0);
        const unsigned long long seq = ++c->seq;
        HIP_OR_FAIL(c, launch_pair_action(c->desc.v_mode, c->stream, c->d_tiles, c->n_tiles, c->d_ranges, c->n_groups,
                                          c->d_slices, M, c->d_walker, 1, c->d_pos, az, c->pair_lds, c->d_partials,
                                          c->d_vslice, c->d_map, &kt, true, c->d_done(), seq,
                                          cor ? c->d_g2part : nullptr, c->g2_blocks, c->d_g2slice,
                                          spatial ? c->d_perm : nullptr));
        if (v_slices) {
            HIP_OR_FAIL(c, hipMemcpyAsync(c->h_stage, c->d_vslice, sizeof(double) * M, hipMemcpyDeviceToHost,
                                          c->stream));
            HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
        } else if (c->wait_done(seq) != 0) {
            return fail(c, PIMC_E_HIP, "pair action: completion not published");
        }
        *action = c->h_map[0];
        if (v_slices)
            for (int m = 0; m < M; ++m) {
                v_slices[2 * m] = 0.0;
                v_slices[2 * m + 1] = c->h_stage[m];
            }
    } else {
        *action = 0.0;
        if (v_slices)
            for (int m = 0; m < 2 * M; ++m) v_slices[m] = 0.0;
    }
    return PIMC_OK;
}

// Per-slice virial sums (pimc_virial.hip), summed over each slice's workgroups
// in order: S [M][kVirialTerms].
static int run_virial(pimc_ctx* c, const int32_t* next_links, int window, std::vector<double>& S) {
    const int M = c->desc.num_slices, N = c->desc.num_alloc;
    if (window < 1) return fail(c, PIMC_E_ARG, "virial window must be >= 1");
    int r;
    AzizDev az = ctx_aziz(c, &r);
    if (r) return r;
    if (c->desc.v_mode == PIMC_V_TABLE) {
        // gradV and grad2V gather lookupdVdr / lookupd2Vdr2 (include/potential.h:996-1017)
        std::string err;
        az.dtable = cached_table(c->desc.aziz_year, 1, az.table_length, &err);
        if (az.dtable) az.d2table = cached_table(c->desc.aziz_year, 2, az.table_length, &err);
        if (!az.dtable || !az.d2table) return fail(c, PIMC_E_NOMEM, err);
    }
    // links: next as given, prev as its inverse; every active bead must link to
    // an active bead (closed worldlines, the estimator's diagonal sector).  The
    // last call's links stay on the device: the same links on the same beads per
    // slice are neither checked nor uploaded again.
    const size_t n_links = 2 * (size_t)M * N;
    const bool same_links = next_links && c->links_ok && c->links_nbs == c->h_nbs &&
                            std::memcmp(c->links_next.data(), next_links, sizeof(int32_t) * n_links) == 0;
    std::vector<int32_t> prev;
    if (next_links && !same_links) {
        c->links_ok = false;
        prev.assign(n_links, -1);
        for (int m = 0; m < M; ++m)
            for (int i = 0; i < c->h_nbs[m]; ++i) {
                const size_t t = (size_t)m * N + i;
                const int a = next_links[2 * t], b = next_links[2 * t + 1];
                if (a < 0 || a >= M || b < 0 || b >= c->h_nbs[a])
                    return fail(c, PIMC_E_ARG, "virial estimator: a next link leaves the closed worldlines");
                const size_t u = (size_t)a * N + b;
                if (prev[2 * u] != -1) return fail(c, PIMC_E_ARG, "virial estimator: next links are not a permutation");
                prev[2 * u] = m;
                prev[2 * u + 1] = i;
            }
    }
    int* d_links = nullptr;
    if (next_links && !same_links) {
        if (c->cap_links < 2 * n_links) {
            free_dev(c, c->d_links);
            c->d_links = nullptr;
            c->cap_links = 0;
            if ((r = ctx_dalloc(c, &c->d_links, 2 * n_links))) return r;
            c->cap_links = 2 * n_links;
        }
        HIP_OR_FAIL(c, hipMemcpyAsync(c->d_links, next_links, sizeof(int) * n_links, hipMemcpyHostToDevice,
                                      c->stream));
        HIP_OR_FAIL(c, hipMemcpyAsync(c->d_links + n_links, prev.data(), sizeof(int) * n_links,
                                      hipMemcpyHostToDevice, c->stream));
        c->links_next.assign(next_links, next_links + n_links);
        c->links_nbs = c->h_nbs;
        c->links_ok = true;
    }
    if (next_links) d_links = c->d_links;
    const int B = (N + 255) / 256;
    const size_t need = (size_t)M * B * kVirialTerms;
    if (c->cap_virial < need) {
        if ((r = ctx_dalloc(c, &c->d_virial, need))) return r;
        c->cap_virial = need;
    }
    VirialArgs A{};
    A.pos = c->d_pos;
    A.plane = c->plane;
    A.nalloc = N;
    A.num_slices = M;
    A.nbs = c->d_nbs;
    A.next = d_links;
    A.prev = d_links ? d_links + n_links : nullptr;
    A.window = window;
    A.t_even = c->gvf[0] > kCorrEps;
    A.t_odd = c->gvf[1] > kCorrEps;
    A.lds_beads = N <= kLdsBeads ? N : 0;
    A.box = c->box;
    A.part = c->d_virial;
    A.chunks = gradv2_chunks(N, M, c->num_cus);
    A.fpart = nullptr;
    if (A.chunks > 1) {
        const size_t fneed = (size_t)M * A.chunks * N * 9;
        if (c->cap_vf < fneed) {
            free_dev(c, c->d_vf);
            c->d_vf = nullptr;
            if ((r = ctx_dalloc(c, &c->d_vf, fneed))) return r;
            c->cap_vf = fneed;
        }
        A.fpart = c->d_vf;
    }
    int blocks = 0;
    c->timed_begin(7);
    HIP_OR_FAIL(c, launch_virial(c->desc.v_mode, c->stream, A, az, &blocks));
    c->timed_end();
    std::vector<double> part(need);
    HIP_OR_FAIL(c, hipMemcpyAsync(part.data(), c->d_virial, sizeof(double) * need, hipMemcpyDeviceToHost,
                                  c->stream));
    HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
    S.assign((size_t)M * kVirialTerms, 0.0);
    for (int m = 0; m < M; ++m)
        for (int b = 0; b < blocks; ++b)
            for (int q = 0; q < kVirialTerms; ++q)
                S[(size_t)m * kVirialTerms + q] += part[((size_t)m * blocks + b) * kVirialTerms + q];
    return PIMC_OK;
}

// The LocalAction virial virtuals per slice from the slice sums S.
static void virial_slice_terms(const pimc_ctx* c, const std::vector<double>& S, double* out6) {
    const double tau = c->desc.tau, lambda = c->desc.lambda;
    for (int m = 0; m < c->desc.num_slices; ++m) {
        const double* s = &S[(size_t)m * kVirialTerms];
        const int eo = m % 2;
        const double gvf = c->gvf[eo];
        const bool cor = gvf > kCorrEps;
        const double t2 = 2.0 * gvf * (tau * tau * tau) * lambda;
        double* o = out6 + 6 * (size_t)m;
        o[0] = c->vf_default[eo] * tau * s[0];                          // rDOTgradUterm1 (:1446-1482)
        o[1] = cor ? s[1] * t2 : 0.0;                                   // rDOTgradUterm2 (:1494-1574)
        o[2] = c->vf_default[eo] * tau * s[2];                          // deltadotgradUterm1 (:1587-1653)
        o[3] = cor ? s[3] * t2 : 0.0;                                   // deltadotgradUterm2 (:1665-1762)
        o[4] = cor ? s[4] * (gvf * (tau * tau * tau) * lambda) : 0.0;   // virialKinCorrection (:1769-1786)
        o[5] = cor ? 6.0 * gvf * tau * lambda * s[4] : 0.0;             // secondderivPotentialActionTau (:776-787)
    }
}

int pimc_virial_terms(pimc_ctx* c, const int32_t* next_links, int window, double* out6) {
    if (!c || !out6) return fail(c, PIMC_E_ARG, "null argument");
    if (!c->have_path) return fail(c, PIMC_E_STATE, "no path uploaded");
    std::vector<double> S;
    int r = run_virial(c, next_links, window, S);
    if (r) return r;
    virial_slice_terms(c, S, out6);
    return PIMC_OK;
}

int pimc_virial_energy(pimc_ctx* c, const int32_t* next_links, int window, double mu, double* out19) {
    if (!c || !out19) return fail(c, PIMC_E_ARG, "null argument");
    if (!c->have_path) return fail(c, PIMC_E_STATE, "no path uploaded");
    const int M = c->desc.num_slices;
    std::vector<double> vs(2 * (size_t)M), terms(6 * (size_t)M), S;
    int r = pimc_potential_slices(c, vs.data(), nullptr);
    if (r) return r;
    if ((r = run_virial(c, next_links, window, S))) return r;
    virial_slice_terms(c, S, terms.data());
    // VirialEnergyEstimator::accumulate (src/estimator.cpp:1086-1251), one call
    long on = 0;
    for (auto v : c->h_on) on += v;
    const int Np = (int)(on / M);
    const double tau = c->desc.tau, lambda = c->desc.lambda;
    const double volume = c->desc.side * c->desc.side * c->desc.side;
    const double beta = 1.0 * M * tau;
    const double tailV = (1.0 * Np * Np / volume) * c->tailV;
    const double thermTerm1 = (0.5 * 3 / tau) * Np;
    const double T1 = 0.5 * 3 * Np / (1.0 * window * tau);
    const double exchangeNorm = 1.0 / (4.0 * window * (tau * tau) * lambda * M);
    double T2 = 0.0, thermE = 0.0, T3 = 0.0, T4 = 0.0, T5 = 0.0, virKin = 0.0, totVop = 0.0, P3 = 0.0, d2U = 0.0;
    for (int m = 0; m < M; ++m) {
        const double* s = &S[(size_t)m * kVirialTerms];
        T2 += s[5];
        thermE += s[6];
        const double* o = &terms[6 * (size_t)m];
        const int eo = m % 2;
        T3 += o[2];
        T4 += o[3];
        // derivPotentialActionTau (src/action.cpp:751-765)
        double dU = c->vf_default[eo] * (vs[2 * m] + vs[2 * m + 1]);
        if (c->gvf[eo] > kCorrEps) dU += 3.0 * c->gvf[eo] * tau * tau * lambda * s[4];
        T5 += dU;
        virKin += o[4];
        if (eo == 0) totVop += vs[2 * m] + vs[2 * m + 1];
        P3 += o[0] + o[1];
        d2U += o[5] / (1.0 * M);
    }
    double P2 = thermE;
    T2 *= exchangeNorm;
    P2 /= (2.0 * lambda * tau * M);
    double Pressure = 3.0 * Np + P2;
    P3 *= (1.0 / (2.0 * M));
    Pressure -= P3;
    Pressure /= (3.0 * tau * volume);
    totVop /= (0.5 * M);
    totVop += tailV;
    T3 /= (2.0 * beta);
    T4 /= (1.0 * beta);
    T5 /= (1.0 * M);
    virKin /= (0.5 * beta);
    thermE *= (1.0 / (4.0 * lambda * tau)) / (tau * M);
    thermE += thermTerm1;
    thermE += T5;
    thermE += tailV;
    const double totEcv = T1 + T2 + T3 + T4 + T5 + tailV;
    const double Kcv = T1 + T2 + T3 + T4 + virKin;
    double dEdB = (-1.0 * T1 - 2.0 * T2 + 2.0 * T4) / tau;
    dEdB += d2U;
    dEdB *= beta * beta / (1.0 * M);
    for (int k = 0; k < 19; ++k) out19[k] = 0.0;
    out19[0] = totEcv - totVop;
    out19[1] = Kcv;
    out19[2] = totVop;
    out19[3] = totEcv - Kcv;
    out19[4] = totEcv;
    out19[5] = totEcv - mu * Np;
    if (Np > 0) {
        out19[6] = (totEcv - totVop) / (1.0 * Np);
        out19[7] = Kcv / (1.0 * Np);
        out19[8] = totVop / (1.0 * Np);
        out19[9] = (totEcv - Kcv) / (1.0 * Np);
        out19[10] = totEcv / (1.0 * Np);
    }
    out19[11] = totEcv * thermE * beta * beta;
    out19[12] = totEcv * beta;
    out19[13] = dEdB;
    out19[14] = totEcv * thermE * beta * beta * totEcv * beta;
    // the reference adds this term under the key "cVCov2", which its constructor
    // never registered: std::map inserts it at index 0, K_op (src/estimator.cpp:1244)
    out19[0] += totEcv * beta * dEdB;
    out19[15] = 0.0;
    out19[16] = totEcv * thermE * beta * beta * dEdB;
    out19[17] = thermE;
    out19[18] = Pressure;
    return PIMC_OK;
}

int pimc_set_launch_shape(pimc_ctx* c, int pairs_per_lane, int grid_per_cu, int stage_in_lds) {
    if (!c) return fail(c, PIMC_E_ARG, "null context");
    if (pairs_per_lane < 0 || pairs_per_lane > 4096 || grid_per_cu < 0 || grid_per_cu > 64 || stage_in_lds < -1 ||
        stage_in_lds > 2)
        return fail(c, PIMC_E_ARG, "launch shape out of range");
    c->tune_ppl = pairs_per_lane;
    c->tune_gmul = grid_per_cu;
    c->tune_lds = stage_in_lds;
    return c->have_path ? rebuild_work(c) : PIMC_OK;
}

int pimc_autotune(pimc_ctx* c, int reps, double* out4) {
    if (!c || reps < 1) return fail(c, PIMC_E_ARG, "bad argument");
    if (!c->have_path) return fail(c, PIMC_E_STATE, "no path uploaded");
    const int ppls[] = {0, 8, 16, 32, 64, 128};   // 0: the automatic choice
    const int gmul[] = {2, 4, 8, 16};
    // staging: index order, none, and spatial order where it applies
    std::vector<int> ldss = {1, 0};
    if (c->lds_beads >= 512 && c->desc.v_mode == PIMC_V_MIXED) ldss.push_back(2);
    c->resolve_timing();
    const bool timing = c->timing;
    const unsigned kinds = c->timing_kinds;
    const double keep_ms = c->kernel_ms[0];
    const long keep_n = c->kernel_count[0];
    c->timing = true;
    c->timing_kinds = ~0u;
    double best = 1e30, act = 0.0;
    int best_p = 0, best_g = 0, best_l = -1, tried = 0, r;
    for (int ppl : ppls)
        for (int gm : gmul)
            for (int l : ldss) {
                if ((r = pimc_set_launch_shape(c, ppl, gm, l))) return r;
                if ((r = pimc_total_action(c, 1, nullptr, &act, nullptr))) return r;   // warm
                c->resolve_timing();
                const double ms0 = c->kernel_ms[0];
                const long n0 = c->kernel_count[0];
                for (int k = 0; k < reps; ++k)
                    if ((r = pimc_total_action(c, 1, nullptr, &act, nullptr))) return r;
                c->resolve_timing();
                const double avg = (c->kernel_ms[0] - ms0) / std::max<long>(1, c->kernel_count[0] - n0);
                ++tried;
                if (avg < best) {
                    best = avg;
                    best_p = ppl;
                    best_g = gm;
                    best_l = l;
                }
            }
    c->timing = timing;
    c->timing_kinds = kinds;
    c->kernel_ms[0] = keep_ms;   // the tuning launches stay out of the kernel totals
    c->kernel_count[0] = keep_n;
    r = pimc_set_launch_shape(c, best_p, best_g, best_l);
    if (out4) {
        out4[0] = best_p;
        out4[1] = best_g;
        out4[2] = best_l;
        out4[3] = best;
    }
    (void)tried;
    return r;
}

int pimc_energy(pimc_ctx* c, const int32_t* next_links, double mu, double* out9) {
    if (!c || !out9) return fail(c, PIMC_E_ARG, "null argument");
    if (!c->have_path) return fail(c, PIMC_E_STATE, "no path uploaded");
    const int M = c->desc.num_slices, N = c->desc.num_alloc;
    std::vector<double> vs(2 * (size_t)M);
    int r = pimc_potential_slices(c, vs.data(), nullptr);
    if (r) return r;
    int* d_next = nullptr;
    if (next_links) {
        // the last call's links stay on the device: the same links again are
        // neither checked nor uploaded
        const size_t nl = 2 * (size_t)M * N;
        const bool same = c->enext_ok && std::memcmp(c->enext.data(), next_links, sizeof(int32_t) * nl) == 0;
        if (!same) {
            c->enext_ok = false;
            for (size_t t = 0; t < (size_t)M * N; ++t) {
                const int a = next_links[2 * t], b = next_links[2 * t + 1];
                const bool null_link = (a == -1 && b == -1);
                if (!null_link && (a < 0 || a >= M || b < 0 || b >= N))
                    return fail(c, PIMC_E_ARG, "next link out of range");
            }
            if (c->cap_enext < nl) {
                free_dev(c, c->d_enext);
                c->d_enext = nullptr;
                c->cap_enext = 0;
                if ((r = ctx_dalloc(c, &c->d_enext, nl))) return r;
                c->cap_enext = nl;
            }
            HIP_OR_FAIL(c, hipMemcpyAsync(c->d_enext, next_links, sizeof(int) * nl, hipMemcpyHostToDevice, c->stream));
            c->enext.assign(next_links, next_links + nl);
            c->enext_ok = true;
        }
        d_next = c->d_enext;
    }
    HIP_OR_FAIL(c, launch_kinetic(c->stream, c->d_pos, c->plane, N, M, c->d_nbs, d_next, c->box, c->d_kin_partials,
                                  1024, c->d_out + 8));
    HIP_OR_FAIL(c, hipMemcpyAsync(c->h_out, c->d_out + 8, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
    const double link_sum = c->h_out[0];
    // EnergyEstimator::accumulate (src/estimator.cpp:940-1029) with the context's action factors
    long on = 0;
    for (auto v : c->h_on) on += v;
    const int Np = (int)(on / M);
    const double tau = c->desc.tau, lambda = c->desc.lambda;
    const double volume = c->desc.side * c->desc.side * c->desc.side;
    const double tailV = (1.0 * Np * Np / volume) * c->tailV;
    const double kinNorm = (1.0 / (4.0 * lambda * tau)) / (tau * M);
    const double classical = (0.5 * 3 / tau) * Np;
    double totK = -link_sum * kinNorm;
    // gradVSquared(slice) from the same pass when the action carries the correction
    std::vector<double> g2(M, 0.0);
    if (c->correction()) {
        HIP_OR_FAIL(c, hipMemcpy(g2.data(), c->d_g2slice, sizeof(double) * M, hipMemcpyDeviceToHost));
    }
    // derivPotentialActionLambda / Tau (src/action.cpp:751-802) and the operator
    // V on every period-th slice
    double t1 = 0.0, t2 = 0.0, vop0 = 0.0, vop1 = 0.0;
    for (int m = 0; m < M; ++m) {
        const int eo = m % 2;
        const double gvf = c->gvf[eo];
        const bool cor = gvf > kCorrEps;
        t1 += cor ? gvf * tau * tau * tau * g2[m] : 0.0;
        double dU = c->vf_default[eo] * (vs[2 * m] + vs[2 * m + 1]);
        if (cor) dU += 3.0 * gvf * tau * tau * lambda * g2[m];
        t2 += dU;
        if (!(m % c->period)) {
            vop0 += vs[2 * m];
            vop1 += vs[2 * m + 1];
        }
    }
    t1 *= lambda / (tau * M);
    t2 /= 1.0 * M;
    vop0 /= (M / c->period);
    vop1 /= (M / c->period);
    totK += (classical + t1);
    const double totV = t2 - t1 + tailV;
    vop1 += tailV;
    // the estimator's columns K V V_ext V_int E E_mu K/N V/N E/N (src/estimator.cpp:922)
    out9[0] = totK;
    out9[1] = totV;
    out9[2] = vop0;
    out9[3] = vop1;
    out9[4] = totK + totV;
    out9[5] = totK + totV - mu * Np;
    out9[6] = Np > 0 ? totK / Np : 0.0;
    out9[7] = Np > 0 ? totV / Np : 0.0;
    out9[8] = Np > 0 ? (totK + totV) / Np : 0.0;
    return PIMC_OK;
}

static int run_vnn(pimc_ctx* c, int n, int shift, const double vf[2], bool commit_in_kernel, double* out3,
                   double* per_bead, double* cor_out = nullptr) {
    int r;
    AzizDev az = ctx_aziz(c, &r);
    if (r) return r;
    VnnArgs A{};
    const bool inline_req = n <= kVnnInline;
    A.beads = inline_req ? nullptr : c->d_req_beads;
    A.n_beads = n;
    A.chunk = kVnnChunk;
    A.chunks = (c->desc.num_alloc + kVnnChunk - 1) / kVnnChunk;
    A.pos = c->d_pos;
    A.plane = c->plane;
    A.nalloc = c->desc.num_alloc;
    A.num_slices = c->desc.num_slices;
    A.on = c->d_on;
    A.cell = c->d_cell;
    A.label = c->d_label;
    A.numlab = c->d_numlab;
    A.hash = c->d_hash;
    A.commit_in_kernel = commit_in_kernel ? 1 : 0;
    A.box = c->box;
    A.worm = c->worm;
    A.shift = shift;
    A.vfactor0 = vf ? vf[0] : c->vf_default[0];
    A.vfactor1 = vf ? vf[1] : c->vf_default[1];
    // gradient correction of the bead actions (shift 1 only, as the reference)
    const bool cor = c->correction() && (shift == 1 || cor_out);
    A.cor_any_shift = cor_out ? 1 : 0;
    A.per_bead_cor = cor_out ? c->d_per_bead_cor : nullptr;
    A.gnn_part = nullptr;
    A.gnn_chunks = (c->desc.num_alloc + kGnnGroup - 1) / kGnnGroup;
    A.gvf0 = c->gvf[0];
    A.gvf1 = c->gvf[1];
    A.lambda = c->desc.lambda;
    if (cor) {
        const size_t need = (size_t)c->cap_req * 2 * A.gnn_chunks * 4;
        if ((size_t)c->cap_gnn < need) {
            int r2;
            if ((r2 = ctx_dalloc(c, &c->d_gnn_part, need))) return r2;
            c->cap_gnn = (int)need;
        }
        A.gnn_part = c->d_gnn_part;
    }
    A.tau = c->desc.tau;
    A.partials = c->d_vnn_partials;
    A.ticket = c->d_vnn_ticket;   // one kernel unless the gradient correction adds its three-body pass
    A.per_bead = c->d_per_bead;
    A.out = c->d_map + 8;
    A.done = c->d_done();
    A.seq = ++c->seq;
    VnnInline req;
    // short requests without the correction or per-bead outputs: one workgroup
    // per bead publishes its two sums; the weighted bead terms are added here in
    // bead order
    if (inline_req && A.chunks == 1 && !cor && !per_bead && !cor_out && out3) {
        std::memcpy(req.beads, c->h_req_beads, sizeof(VnnBead) * n);
        const KernelTimer kt = c->kernel_timer(1);
        // a packet on the device's AQL queue when the call is not timed (the
        // timing events live on the stream) and the context's stream is quiet:
        // the two never run work of one context at once
        const bool packet = !kt.start && c->direct_ready(c->desc.v_mode) && c->stream_quiet();
        if (packet) {
            VnnDirectOut* outp = c->d_vdir;
            const pimc::ArgPiece pieces[4] = {{&A, 0, (uint32_t)sizeof(VnnArgs)},
                                              {req.beads, pimc::kVdirOffReq, (uint32_t)(sizeof(VnnBead) * n)},
                                              {&az, pimc::kVdirOffAz, (uint32_t)sizeof(AzizDev)},
                                              {&outp, pimc::kVdirOffOut, (uint32_t)sizeof(outp)}};
            pimc::direct_dispatch(&c->dring, c->dk_vdir[c->desc.v_mode], pieces, 4, (uint32_t)n,
                                  (uint32_t)kVnnDirectThreads);
            ++c->direct_count;
        } else {
            HIP_OR_FAIL(c, launch_vnn_direct(c->desc.v_mode, c->stream, A, req, az, c->d_vdir, &kt));
        }
        if (c->wait_direct(n, A.seq, packet) != 0) return fail(c, PIMC_E_HIP, "delta action: results not published");
        const double tau_s = shift * c->desc.tau;
        double ao = 0.0, an = 0.0, du = 0.0;
        for (int k = 0; k < n; ++k) {
            const double co = ((req.beads[k].slice & 1) ? A.vfactor1 : A.vfactor0) * tau_s;
            const double vo = c->h_vdir[k].v_old, vn = c->h_vdir[k].v_new;
            ao += co * vo;
            an += co * vn;
            du += co * (vn - vo);
        }
        out3[0] = ao;
        out3[1] = an;
        out3[2] = du;
        return PIMC_OK;
    }
    if (inline_req) {
        std::memcpy(req.beads, c->h_req_beads, sizeof(VnnBead) * n);
    } else {
        HIP_OR_FAIL(c, hipMemcpyAsync(c->d_req_beads, c->h_req_beads, sizeof(VnnBead) * n, hipMemcpyHostToDevice,
                                      c->stream));
    }
    const KernelTimer kt = c->kernel_timer(1);
    const KernelTimer kg = cor ? c->kernel_timer(6) : KernelTimer{};
    const int gnn_lds = c->desc.num_alloc <= kLdsBeads ? c->desc.num_alloc : 0;
    HIP_OR_FAIL(c, launch_vnn(c->desc.v_mode, c->stream, A, req, az, &kt, cor ? c->d_gnn_part : nullptr, gnn_lds,
                              &kg));
    if (per_bead || cor_out) {
        if (per_bead)
            HIP_OR_FAIL(c, hipMemcpyAsync(c->h_per_bead, c->d_per_bead, sizeof(double) * 2 * n,
                                          hipMemcpyDeviceToHost, c->stream));
        if (cor_out)
            HIP_OR_FAIL(c, hipMemcpyAsync(c->h_per_bead_cor, c->d_per_bead_cor, sizeof(double) * 2 * n,
                                          hipMemcpyDeviceToHost, c->stream));
        HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
        if (cor_out)
            for (int k = 0; k < n; ++k) cor_out[k] = cor ? c->h_per_bead_cor[2 * k] : 0.0;
    } else if (c->wait_done(A.seq) != 0) {
        // the arrival counter may have been left part-way: clear it for the next call
        (void)hipMemsetAsync(c->d_vnn_ticket, 0, sizeof(unsigned), c->stream);
        return fail(c, PIMC_E_HIP, "delta action: completion not published");
    }
    if (out3)
        for (int k = 0; k < 3; ++k) out3[k] = c->h_map[8 + k];
    if (per_bead) std::memcpy(per_bead, c->h_per_bead, sizeof(double) * 2 * n);
    return PIMC_OK;
}

int pimc_action_list(pimc_ctx* c, const int32_t* sp, int n, int shift, const double vfactor[2], double* action,
                     double* per_bead) {
    if (!c || !action) return fail(c, PIMC_E_ARG, "null argument");
    if (!c->have_path) return fail(c, PIMC_E_STATE, "no path uploaded");
    if (shift < 1) return fail(c, PIMC_E_ARG, "shift must be >= 1");
    int r;
    if ((r = check_beads(c, sp, n))) return r;
    if (n == 0) {
        *action = 0.0;
        return PIMC_OK;
    }
    if ((r = ensure_req(c, n))) return r;
    for (int k = 0; k < n; ++k) {
        VnnBead& q = c->h_req_beads[k];
        q.walker = 0;
        q.slice = sp[2 * k];
        q.ptcl = sp[2 * k + 1];
        q.flags = 0;
        for (int d = 0; d < 3; ++d) q.old_pos[d] = q.new_pos[d] = 0.0;
    }
    double out3[3];
    std::vector<double> pb(per_bead ? 2 * (size_t)n : 0);
    if ((r = run_vnn(c, n, shift, vfactor, false, out3, per_bead ? pb.data() : nullptr))) return r;
    *action = out3[0];
    if (per_bead)
        for (int k = 0; k < n; ++k) per_bead[k] = pb[2 * k];
    return PIMC_OK;
}

int pimc_action_parts(pimc_ctx* c, const int32_t* sp, int n, int shift, const double vfactor[2], double* bare,
                      double* cor) {
    if (!c) return fail(c, PIMC_E_ARG, "null context");
    if (!c->have_path) return fail(c, PIMC_E_STATE, "no path uploaded");
    if (shift < 1) return fail(c, PIMC_E_ARG, "shift must be >= 1");
    int r;
    if ((r = check_beads(c, sp, n))) return r;
    if (n == 0 || (!bare && !cor)) return PIMC_OK;
    if ((r = ensure_req(c, n))) return r;
    for (int k = 0; k < n; ++k) {
        VnnBead& q = c->h_req_beads[k];
        q.walker = 0;
        q.slice = sp[2 * k];
        q.ptcl = sp[2 * k + 1];
        q.flags = 0;
        for (int d = 0; d < 3; ++d) q.old_pos[d] = q.new_pos[d] = 0.0;
    }
    std::vector<double> pb(2 * (size_t)n), cb(cor ? (size_t)n : 0);
    if ((r = run_vnn(c, n, shift, vfactor, false, nullptr, pb.data(), cor ? cb.data() : nullptr))) return r;
    const double tau_s = shift * c->desc.tau;
    if (bare)
        for (int k = 0; k < n; ++k) {
            const double vf = vfactor ? vfactor[sp[2 * k] & 1] : c->vf_default[sp[2 * k] & 1];
            bare[k] = vf * tau_s * pb[2 * k];
        }
    if (cor) std::memcpy(cor, cb.data(), sizeof(double) * n);
    return PIMC_OK;
}

int pimc_action_range(pimc_ctx* c, const int32_t start[2], const int32_t end[2], const int32_t* next_links,
                      int shift, const double vfactor[2], double* action) {
    if (!c || !start || !end || !action) return fail(c, PIMC_E_ARG, "null argument");
    const int M = c->desc.num_slices, N = c->desc.num_alloc;
    auto nxt = [&](int m, int i, int* nm, int* ni) {
        if (next_links) {
            const size_t t = (size_t)m * N + i;
            *nm = next_links[2 * t];
            *ni = next_links[2 * t + 1];
        } else {
            *nm = (m + 1) % M;
            *ni = i;
        }
    };
    if (start[0] < 0 || start[0] >= M || start[1] < 0 || start[1] >= N || end[0] < 0 || end[0] >= M || end[1] < 0 ||
        end[1] >= N)
        return fail(c, PIMC_E_ARG, "range start or end out of range");
    int sm, si;
    nxt(end[0], end[1], &sm, &si);
    std::vector<int32_t> list;
    int m = start[0], i = start[1];
    // a walk from start reaches end's successor within M*N steps or never
    const long limit = (long)M * N;
    long steps = 0;
    do {
        if (m < 0 || m >= M || i < 0 || i >= N) return fail(c, PIMC_E_ARG, "range walks off the path");
        // LocalAction skips inactive beads (src/action.cpp:532-541)
        if (c->h_on[(size_t)m * N + i]) {
            list.push_back(m);
            list.push_back(i);
        }
        int a, b;
        nxt(m, i, &a, &b);
        m = a;
        i = b;
        if (++steps > limit) return fail(c, PIMC_E_ARG, "range does not reach end's successor");
    } while (!(m == sm && i == si));
    return pimc_action_list(c, list.data(), (int)list.size() / 2, shift, vfactor, action, nullptr);
}

int pimc_delta_action(pimc_ctx* c, const int32_t* sp, const double* old_pos, const double* new_pos, int n, int shift,
                      const double vfactor[2], int commit, double* out3) {
    if (!c || !out3) return fail(c, PIMC_E_ARG, "null argument");
    if (!c->have_path) return fail(c, PIMC_E_STATE, "no path uploaded");
    if (shift < 1) return fail(c, PIMC_E_ARG, "shift must be >= 1");
    int r;
    if ((r = check_beads(c, sp, n))) return r;
    if (n == 0) {
        out3[0] = out3[1] = out3[2] = 0.0;
        return PIMC_OK;
    }
    if (!new_pos) return fail(c, PIMC_E_ARG, "null new positions");
    // one-pass evaluation needs the moved beads on distinct slices
    std::vector<char> seen(c->desc.num_slices, 0);
    for (int k = 0; k < n; ++k) {
        if (seen[sp[2 * k]]) return fail(c, PIMC_E_ARG, "delta_action beads must sit on distinct slices");
        seen[sp[2 * k]] = 1;
    }
    if ((r = ensure_req(c, n))) return r;
    for (int k = 0; k < n; ++k) {
        VnnBead& q = c->h_req_beads[k];
        q.walker = 0;
        q.slice = sp[2 * k];
        q.ptcl = sp[2 * k + 1];
        q.flags = 2 | (old_pos ? 1 : 0) | (commit ? 4 : 0);
        for (int d = 0; d < 3; ++d) {
            q.old_pos[d] = old_pos ? old_pos[3 * k + d] : 0.0;
            q.new_pos[d] = new_pos[3 * k + d];
        }
    }
    if (commit) ++c->path_version;
    return run_vnn(c, n, shift, vfactor, commit != 0, out3, nullptr);
}

int pimc_cell_dims(pimc_ctx* c, int32_t dims3[3], int32_t* hash_cap) {
    if (!c || !dims3) return fail(c, PIMC_E_ARG, "null argument");
    for (int i = 0; i < 3; ++i) dims3[i] = c->box.ngrid;
    if (hash_cap) *hash_cap = c->desc.num_alloc;
    return PIMC_OK;
}

int pimc_cell_lists(pimc_ctx* c, int32_t* grid, int32_t* bead_label, int32_t* num_labels, int32_t* hash) {
    if (!c) return fail(nullptr, PIMC_E_ARG, "null ctx");
    if (!c->have_path) return fail(c, PIMC_E_STATE, "no path uploaded");
    const int M = c->desc.num_slices, N = c->desc.num_alloc;
    const size_t MN = (size_t)M * N;
    const int ng = c->box.ngrid;
    if (grid) {
        std::vector<int> cell(MN);
        HIP_OR_FAIL(c, hipMemcpy(cell.data(), c->d_cell, sizeof(int) * MN, hipMemcpyDeviceToHost));
        for (size_t t = 0; t < MN; ++t) {
            const int cn = cell[t];
            if (cn < 0) {
                grid[3 * t] = grid[3 * t + 1] = grid[3 * t + 2] = -1;
            } else {
                grid[3 * t] = cn / (ng * ng);
                grid[3 * t + 1] = (cn / ng) % ng;
                grid[3 * t + 2] = cn % ng;
            }
        }
    }
    if (bead_label) HIP_OR_FAIL(c, hipMemcpy(bead_label, c->d_label, sizeof(int) * MN, hipMemcpyDeviceToHost));
    if (num_labels)
        HIP_OR_FAIL(c, hipMemcpy(num_labels, c->d_numlab, sizeof(int) * (size_t)c->cells * M, hipMemcpyDeviceToHost));
    if (hash) HIP_OR_FAIL(c, hipMemcpy(hash, c->d_hash, sizeof(int) * (size_t)c->cells * MN, hipMemcpyDeviceToHost));
    return PIMC_OK;
}

int pimc_interaction_list(pimc_ctx* c, int slice, int ptcl, int32_t* list2, double* sep3, int* count) {
    if (!c || !count) return fail(c, PIMC_E_ARG, "null argument");
    if (!c->have_path) return fail(c, PIMC_E_STATE, "no path uploaded");
    const int M = c->desc.num_slices, N = c->desc.num_alloc;
    if (slice < 0 || slice >= M || ptcl < 0 || ptcl >= N) return fail(c, PIMC_E_ARG, "bead out of range");
    if (!c->h_on[(size_t)slice * N + ptcl]) return fail(c, PIMC_E_ARG, "bead is not active");
    int* d_list = nullptr;
    double* d_sep = nullptr;
    int* d_cnt = nullptr;
    HIP_OR_FAIL(c, hipMalloc(&d_list, sizeof(int) * 2 * N));
    HIP_OR_FAIL(c, hipMalloc(&d_sep, sizeof(double) * 3 * N));
    HIP_OR_FAIL(c, hipMalloc(&d_cnt, sizeof(int)));
    hipError_t e = launch_interaction_list(c->stream, slice, ptcl, c->d_pos, c->plane, N, M, c->d_cell, c->d_numlab,
                                           c->d_hash, c->d_grid_nn, c->box, d_list, d_sep, d_cnt);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    int cnt = 0;
    if (e == hipSuccess) e = hipMemcpy(&cnt, d_cnt, sizeof(int), hipMemcpyDeviceToHost);
    if (e == hipSuccess && list2 && cnt > 0) e = hipMemcpy(list2, d_list, sizeof(int) * 2 * cnt, hipMemcpyDeviceToHost);
    if (e == hipSuccess && sep3 && cnt > 0) e = hipMemcpy(sep3, d_sep, sizeof(double) * 3 * cnt, hipMemcpyDeviceToHost);
    (void)hipFree(d_list);
    (void)hipFree(d_sep);
    (void)hipFree(d_cnt);
    if (e != hipSuccess) return fail(c, PIMC_E_HIP, std::string("interaction list: ") + hipGetErrorString(e));
    *count = cnt;
    return PIMC_OK;
}

int pimc_set_action_factors(pimc_ctx* c, const double vfactor[2], const double gradvfactor[2], int period) {
    if (!c || !vfactor || !gradvfactor) return fail(c, PIMC_E_ARG, "null argument");
    if (period < 1) return fail(c, PIMC_E_ARG, "period must be >= 1");
    if (!(gradvfactor[0] >= 0.0) || !(gradvfactor[1] >= 0.0)) return fail(c, PIMC_E_ARG, "negative gradVFactor");
    if ((gradvfactor[0] > kCorrEps || gradvfactor[1] > kCorrEps) && c->desc.num_alloc > 16384)
        return fail(c, PIMC_E_ARG, "the gradient correction holds an interaction list of at most 16384 beads");
    c->vf_default[0] = vfactor[0];
    c->vf_default[1] = vfactor[1];
    c->gvf[0] = gradvfactor[0];
    c->gvf[1] = gradvfactor[1];
    c->period = period;
    c->walker_valid = false;
    ++c->path_version;
    return PIMC_OK;
}

int pimc_set_timing(pimc_ctx* c, int enable) {
    if (!c) return fail(nullptr, PIMC_E_ARG, "null ctx");
    c->resolve_timing();
    c->timing = enable != 0;
    c->timing_kinds = ~0u;
    if (c->timing && !c->ensure_event_pool()) return fail(c, PIMC_E_HIP, "hipEventCreate failed");
    return PIMC_OK;
}

int pimc_set_timing_kinds(pimc_ctx* c, unsigned kinds) {
    if (!c) return fail(nullptr, PIMC_E_ARG, "null ctx");
    c->resolve_timing();
    c->timing = kinds != 0;
    c->timing_kinds = kinds;
    if (c->timing && !c->ensure_event_pool()) return fail(c, PIMC_E_HIP, "hipEventCreate failed");
    return PIMC_OK;
}

int pimc_kernel_time(pimc_ctx* c, const char* name, double* total_ms, long* count) {
    if (!c || !name) return fail(c, PIMC_E_ARG, "null argument");
    c->resolve_timing();
    if (std::strcmp(name, "delta_packets") == 0) {
        if (total_ms) *total_ms = 0.0;
        if (count) *count = c->direct_count;
        return PIMC_OK;
    }
    if (std::strcmp(name, "isf_chunks") == 0) {
        if (total_ms) *total_ms = 0.0;
        if (count) *count = c->isf_chunks;
        return PIMC_OK;
    }
    for (int k = 0; k < pimc_ctx::kKinds; ++k)
        if (std::strcmp(name, kKernelNames[k]) == 0) {
            if (total_ms) *total_ms = c->kernel_ms[k];
            if (count) *count = c->kernel_count[k];
            return PIMC_OK;
        }
    return fail(c, PIMC_E_ARG, "unknown kernel name");
}

int pimc_device_synchronize(void) {
    hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) return fail(nullptr, PIMC_E_HIP, hipGetErrorString(e));
    return PIMC_OK;
}

int pimc_host_alloc(size_t bytes, void** ptr) {
    if (!ptr || bytes == 0) return fail(nullptr, PIMC_E_ARG, "pimc_host_alloc: null pointer or zero bytes");
    *ptr = nullptr;
    hipError_t e = hipHostMalloc(ptr, bytes, hipHostMallocDefault);
    if (e != hipSuccess) {
        *ptr = nullptr;
        return fail(nullptr, PIMC_E_NOMEM, std::string("pimc_host_alloc: ") + hipGetErrorString(e));
    }
    return PIMC_OK;
}

int pimc_host_free(void* ptr) {
    if (!ptr) return PIMC_OK;
    hipError_t e = hipHostFree(ptr);
    if (e != hipSuccess) return fail(nullptr, PIMC_E_HIP, std::string("pimc_host_free: ") + hipGetErrorString(e));
    return PIMC_OK;
}

}  // extern "C"
