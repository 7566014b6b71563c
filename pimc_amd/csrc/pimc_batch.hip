// pimc_batch.hip -- the batched form: the total potential action
// (LocalAction::potentialAction(), src/action.cpp:456-472) of many independent
// walkers with uneven N and M in one persistent launch.
//
// Every walker's slices are cut into row tiles of a fixed number of pairs per
// lane (the persistent pair kernel of pimc_pair.hip), so a walker's sums never
// depend on which other walkers share the batch.  A fixed grid of G workgroups
// walks contiguous tile ranges balanced by pair count; a workgroup restages a
// slice in LDS only when the slice changes.  Slice and walker sums follow in two
// fixed-order passes.  The three launches can be captured once into a hipGraph
// and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "pimc_host.hpp"

namespace pimc {

// One wave per slice: items of the slice (lanes strided), fixed shuffle tree.
__global__ __launch_bounds__(256) void k_slice_sums(int n_items, const SliceDesc* __restrict__ slices, int n_slices,
                                                    const double* __restrict__ partials,
                                                    double* __restrict__ v_slice) {
    const int sl = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (sl >= n_slices) return;
    const int b = slices[sl].first_item;
    const int e = (sl + 1 < n_slices) ? slices[sl + 1].first_item : n_items;
    double v = 0.0;
    for (int k = b + lane; k < e; k += 64) v += partials[k];
    v = wave_sum(v);
    if (lane == 0) v_slice[sl] = v;
}

// One wave per walker: vfactor[m%2] * shift*tau * V(m) over its slices.
__global__ __launch_bounds__(256) void k_walker_sums(const WalkerDesc* __restrict__ walkers, int n_walkers,
                                                     const double* __restrict__ v_slice, double* __restrict__ action) {
    const int w = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (w >= n_walkers) return;
    const WalkerDesc& W = walkers[w];
    const double tau_s = W.shift * W.tau;
    double tot = 0.0;
    for (int m = lane; m < W.num_slices; m += 64) {
        const double vf = (m & 1) ? W.vfactor1 : W.vfactor0;
        tot += vf * tau_s * (0.0 + v_slice[W.first_slice + m]);
    }
    tot = wave_sum(tot);
    if (lane == 0) action[w] = tot;
}

}  // namespace pimc

using namespace pimc;

struct pimc_batch {
    std::vector<pimc_desc> desc;
    std::vector<WalkerDesc> walkers;
    std::vector<int64_t> pos_off;     // per walker, in doubles
    std::vector<int> first_slice;
    int total_slices = 0;
    int64_t total_doubles = 0;
    int v_mode = PIMC_V_MIXED;
    int year = 1979;
    AzizDev az{};
    double max_sep_all = 0;
    // tiles
    int pairs_per_lane = 64;
    int grid = 0;                     // persistent workgroups (0: 4 per CU)
    int num_cus = 256;
    // device
    hipStream_t stream = nullptr;
    double* d_pos = nullptr;
    SliceDesc* d_slices = nullptr;
    WalkerDesc* d_walkers = nullptr;
    RowTile* d_items = nullptr;
    int* d_ranges = nullptr;
    double* d_partials = nullptr;
    double* d_vslice = nullptr;
    double* d_action = nullptr;
    int n_items = 0, lds_beads = 0, n_groups = 0;
    int64_t total_pairs = 0;
    std::vector<void*> allocs;
    std::vector<double> h_stage;
    // graph
    bool use_graph = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int last_shift = 1;
    double last_vf[2] = {1, 1};
    float last_ms = 0.f;
    std::string err;
    ~pimc_batch() {
        if (stream) (void)hipStreamSynchronize(stream);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (void* p : allocs) (void)hipFree(p);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

int bfail(pimc_batch* b, int code, const std::string& m) {
    if (b) b->err = m;
    return code;
}

#define BHIP(b, expr)                                                                                      \
    do {                                                                                                   \
        hipError_t _e = (expr);                                                                            \
        if (_e != hipSuccess) return bfail((b), PIMC_E_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

template <typename T>
int balloc(pimc_batch* b, T** p, size_t n) {
    hipError_t e = hipMalloc((void**)p, sizeof(T) * std::max<size_t>(n, 1));
    if (e != hipSuccess) return bfail(b, PIMC_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    b->allocs.push_back((void*)*p);
    return PIMC_OK;
}

void drop_graph(pimc_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

// Work list: per slice, row tiles of about pairs_per_lane pairs per lane; grid
// ranges balanced by pair count.
int build_work(pimc_batch* b) {
    const int W = (int)b->desc.size();
    std::vector<SliceDesc> sl(b->total_slices);
    std::vector<RowTile> items;
    int maxn = 0;
    b->total_pairs = 0;
    for (int w = 0; w < W; ++w) {
        const int M = b->desc[w].num_slices, N = b->desc[w].num_alloc;
        maxn = std::max(maxn, N);
        for (int m = 0; m < M; ++m) {
            const int si = b->first_slice[w] + m;
            SliceDesc& s = sl[si];
            s.n = N;
            s.slice = m;
            s.walker = w;
            s.first_item = (int)items.size();
            s.pos = b->pos_off[w] + (int64_t)m * N;
            s.plane = (int64_t)M * N;
            b->total_pairs += append_row_tiles(si, s, b->pairs_per_lane, items);
        }
    }
    b->n_items = (int)items.size();
    b->lds_beads = maxn <= kLdsBeads ? std::max(maxn, 1) : 0;
    int G = b->grid > 0 ? b->grid : pair_grid_for(b->lds_beads, b->num_cus);
    G = std::max(1, std::min(G, std::max(b->n_items, 1)));
    b->n_groups = G;
    std::vector<int> ranges;
    balance_row_tiles(items, G, ranges);
    for (void* p : {(void*)b->d_items, (void*)b->d_ranges, (void*)b->d_partials}) {
        if (!p) continue;
        (void)hipFree(p);
        b->allocs.erase(std::remove(b->allocs.begin(), b->allocs.end(), p), b->allocs.end());
    }
    int r;
    if ((r = balloc(b, &b->d_items, b->n_items))) return r;
    if ((r = balloc(b, &b->d_ranges, G + 1))) return r;
    if ((r = balloc(b, &b->d_partials, b->n_items))) return r;
    BHIP(b, hipMemcpyAsync(b->d_slices, sl.data(), sizeof(SliceDesc) * sl.size(), hipMemcpyHostToDevice, b->stream));
    BHIP(b, hipMemcpyAsync(b->d_items, items.data(), sizeof(RowTile) * items.size(), hipMemcpyHostToDevice,
                           b->stream));
    BHIP(b, hipMemcpyAsync(b->d_ranges, ranges.data(), sizeof(int) * ranges.size(), hipMemcpyHostToDevice,
                           b->stream));
    BHIP(b, hipStreamSynchronize(b->stream));
    drop_graph(b);
    return PIMC_OK;
}

hipError_t enqueue(pimc_batch* b, const AzizDev& az) {
    hipError_t e = launch_pair_action(b->v_mode, b->stream, b->d_items, b->n_items, b->d_ranges, b->n_groups,
                                      b->d_slices, b->total_slices, b->d_walkers, (int)b->desc.size(), b->d_pos, az,
                                      b->lds_beads, b->d_partials, nullptr, nullptr, nullptr, false);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_slice_sums, dim3((b->total_slices + 3) / 4), dim3(256), 0, b->stream, b->n_items,
                       b->d_slices, b->total_slices, b->d_partials, b->d_vslice);
    const int W = (int)b->desc.size();
    hipLaunchKernelGGL(k_walker_sums, dim3((W + 3) / 4), dim3(256), 0, b->stream, b->d_walkers, W, b->d_vslice,
                       b->d_action);
    return hipGetLastError();
}

int batch_aziz(pimc_batch* b, AzizDev* az) {
    *az = b->az;
    if (b->v_mode == PIMC_V_TABLE) {
        std::string err;
        int len = 0;
        for (auto& w : b->walkers) len = std::max(len, w.table_length);
        az->table = cached_table(b->year, 0, len, &err);
        if (!az->table) return bfail(b, PIMC_E_NOMEM, err);
    }
    return PIMC_OK;
}

int set_weights(pimc_batch* b, int shift, const double vf[2]) {
    bool changed = shift != b->last_shift || (vf && (vf[0] != b->last_vf[0] || vf[1] != b->last_vf[1]));
    if (!changed && b->walkers[0].shift == shift) return PIMC_OK;
    for (auto& w : b->walkers) {
        w.shift = shift;
        w.vfactor0 = vf ? vf[0] : 1.0;
        w.vfactor1 = vf ? vf[1] : 1.0;
    }
    b->last_shift = shift;
    b->last_vf[0] = vf ? vf[0] : 1.0;
    b->last_vf[1] = vf ? vf[1] : 1.0;
    BHIP(b, hipMemcpyAsync(b->d_walkers, b->walkers.data(), sizeof(WalkerDesc) * b->walkers.size(),
                           hipMemcpyHostToDevice, b->stream));
    return PIMC_OK;
}

}  // namespace

extern "C" {

int pimc_batch_create(const pimc_desc* walkers, int n_walkers, pimc_batch** out) {
    if (!walkers || n_walkers < 1 || !out) return PIMC_E_ARG;
    *out = nullptr;
    std::unique_ptr<pimc_batch> b(new pimc_batch());
    b->year = walkers[0].aziz_year;
    b->v_mode = walkers[0].v_mode;
    AzizParams p;
    if (!aziz_params(b->year, &p)) return PIMC_E_ARG;
    if (b->v_mode < 0 || b->v_mode > 2) return PIMC_E_ARG;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return PIMC_E_HIP;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) == hipSuccess) b->num_cus = prop.multiProcessorCount;
    int64_t off = 0;
    int fs = 0;
    double max_sep_all = 0;
    for (int w = 0; w < n_walkers; ++w) {
        const pimc_desc& d = walkers[w];
        if (d.num_slices < 1 || d.num_alloc < 1 || !(d.side > 0) || !(d.tau > 0) || !(d.rc > 0)) return PIMC_E_ARG;
        if (d.aziz_year != b->year) return PIMC_E_ARG;
        b->desc.push_back(d);
        WalkerDesc wd;
        std::memset(&wd, 0, sizeof(wd));
        wd.box.L = d.side;
        wd.box.invL = 1.0 / d.side;
        wd.box.rc2 = d.rc * d.rc;
        wd.box.ngrid = 1;
        wd.box.grid_size = d.side;
        const double h = d.side / 2.0;
        double ms = 0.0;
        ms += h * h;
        ms += h * h;
        ms += h * h;
        const double max_sep = std::sqrt(ms);
        max_sep_all = std::max(max_sep_all, max_sep);
        wd.table_length = int(max_sep / ((1.0E-6) * p.rm));
        wd.num_slices = d.num_slices;
        wd.first_slice = fs;
        wd.shift = 1;
        wd.vfactor0 = wd.vfactor1 = 1.0;
        wd.tau = d.tau;
        wd.worm = WormDev{{-1, -1}, {-1, -1}, {-1, -1}, {-1, -1}};
        b->walkers.push_back(wd);
        b->pos_off.push_back(off);
        b->first_slice.push_back(fs);
        off += 3 * (int64_t)d.num_slices * d.num_alloc;
        fs += d.num_slices;
    }
    b->total_slices = fs;
    b->total_doubles = off;
    b->max_sep_all = max_sep_all;
    b->az = make_aziz_dev(p, max_sep_all);
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) return PIMC_E_HIP;
    pimc_batch* raw = b.get();
    int r;
    if ((r = balloc(raw, &raw->d_pos, off))) return r;
    if ((r = balloc(raw, &raw->d_slices, fs))) return r;
    if ((r = balloc(raw, &raw->d_walkers, n_walkers))) return r;
    if ((r = balloc(raw, &raw->d_vslice, fs))) return r;
    if ((r = balloc(raw, &raw->d_action, n_walkers))) return r;
    if (hipMemsetAsync(raw->d_pos, 0, sizeof(double) * off, raw->stream) != hipSuccess) return PIMC_E_HIP;
    if (hipMemcpyAsync(raw->d_walkers, raw->walkers.data(), sizeof(WalkerDesc) * n_walkers, hipMemcpyHostToDevice,
                       raw->stream) != hipSuccess)
        return PIMC_E_HIP;
    if ((r = build_work(raw))) return r;
    *out = b.release();
    return PIMC_OK;
}

void pimc_batch_destroy(pimc_batch* b) { delete b; }

const char* pimc_batch_last_error(const pimc_batch* b) { return b ? b->err.c_str() : ""; }

int pimc_batch_upload(pimc_batch* b, int walker, const double* beads) {
    if (!b || !beads || walker < 0 || walker >= (int)b->desc.size()) return bfail(b, PIMC_E_ARG, "bad argument");
    const int M = b->desc[walker].num_slices, N = b->desc[walker].num_alloc;
    const size_t MN = (size_t)M * N;
    b->h_stage.resize(3 * MN);
    for (size_t t = 0; t < MN; ++t) {
        b->h_stage[t] = beads[3 * t];
        b->h_stage[t + MN] = beads[3 * t + 1];
        b->h_stage[t + 2 * MN] = beads[3 * t + 2];
    }
    BHIP(b, hipMemcpyAsync(b->d_pos + b->pos_off[walker], b->h_stage.data(), sizeof(double) * 3 * MN,
                           hipMemcpyHostToDevice, b->stream));
    BHIP(b, hipStreamSynchronize(b->stream));
    return PIMC_OK;
}

int pimc_batch_set_tiles(pimc_batch* b, int pairs_per_lane, int grid) {
    if (!b || pairs_per_lane < 1 || pairs_per_lane > 4096 || grid < 0) return bfail(b, PIMC_E_ARG, "bad tiles");
    b->pairs_per_lane = pairs_per_lane;
    b->grid = grid;
    return build_work(b);
}

int pimc_batch_use_graph(pimc_batch* b, int enable) {
    if (!b) return PIMC_E_ARG;
    b->use_graph = enable != 0;
    if (!b->use_graph) drop_graph(b);
    return PIMC_OK;
}

int pimc_batch_total_action(pimc_batch* b, int shift, const double vfactor[2], double* actions, double* v_slices) {
    if (!b || !actions || shift < 1) return bfail(b, PIMC_E_ARG, "bad argument");
    int r;
    if ((r = set_weights(b, shift, vfactor))) return r;
    AzizDev az;
    if ((r = batch_aziz(b, &az))) return r;
    // the two timing events are destroyed on every return path
    struct Events {
        hipEvent_t e0 = nullptr, e1 = nullptr;
        ~Events() {
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
        }
    } ev;
    BHIP(b, hipEventCreate(&ev.e0));
    BHIP(b, hipEventCreate(&ev.e1));
    hipEvent_t e0 = ev.e0, e1 = ev.e1;
    BHIP(b, hipEventRecord(e0, b->stream));
    if (b->use_graph) {
        if (!b->exec) {
            BHIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            hipError_t le = enqueue(b, az);
            hipGraph_t g = nullptr;
            hipError_t ce = hipStreamEndCapture(b->stream, &g);
            if (le != hipSuccess || ce != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                return bfail(b, PIMC_E_HIP, std::string("graph capture: ") + hipGetErrorString(le != hipSuccess ? le : ce));
            }
            b->graph = g;
            BHIP(b, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
        }
        BHIP(b, hipGraphLaunch(b->exec, b->stream));
    } else {
        BHIP(b, enqueue(b, az));
    }
    BHIP(b, hipEventRecord(e1, b->stream));
    const int W = (int)b->desc.size();
    BHIP(b, hipMemcpyAsync(actions, b->d_action, sizeof(double) * W, hipMemcpyDeviceToHost, b->stream));
    if (v_slices)
        BHIP(b, hipMemcpyAsync(v_slices, b->d_vslice, sizeof(double) * b->total_slices, hipMemcpyDeviceToHost,
                               b->stream));
    BHIP(b, hipStreamSynchronize(b->stream));
    (void)hipEventElapsedTime(&b->last_ms, e0, e1);
    return PIMC_OK;
}

int pimc_batch_info(const pimc_batch* b, int64_t* total_pairs, int* n_items, int* n_groups, double* last_ms) {
    if (!b) return PIMC_E_ARG;
    if (total_pairs) *total_pairs = b->total_pairs;
    if (n_items) *n_items = b->n_items;
    if (n_groups) *n_groups = b->n_groups;
    if (last_ms) *last_ms = b->last_ms;
    return PIMC_OK;
}

// Autotuner: time every (pairs_per_lane, grid) candidate `reps` times (graph
// replays, events on the batch stream), keep the fastest.  out4 = {best ppl,
// best grid, best ms, candidates tried}.
int pimc_batch_autotune(pimc_batch* b, int reps, double* out4) {
    if (!b || reps < 1) return bfail(b, PIMC_E_ARG, "bad argument");
    const int ppls[] = {4, 8, 16, 32, 64, 128};
    const int gmul[] = {2, 4, 8, 16};
    const bool graph = b->use_graph;
    b->use_graph = true;
    std::vector<double> act(b->desc.size());
    double best = 1e30;
    int best_p = b->pairs_per_lane, best_g = b->grid, tried = 0;
    const double vf[2] = {b->last_vf[0], b->last_vf[1]};
    for (int ppl : ppls)
        for (int gm : gmul) {
            int r;
            if ((r = pimc_batch_set_tiles(b, ppl, gm * b->num_cus))) return r;
            if ((r = pimc_batch_total_action(b, b->last_shift, vf, act.data(), nullptr))) return r;  // warm + capture
            double tot = 0.0;
            for (int k = 0; k < reps; ++k) {
                if ((r = pimc_batch_total_action(b, b->last_shift, vf, act.data(), nullptr))) return r;
                tot += b->last_ms;
            }
            ++tried;
            if (tot / reps < best) {
                best = tot / reps;
                best_p = ppl;
                best_g = gm * b->num_cus;
            }
        }
    b->use_graph = graph;
    int r = pimc_batch_set_tiles(b, best_p, best_g);
    if (out4) {
        out4[0] = best_p;
        out4[1] = best_g;
        out4[2] = best;
        out4[3] = tried;
    }
    return r;
}

}  // extern "C"
