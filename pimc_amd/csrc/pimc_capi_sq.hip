// pimc_capi_sq.hip -- host side of the structure-factor entry points and of the
// batched PotentialBase (Aziz) entry points of include/pimc_gpu.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <memory>
#include <vector>

#include "pimc_host.hpp"

using namespace pimc;

namespace {

int sfail(pimc_ctx* c, int code, const std::string& m) {
    if (c) {
        c->err = m;
        ++c->fail_count;
    }
    return code;
}

#define SHIP(c, expr)                                                                          \
    do {                                                                                       \
        hipError_t _e = (expr);                                                                \
        if (_e != hipSuccess) return sfail((c), PIMC_E_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

template <typename T>
int grow(pimc_ctx* c, T** p, size_t* cap, size_t need) {
    if (need <= *cap && *p) return PIMC_OK;
    if (*p) {
        (void)hipFree(*p);
        c->dev_allocs.erase(std::remove(c->dev_allocs.begin(), c->dev_allocs.end(), (void*)*p), c->dev_allocs.end());
        *p = nullptr;
    }
    size_t n = std::max<size_t>(need, 1);
    hipError_t e = hipMalloc((void**)p, sizeof(T) * n);
    if (e != hipSuccess) return sfail(c, PIMC_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    c->dev_allocs.push_back((void*)*p);
    *cap = n;
    return PIMC_OK;
}

int true_num_particles(pimc_ctx* c) {
    long on = 0;
    for (auto v : c->h_on) on += v;
    return (int)(on / c->desc.num_slices);
}

// rho_q(m) for a group of q-vectors into c->d_rho ([M][nq][2])
// Room for nq q-vectors in both device q lists.
int ensure_q(pimc_ctx* c, int nq) {
    int r;
    size_t capq = (size_t)c->cap_q;
    if ((size_t)nq * 3 > capq) {
        size_t cq = capq, cn = capq;
        if ((r = grow(c, &c->d_q, &cq, (size_t)nq * 3))) return r;
        if ((r = grow(c, &c->d_qn, &cn, (size_t)nq * 3))) return r;
        c->cap_q = (int)std::min(cq, cn);
    }
    return PIMC_OK;
}

// The density-mode arguments of the context's worldlines for nq q-vectors,
// modes into c->d_rho; for lattice q, qn (host, nq x 3) sets the phasor ranges.
RhoArgs rho_args(pimc_ctx* c, int nq, const int* qn) {
    RhoArgs A{};
    A.pos = c->d_pos;
    A.plane = c->plane;
    A.nalloc = c->desc.num_alloc;
    A.num_slices = c->desc.num_slices;
    A.nbs = c->d_nbs;
    A.invL = 1.0 / c->desc.side;
    A.nq = nq;
    A.rho = c->d_rho;
    if (qn) {
        int nm[3] = {0, 0, 0};
        for (int k = 0; k < nq; ++k)
            for (int d = 0; d < 3; ++d) nm[d] = std::max(nm[d], std::abs(qn[3 * k + d]));
        for (int d = 0; d < 3; ++d) A.nmax[d] = nm[d];
    }
    return A;
}

int compute_rho(pimc_ctx* c, bool lattice, const std::vector<int>& qn, const std::vector<double>& qv, int nq) {
    const int M = c->desc.num_slices;
    int r;
    if ((r = ensure_q(c, nq))) return r;
    if ((r = grow(c, &c->d_rho, &c->cap_rho, (size_t)M * nq * 2))) return r;
    RhoArgs A = rho_args(c, nq, lattice ? qn.data() : nullptr);
    if (lattice) {
        SHIP(c, hipMemcpyAsync(c->d_qn, qn.data(), sizeof(int) * 3 * nq, hipMemcpyHostToDevice, c->stream));
        A.qn = c->d_qn;
    } else {
        SHIP(c, hipMemcpyAsync(c->d_q, qv.data(), sizeof(double) * 3 * nq, hipMemcpyHostToDevice, c->stream));
        A.q = c->d_q;
    }
    c->timed_begin(2);
    SHIP(c, launch_rho(lattice, c->stream, A));
    c->timed_end();
    return PIMC_OK;
}

// F(q, tau) of a long lattice q list written straight into page-locked memory:
// the list runs in chunks, each chunk's density modes and lags on the context's
// stream, and each chunk's rows are copied out on a second stream while the
// next chunk computes (100 401 q x 250 lags make a 200 MB result whose copy
// takes ~6 ms after ~18 ms of kernels).  A q's modes and lags come from the same
// expressions whatever q share its launch, so the values are those of the
// one-launch path bit for bit.
// q per chunk: at 100 401 q x 250 lags, chunks of 34 000 / 25 600 / 18 432 q
// took 22.0 / 22.2 / 22.6 ms against 23.9 ms in one launch (interleaved), and
// 14 400 (4 q per lane in k_rho_lattice) 28.5 ms; chunks on two alternating
// streams with a mode buffer each were no faster (22.2 / 22.5 / 22.8 ms)
constexpr int kIsfChunkQ = 34000;                  // 8 q per lane in k_rho_lattice from M = 121 on
constexpr int kIsfChunkMinQ = 1024;                // shorter lists: one launch per kernel
constexpr size_t kIsfChunkMinBytes = 4u << 20;     // smaller results: one launch per kernel
constexpr int kIsfChunksMax = 64;

// q per chunk: lists of up to two chunks of kIsfChunkQ run as two halves
// (5377 q x 167 lags: 7.2 MB, whose copy is half the one-launch wall)
int isf_chunk_q(int n) {
    return n > 2 * kIsfChunkQ ? std::max(kIsfChunkQ, (n + kIsfChunksMax - 1) / kIsfChunksMax) : (n + 1) / 2;
}

bool isf_chunks_pay(pimc_ctx* c, int n, int M, const double* fqt) {
    if (n < 2 * kIsfChunkMinQ || (size_t)n * M * sizeof(double) < kIsfChunkMinBytes) return false;
    // a timed call keeps one launch per kernel, so its durations stay per call
    if (c->timing && (c->timing_kinds & ((1u << 2) | (1u << 4)))) return false;
    // a pageable destination is copied through staging with the host waiting
    hipPointerAttribute_t at{};
    const hipError_t e = hipPointerGetAttributes(&at, fqt);
    (void)hipGetLastError();   // a pageable pointer may report an error here
    return e == hipSuccess && at.type == hipMemoryTypeHost;
}

int isf_lattice_chunked(pimc_ctx* c, const std::vector<int>& qn, int n, int M, double inv_n, double* fqt) {
    const int per = isf_chunk_q(n);
    const int chunks = (n + per - 1) / per;
    int r;
    if ((r = ensure_q(c, n))) return r;
    if ((r = grow(c, &c->d_rho, &c->cap_rho, (size_t)M * per * 2))) return r;
    if ((r = grow(c, &c->d_fqt, &c->cap_fqt, (size_t)n * M))) return r;
    if (!c->copy_stream) SHIP(c, hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking));
    while ((int)c->chunk_ev.size() < chunks) {
        hipEvent_t ev = nullptr;
        SHIP(c, hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        c->chunk_ev.push_back(ev);
    }
    SHIP(c, hipMemcpyAsync(c->d_qn, qn.data(), sizeof(int) * 3 * (size_t)n, hipMemcpyHostToDevice, c->stream));
    hipError_t e = hipSuccess;
    for (int k = 0; k < chunks && e == hipSuccess; ++k) {
        const int off = k * per, nc = std::min(per, n - off);
        RhoArgs A = rho_args(c, nc, qn.data() + 3 * (size_t)off);
        A.qn = c->d_qn + 3 * (size_t)off;
        double* rows = c->d_fqt + (size_t)off * M;
        // the chunk's modes overwrite d_rho after the previous chunk's lags (one stream)
        e = launch_rho(true, c->stream, A);
        if (e == hipSuccess) e = launch_isf(c->stream, c->d_rho, nc, M, inv_n, rows, 0);
        if (e == hipSuccess) e = hipEventRecord(c->chunk_ev[k], c->stream);
        if (e == hipSuccess) e = hipStreamWaitEvent(c->copy_stream, c->chunk_ev[k], 0);
        if (e == hipSuccess)
            e = hipMemcpyAsync(fqt + (size_t)off * M, rows, sizeof(double) * (size_t)nc * M, hipMemcpyDeviceToHost,
                               c->copy_stream);
        if (e == hipSuccess) ++c->isf_chunks;
    }
    // both streams end quiet, after an error too (no copy may land later)
    const hipError_t e1 = hipStreamSynchronize(c->copy_stream);
    const hipError_t e2 = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = e1 != hipSuccess ? e1 : e2;
    if (e != hipSuccess) return sfail(c, PIMC_E_HIP, std::string("F(q,tau) in chunks: ") + hipGetErrorString(e));
    return PIMC_OK;
}

}  // namespace

namespace pimc {

bool lattice_q(const double* q, double L, int* n3) {
    const double f = L / (2.0 * M_PI);
    for (int d = 0; d < 3; ++d) {
        const double x = q[d] * f;
        const double r = std::nearbyint(x);
        if (std::fabs(r) > rho_max_nmax()) return false;
        const double back = r * 2.0 * M_PI / L;
        if (std::fabs(back - q[d]) > 1e-9 * std::max(1.0, std::fabs(q[d]))) return false;
        n3[d] = (int)r;
    }
    return true;
}

}  // namespace pimc

extern "C" {

int pimc_ssf(pimc_ctx* c, const double* q, int nq, double* sq) {
    if (!c || (nq > 0 && (!q || !sq))) return sfail(c, PIMC_E_ARG, "null argument");
    if (!c->have_path) return sfail(c, PIMC_E_STATE, "no path uploaded");
    if (nq <= 0) return PIMC_OK;
    const int M = c->desc.num_slices;
    if (M > 512) return sfail(c, PIMC_E_UNSUPPORTED, "S(q) supports M <= 512 slices");
    const int Np = true_num_particles(c);
    if (Np <= 0) return sfail(c, PIMC_E_STATE, "no active particles");
    const double inv_n = 1.0 / Np;
    std::vector<int> idx_l, idx_g, qn;
    std::vector<double> qg;
    for (int k = 0; k < nq; ++k) {
        int n3[3];
        if (lattice_q(q + 3 * k, c->desc.side, n3)) {
            idx_l.push_back(k);
            qn.insert(qn.end(), n3, n3 + 3);
        } else {
            idx_g.push_back(k);
            qg.insert(qg.end(), q + 3 * k, q + 3 * k + 3);
        }
    }
    int r;
    std::vector<double> host;
    if (!idx_l.empty()) {
        const int nl = (int)idx_l.size();
        if ((r = compute_rho(c, true, qn, qg, nl))) return r;
        if ((r = grow(c, &c->d_fqt, &c->cap_fqt, (size_t)nl))) return r;
        c->timed_begin(3);
        SHIP(c, launch_isf(c->stream, c->d_rho, nl, M, inv_n, c->d_fqt, 1));
        c->timed_end();
        host.resize(nl);
        SHIP(c, hipMemcpyAsync(host.data(), c->d_fqt, sizeof(double) * nl, hipMemcpyDeviceToHost, c->stream));
        SHIP(c, hipStreamSynchronize(c->stream));
        for (int k = 0; k < nl; ++k) sq[idx_l[k]] = host[k];
    }
    if (!idx_g.empty()) {
        const int ng = (int)idx_g.size();
        size_t capq = (size_t)c->cap_q;
        if ((size_t)ng * 3 > capq) {
            size_t cq = capq, cn = capq;
            if ((r = grow(c, &c->d_q, &cq, (size_t)ng * 3))) return r;
            if ((r = grow(c, &c->d_qn, &cn, (size_t)ng * 3))) return r;
            c->cap_q = (int)std::min(cq, cn);
        }
        if ((r = grow(c, &c->d_rho, &c->cap_rho, (size_t)ng * M))) return r;
        if ((r = grow(c, &c->d_fqt, &c->cap_fqt, (size_t)ng))) return r;
        SHIP(c, hipMemcpyAsync(c->d_q, qg.data(), sizeof(double) * 3 * ng, hipMemcpyHostToDevice, c->stream));
        c->timed_begin(3);
        SHIP(c, launch_ssf_pairs(c->stream, c->d_pos, c->plane, c->desc.num_alloc, M, c->d_nbs, c->d_q, ng, c->box,
                                 c->d_rho, inv_n, c->d_fqt));
        c->timed_end();
        host.resize(ng);
        SHIP(c, hipMemcpyAsync(host.data(), c->d_fqt, sizeof(double) * ng, hipMemcpyDeviceToHost, c->stream));
        SHIP(c, hipStreamSynchronize(c->stream));
        for (int k = 0; k < ng; ++k) sq[idx_g[k]] = host[k];
    }
    return PIMC_OK;
}

int pimc_isf(pimc_ctx* c, const double* q, int nq, double* fqt) {
    if (!c || (nq > 0 && (!q || !fqt))) return sfail(c, PIMC_E_ARG, "null argument");
    if (!c->have_path) return sfail(c, PIMC_E_STATE, "no path uploaded");
    if (nq <= 0) return PIMC_OK;
    const int M = c->desc.num_slices;
    if (M > 512) return sfail(c, PIMC_E_UNSUPPORTED, "F(q,tau) supports M <= 512 slices");
    const int Np = true_num_particles(c);
    if (Np <= 0) return sfail(c, PIMC_E_STATE, "no active particles");
    const double inv_n = 1.0 / Np;
    for (int pass = 0; pass < 2; ++pass) {
        std::vector<int> idx, qn;
        std::vector<double> qg;
        for (int k = 0; k < nq; ++k) {
            int n3[3];
            const bool lat = lattice_q(q + 3 * k, c->desc.side, n3);
            if (lat != (pass == 0)) continue;
            idx.push_back(k);
            if (lat) qn.insert(qn.end(), n3, n3 + 3);
            else qg.insert(qg.end(), q + 3 * k, q + 3 * k + 3);
        }
        if (idx.empty()) continue;
        const int n = (int)idx.size();
        int r;
        if (pass == 0 && n == nq && isf_chunks_pay(c, n, M, fqt)) {
            if ((r = isf_lattice_chunked(c, qn, n, M, inv_n, fqt))) return r;
            continue;
        }
        if ((r = compute_rho(c, pass == 0, qn, qg, n))) return r;
        if ((r = grow(c, &c->d_fqt, &c->cap_fqt, (size_t)n * M))) return r;
        c->timed_begin(4);
        SHIP(c, launch_isf(c->stream, c->d_rho, n, M, inv_n, c->d_fqt, 0));
        c->timed_end();
        if (n == nq) {
            // every q in this pass, in order: straight into the caller's array
            SHIP(c, hipMemcpyAsync(fqt, c->d_fqt, sizeof(double) * n * M, hipMemcpyDeviceToHost, c->stream));
            SHIP(c, hipStreamSynchronize(c->stream));
            continue;
        }
        std::unique_ptr<double[]> host(new double[(size_t)n * M]);
        SHIP(c, hipMemcpyAsync(host.get(), c->d_fqt, sizeof(double) * n * M, hipMemcpyDeviceToHost, c->stream));
        SHIP(c, hipStreamSynchronize(c->stream));
        for (int k = 0; k < n; ++k) std::memcpy(fqt + (size_t)idx[k] * M, host.get() + (size_t)k * M, sizeof(double) * M);
    }
    return PIMC_OK;
}

}  // extern "C"

// ===========================================================================
// Batched PotentialBase for the Aziz pair potential
// ===========================================================================
struct pimc_aziz {
    AzizParams p{};
    AzizDev az{};
    int v_mode = PIMC_V_FP32;
    int year = 1979;
    double tailV = 0;
    hipStream_t stream = nullptr;
    double* d_in = nullptr;
    double* d_out = nullptr;
    size_t cap = 0;
    std::string err;
};

namespace {

int afail(pimc_aziz* a, int code, const std::string& m) {
    if (a) a->err = m;
    return code;
}

int aziz_run(pimc_aziz* a, const double* sep, double* out, int n, int what) {
    if (!a || (n > 0 && (!sep || !out))) return afail(a, PIMC_E_ARG, "null argument");
    if (n < 0) return afail(a, PIMC_E_ARG, "negative count");
    if (n == 0) return PIMC_OK;
    const size_t need = 3 * (size_t)n;
    if (need > a->cap) {
        if (a->d_in) (void)hipFree(a->d_in);
        if (a->d_out) (void)hipFree(a->d_out);
        a->d_in = a->d_out = nullptr;
        if (hipMalloc(&a->d_in, sizeof(double) * need) != hipSuccess ||
            hipMalloc(&a->d_out, sizeof(double) * need) != hipSuccess) {
            a->cap = 0;
            return afail(a, PIMC_E_NOMEM, "hipMalloc failed");
        }
        a->cap = need;
    }
    std::string err;
    AzizDev az = a->az;
    const double* d1 = nullptr;
    const double* d2 = nullptr;
    if (what == 0 && a->v_mode == PIMC_V_TABLE) {
        az.table = cached_table(a->year, 0, az.table_length, &err);
        if (!az.table) return afail(a, PIMC_E_NOMEM, err);
    }
    if (what == 1) {
        d1 = cached_table(a->year, 1, az.table_length, &err);
        if (!d1) return afail(a, PIMC_E_NOMEM, err);
    }
    if (what == 2) {
        d2 = cached_table(a->year, 2, az.table_length, &err);
        if (!d2) return afail(a, PIMC_E_NOMEM, err);
    }
    hipError_t e = hipMemcpyAsync(a->d_in, sep, sizeof(double) * need, hipMemcpyHostToDevice, a->stream);
    if (e == hipSuccess) e = launch_aziz_batch(a->v_mode, a->stream, a->d_in, n, az, d1, d2, what, a->d_out);
    const size_t outn = (what == 1) ? need : (size_t)n;
    if (e == hipSuccess) e = hipMemcpyAsync(out, a->d_out, sizeof(double) * outn, hipMemcpyDeviceToHost, a->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(a->stream);
    if (e != hipSuccess) return afail(a, PIMC_E_HIP, hipGetErrorString(e));
    return PIMC_OK;
}

}  // namespace

extern "C" {

int pimc_aziz_create(int year, double max_sep, double rc, int v_mode, pimc_aziz** out) {
    if (!out) return PIMC_E_ARG;
    *out = nullptr;
    pimc_aziz* a = new pimc_aziz();
    if (!aziz_params(year, &a->p) || !(max_sep > 0) || !(rc > 0) || v_mode < 0 || v_mode > 2) {
        delete a;
        return PIMC_E_ARG;
    }
    a->year = year;
    a->v_mode = v_mode;
    a->az = make_aziz_dev(a->p, max_sep);
    a->tailV = aziz_tail(a->p, rc);
    if (hipStreamCreateWithFlags(&a->stream, hipStreamNonBlocking) != hipSuccess) {
        delete a;
        return PIMC_E_HIP;
    }
    *out = a;
    return PIMC_OK;
}

void pimc_aziz_destroy(pimc_aziz* a) {
    if (!a) return;
    if (a->stream) (void)hipStreamSynchronize(a->stream);
    if (a->d_in) (void)hipFree(a->d_in);
    if (a->d_out) (void)hipFree(a->d_out);
    if (a->stream) (void)hipStreamDestroy(a->stream);
    delete a;
}

int pimc_aziz_V(pimc_aziz* a, const double* sep3, double* out, int n) { return aziz_run(a, sep3, out, n, 0); }
int pimc_aziz_gradV(pimc_aziz* a, const double* sep3, double* out3, int n) { return aziz_run(a, sep3, out3, n, 1); }
int pimc_aziz_grad2V(pimc_aziz* a, const double* sep3, double* out, int n) { return aziz_run(a, sep3, out, n, 2); }
double pimc_aziz_tailV(const pimc_aziz* a) { return a ? a->tailV : 0.0; }
int pimc_aziz_table_length(const pimc_aziz* a) { return a ? a->az.table_length : 0; }

}  // extern "C"
