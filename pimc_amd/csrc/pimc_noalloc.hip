// pimc_noalloc.hip -- S(q) without any allocation: pimc_ssf_noalloc.
//
// pimc_ssf grows device buffers (q, rho_q(m), results) to the largest q list a
// context has seen.  This variant keeps all of that in one file-scope device
// array, `scratch`, of kScratchBytes per GPU, and walks the caller's q list in
// runs that fit it, so it never calls hipMalloc, hipFree or any other
// allocating routine (device or host).  It runs the same kernels in the same
// per-q order as pimc_ssf, so its results are bit-identical to pimc_ssf's.
//
// Every call on one GPU shares that GPU's array and leaves its contents behind
// (every word a call reads was written by the same call first): calls that
// overlap on one GPU, from different threads or different contexts' streams,
// race on it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>

#include "pimc_host.hpp"

namespace pimc {

constexpr size_t kScratchBytes = size_t(32) << 20;   // 32 MiB per GPU
constexpr int kScratchDevices = 64;

static __device__ __align__(256) char scratch[kScratchBytes];

// q [n][3] -> lattice components rint(q L / 2 pi), the host's classification
// (lattice_q) done again on the device with the same FP64 product and rounding.
__global__ void k_q_to_lattice(const double* __restrict__ q, int n, double f, int* __restrict__ qn) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < 3 * n) qn[t] = (int)rint(q[t] * f);
}

}  // namespace pimc

using namespace pimc;

namespace {

std::mutex g_scratch_mu;
void* g_scratch_ptr[kScratchDevices] = {};

int nfail(pimc_ctx* c, int code, const std::string& m) {
    if (c) {
        c->err = m;
        ++c->fail_count;
    }
    return code;
}

#define NA_HIP(c, expr)                                                                        \
    do {                                                                                       \
        hipError_t _e = (expr);                                                                \
        if (_e != hipSuccess) return nfail((c), PIMC_E_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

// The device address of `scratch` on device `dev` (current), looked up once per
// device.  In host code the array's name is only a stand-in for the symbol.
int scratch_for(pimc_ctx* c, int dev, char** out) {
    if (dev < 0 || dev >= kScratchDevices) return nfail(c, PIMC_E_UNSUPPORTED, "device index beyond the scratch table");
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return nfail(c, PIMC_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *out = static_cast<char*>(g_scratch_ptr[dev]);
    return PIMC_OK;
}

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// Scratch bytes of a run of n q-vectors over M slices: the q copy, the lattice
// components, rho_q(m) (lattice) or the per-slice pair sums (general), results.
size_t run_bytes(int n, int M) {
    return align256(sizeof(double) * 3 * (size_t)n) + align256(sizeof(int) * 3 * (size_t)n) +
           align256(sizeof(double) * 2 * (size_t)M * n) + align256(sizeof(double) * (size_t)n);
}

// One run of n consecutive q of one kind, q[0..n), results into sq[0..n).
int ssf_run(pimc_ctx* c, char* base, bool lattice, const double* q, int n, const int nmax[3], double inv_n,
            double* sq) {
    const int M = c->desc.num_slices;
    char* p = base;
    double* d_q = reinterpret_cast<double*>(p);
    p += align256(sizeof(double) * 3 * (size_t)n);
    int* d_qn = reinterpret_cast<int*>(p);
    p += align256(sizeof(int) * 3 * (size_t)n);
    double* d_rho = reinterpret_cast<double*>(p);
    p += align256(sizeof(double) * 2 * (size_t)M * n);
    double* d_out = reinterpret_cast<double*>(p);
    NA_HIP(c, hipMemcpyAsync(d_q, q, sizeof(double) * 3 * n, hipMemcpyHostToDevice, c->stream));
    if (lattice) {
        hipLaunchKernelGGL(k_q_to_lattice, dim3((unsigned)((3 * n + 255) / 256)), dim3(256), 0, c->stream, d_q, n,
                           c->desc.side / (2.0 * M_PI), d_qn);
        NA_HIP(c, hipGetLastError());
        RhoArgs A{};
        A.pos = c->d_pos;
        A.plane = c->plane;
        A.nalloc = c->desc.num_alloc;
        A.num_slices = M;
        A.nbs = c->d_nbs;
        A.invL = 1.0 / c->desc.side;
        A.nq = n;
        A.qn = d_qn;
        A.rho = d_rho;
        for (int d = 0; d < 3; ++d) A.nmax[d] = nmax[d];
        NA_HIP(c, launch_rho(true, c->stream, A));
        NA_HIP(c, launch_isf(c->stream, d_rho, n, M, inv_n, d_out, 1));
    } else {
        NA_HIP(c, launch_ssf_pairs(c->stream, c->d_pos, c->plane, c->desc.num_alloc, M, c->d_nbs, d_q, n, c->box,
                                   d_rho, inv_n, d_out));
    }
    NA_HIP(c, hipMemcpyAsync(sq, d_out, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    NA_HIP(c, hipStreamSynchronize(c->stream));
    return PIMC_OK;
}

int ssf_noalloc(pimc_ctx* c, const double* q, int nq, double* sq) {
    if (!c || (nq > 0 && (!q || !sq))) return nfail(c, PIMC_E_ARG, "null argument");
    if (!c->have_path) return nfail(c, PIMC_E_STATE, "no path uploaded");
    if (nq < 0) return nfail(c, PIMC_E_ARG, "negative q count");
    if (nq == 0) return PIMC_OK;
    const int M = c->desc.num_slices;
    if (M > 512) return nfail(c, PIMC_E_UNSUPPORTED, "S(q) supports M <= 512 slices");
    // q-vectors per run: as many as the scratch holds (at M <= 512, over 4000)
    int cap = (int)std::min<size_t>(kScratchBytes / (sizeof(double) * 2 * (size_t)M + 64), 1 << 20);
    while (cap > 1 && run_bytes(cap, M) > kScratchBytes) --cap;
    if (cap < 1 || run_bytes(1, M) > kScratchBytes)
        return nfail(c, PIMC_E_ARG, "one q-vector over M slices exceeds the 32 MiB scratch");
    long on = 0;
    for (auto v : c->h_on) on += v;
    const int Np = (int)(on / M);
    if (Np <= 0) return nfail(c, PIMC_E_STATE, "no active particles");
    const double inv_n = 1.0 / Np;
    char* base = nullptr;
    int r;
    if ((r = scratch_for(c, c->device, &base))) return r;
    // runs of consecutive q of one kind (lattice or general), at most cap long
    int k = 0;
    while (k < nq) {
        int n3[3];
        const bool lat = lattice_q(q + 3 * k, c->desc.side, n3);
        int nmax[3] = {0, 0, 0};
        int e = k;
        while (e < nq && e - k < cap) {
            int m3[3];
            if (lattice_q(q + 3 * e, c->desc.side, m3) != lat) break;
            if (lat)
                for (int d = 0; d < 3; ++d) nmax[d] = std::max(nmax[d], std::abs(m3[d]));
            ++e;
        }
        if ((r = ssf_run(c, base, lat, q + 3 * k, e - k, nmax, inv_n, sq + k))) return r;
        k = e;
    }
    return PIMC_OK;
}

}  // namespace

extern "C" int pimc_ssf_noalloc(pimc_ctx* c, const double* q, int nq, double* sq) {
    if (!c) return PIMC_E_ARG;
    int cur = 0;
    if (hipGetDevice(&cur) != hipSuccess) return nfail(c, PIMC_E_HIP, "hipGetDevice failed");
    if (cur != c->device && hipSetDevice(c->device) != hipSuccess)
        return nfail(c, PIMC_E_HIP, "hipSetDevice to the context's device failed");
    const int r = ssf_noalloc(c, q, nq, sq);
    if (cur != c->device) (void)hipSetDevice(cur);
    return r;
}
