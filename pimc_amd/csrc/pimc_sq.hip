// pimc_sq.hip -- density modes, S(q) and F(q,tau) for gfx950.
//
// The reference's estimators are direct pair sums:
//   S:  sf(q)      = sum_m [N_m + 2 sum_{i<j} cos(q . r_ij)]          (src/estimator.cpp:3705-3737)
//   F:  isf(q,tau) = sum_m sum_{i in m} sum_{j in m+tau} cos(q.r_i - q.r_j)  (:3923-3961)
// Both equal sums over the density mode rho_q(m) = sum_i exp(i q . r_i(m)):
//   sf(q) = sum_m |rho_q(m)|^2,   isf(q,tau) = sum_m Re rho_q(m) conj(rho_q(m+tau)),
// exactly for F and, for S, whenever q is a reciprocal-lattice vector of the
// periodic box (then the minimum image drops out of the phase).  The kernels
// build rho_q(m) once (O(n_q M N)) and then correlate in imaginary time
// (O(n_q M^2)) instead of O(n_q M^2 N^2) trigonometric pair terms.
//
// Lattice q = (2 pi / L) n: exp(i q.r) = prod_d exp(2 pi i n_d s_d), s_d = x_d / L.
// Each workgroup stages per-bead axis phasors exp(2 pi i n s_d) for -nmax <= n <=
// nmax in LDS (phase reduced in FP64, FP32 sincospi; the negative n are the
// conjugates), then each lane accumulates a run of q-vectors with FP32 complex
// products and FP64 per-chunk sums; the X Y product of a row of q sharing
// (n_x, n_y) is formed once per bead (k_rho_lattice).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>

#include "pimc_kernels.hpp"

namespace pimc {

constexpr int kRhoThreads = 256;
constexpr int kRhoBeadChunk = 32;  // beads per LDS phasor stage
constexpr int kMaxNmax = 40;       // largest |n_d| of a lattice q
constexpr int kReseed = 4;         // exact sincospi every kReseed powers, products in between

// rho_q(m) for lattice q.  Lane t of a workgroup takes q = base + t + 256 k
// (k < QPT), so a wave walks consecutive q: mostly one (n_x, n_y) and consecutive
// n_z, which makes the x/y phasor reads LDS broadcasts and the z reads
// consecutive words.  Tables are bead-major: ph[axis][bead][n + nmax].
//
// Rows: consecutive q of the list that share (n_x, n_y) -- in the reference's
// order a q set is runs of n_z -- share the product X(n_x) Y(n_y) of each bead.
// A workgroup first numbers the rows its q touch (a block scan over the q
// range); when they fit the launch's row capacity (rows_cap, at most
// kRhoRowsMax, sized with the q per workgroup), each bead chunk forms every row's product
// once into LDS and a lane's term per (q, bead) is one complex product with
// Z(n_z), 4 FMA instead of 8.  The products are formed by the same FP32
// expressions as in the per-lane form, and the FP32 partial of every 32 beads
// is added into the FP64 sum either way, so both forms give the same bits.
constexpr int kRhoRowsMax = 128;   // rows of one workgroup's q range that share products

template <int QPT>
__global__ __launch_bounds__(kRhoThreads) void k_rho_lattice(RhoArgs A, int rows_cap) {
    extern __shared__ float2 ph[];
    const int m = blockIdx.y;
    const int n_beads = A.nbs[m];
    const int h0 = A.nmax[0], h1 = A.nmax[1], h2 = A.nmax[2];
    const int w0 = 2 * h0 + 1, w1 = 2 * h1 + 1, w2 = 2 * h2 + 1;
    float2* phx = ph;
    float2* phy = phx + w0 * kRhoBeadChunk;
    float2* phz = phy + w1 * kRhoBeadChunk;
    float2* xy = phz + w2 * kRhoBeadChunk;                                // [rows_cap][kRhoBeadChunk]
    int2* rows = reinterpret_cast<int2*>(xy + rows_cap * kRhoBeadChunk);  // [rows_cap] (n_x, n_y)
    int* scan = reinterpret_cast<int*>(xy);                                // [kRhoThreads * QPT], before xy is used
    __shared__ int wave_tot[kRhoThreads / 64];
    const double* px = A.pos + (int64_t)m * A.nalloc;
    const double* py = px + A.plane;
    const double* pz = py + A.plane;

    const int q0 = blockIdx.x * kRhoThreads * QPT;
    const int qbase = q0 + threadIdx.x;
    int ix[QPT], iy[QPT], iz[QPT], ro[QPT];
#pragma unroll
    for (int k = 0; k < QPT; ++k) {
        const int qi = qbase + k * kRhoThreads;
        int nx = 0, ny = 0, nz = 0;
        if (qi < A.nq) {
            nx = A.qn[3 * qi];
            ny = A.qn[3 * qi + 1];
            nz = A.qn[3 * qi + 2];
        }
        ix[k] = nx + h0;
        iy[k] = ny + h1;
        iz[k] = nz + h2;
        ro[k] = 0;
    }
    bool use_rows = false;
    int n_rows = 0;
    if (rows_cap > 0) {
#pragma unroll
        for (int k = 0; k < QPT; ++k) {
            const int qi = qbase + k * kRhoThreads;
            const int start = qi < A.nq &&
                              (qi == q0 || ix[k] - h0 != A.qn[3 * qi - 3] || iy[k] - h1 != A.qn[3 * qi - 2]);
            scan[threadIdx.x + k * kRhoThreads] = start;
        }
        __syncthreads();
        // inclusive prefix of the row starts over the workgroup's q range: lane t
        // owns entries [t QPT, (t + 1) QPT), then a wave scan and the waves in order
        int own[QPT];
        int run = 0;
#pragma unroll
        for (int k = 0; k < QPT; ++k) {
            run += scan[threadIdx.x * QPT + k];
            own[k] = run;
        }
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        int incl = run;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int v = __shfl_up(incl, off, 64);
            if (lane >= off) incl += v;
        }
        if (lane == 63) wave_tot[wave] = incl;
        __syncthreads();
        int before = incl - run;
        for (int w = 0; w < wave; ++w) before += wave_tot[w];
        n_rows = wave_tot[0] + wave_tot[1] + wave_tot[2] + wave_tot[3];
        __syncthreads();   // every lane has read its entries
#pragma unroll
        for (int k = 0; k < QPT; ++k) scan[threadIdx.x * QPT + k] = before + own[k];
        __syncthreads();
        use_rows = n_rows <= rows_cap;
#pragma unroll
        for (int k = 0; k < QPT; ++k) {
            const int l = threadIdx.x + k * kRhoThreads;
            const int r = scan[l] - 1;   // the row of q_l (starts counted up to and with it)
            const bool first = l == 0 || scan[l - 1] != scan[l];
            ro[k] = (use_rows && r >= 0 && r < rows_cap) ? r : 0;
            if (use_rows && first && qbase + k * kRhoThreads < A.nq && r >= 0 && r < rows_cap)
                rows[r] = make_int2(ix[k] - h0, iy[k] - h1);
        }
    }
    double are[QPT], aim[QPT];
#pragma unroll
    for (int k = 0; k < QPT; ++k) are[k] = aim[k] = 0.0;

    for (int b0 = 0; b0 < n_beads; b0 += kRhoBeadChunk) {
        const int nb = min(kRhoBeadChunk, n_beads - b0);
        __syncthreads();
        // one lane per (axis, bead) row: exp(2 pi i n s) for n = 0..nmax, and the
        // conjugates at -n
        for (int r = threadIdx.x; r < 3 * kRhoBeadChunk; r += kRhoThreads) {
            const int axis = r / kRhoBeadChunk, bb = r % kRhoBeadChunk;
            const int h = axis == 0 ? h0 : (axis == 1 ? h1 : h2);
            float2* row = (axis == 0 ? phx : (axis == 1 ? phy : phz)) + bb * (2 * h + 1) + h;
            if (bb >= nb) {
                for (int n = -h; n <= h; ++n) row[n] = make_float2(0.f, 0.f);
                continue;
            }
            const double x = (axis == 0 ? px : (axis == 1 ? py : pz))[b0 + bb];
            const double sfrac = x * A.invL;           // position in box units
            float s1, c1;
            {
                const double red = sfrac - rint(sfrac);
                sincospif((float)(2.0 * red), &s1, &c1);
            }
            float cr = 1.f, ci = 0.f;
            for (int n = 0; n <= h; ++n) {
                if (n % kReseed == 0) {
                    const double t = (double)n * sfrac;
                    const double red = t - rint(t);
                    sincospif((float)(2.0 * red), &ci, &cr);
                } else {
                    const float nr = fmaf(cr, c1, -ci * s1);
                    const float ni = fmaf(cr, s1, ci * c1);
                    cr = nr;
                    ci = ni;
                }
                row[n] = make_float2(cr, ci);
                row[-n] = make_float2(cr, -ci);
            }
        }
        __syncthreads();
        float sre[QPT], sim[QPT];
#pragma unroll
        for (int k = 0; k < QPT; ++k) sre[k] = sim[k] = 0.f;
        if (use_rows) {
            // every row's X Y product of the chunk's beads, once
            for (int t = threadIdx.x; t < n_rows * kRhoBeadChunk; t += kRhoThreads) {
                const int r = t / kRhoBeadChunk, bb = t % kRhoBeadChunk;
                const int2 rn = rows[r];
                const float2 ex = phx[bb * w0 + rn.x + h0];
                const float2 ey = phy[bb * w1 + rn.y + h1];
                const float xr = fmaf(ex.x, ey.x, -ex.y * ey.y);
                const float xi = fmaf(ex.x, ey.y, ex.y * ey.x);
                xy[t] = make_float2(xr, xi);
            }
            __syncthreads();
            for (int bb = 0; bb < nb; ++bb) {
#pragma unroll
                for (int k = 0; k < QPT; ++k) {
                    const float2 p = xy[ro[k] * kRhoBeadChunk + bb];
                    const float2 ez = phz[bb * w2 + iz[k]];
                    sre[k] = fmaf(p.x, ez.x, sre[k]);
                    sre[k] = fmaf(-p.y, ez.y, sre[k]);
                    sim[k] = fmaf(p.x, ez.y, sim[k]);
                    sim[k] = fmaf(p.y, ez.x, sim[k]);
                }
            }
        } else {
            for (int bb = 0; bb < nb; ++bb) {
#pragma unroll
                for (int k = 0; k < QPT; ++k) {
                    const float2 ex = phx[bb * w0 + ix[k]];
                    const float2 ey = phy[bb * w1 + iy[k]];
                    const float2 ez = phz[bb * w2 + iz[k]];
                    const float xr = fmaf(ex.x, ey.x, -ex.y * ey.y);
                    const float xi = fmaf(ex.x, ey.y, ex.y * ey.x);
                    sre[k] = fmaf(xr, ez.x, sre[k]);
                    sre[k] = fmaf(-xi, ez.y, sre[k]);
                    sim[k] = fmaf(xr, ez.y, sim[k]);
                    sim[k] = fmaf(xi, ez.x, sim[k]);
                }
            }
        }
#pragma unroll
        for (int k = 0; k < QPT; ++k) {
            are[k] += (double)sre[k];
            aim[k] += (double)sim[k];
        }
    }
#pragma unroll
    for (int k = 0; k < QPT; ++k) {
        const int qi = qbase + k * kRhoThreads;
        if (qi < A.nq) {
            A.rho[((int64_t)m * A.nq + qi) * 2 + 0] = are[k];
            A.rho[((int64_t)m * A.nq + qi) * 2 + 1] = aim[k];
        }
    }
}

// General q: rho_q(m) = sum_i exp(i q . r_i) with the phase reduced in FP64.
__global__ __launch_bounds__(kRhoThreads) void k_rho_general(RhoArgs A) {
    const int m = blockIdx.y;
    const int qi = blockIdx.x * kRhoThreads + threadIdx.x;
    if (qi >= A.nq) return;
    const double qx = A.q[3 * qi], qy = A.q[3 * qi + 1], qz = A.q[3 * qi + 2];
    const double* px = A.pos + (int64_t)m * A.nalloc;
    const double* py = px + A.plane;
    const double* pz = py + A.plane;
    const int n = A.nbs[m];
    double re = 0.0, im = 0.0;
    constexpr double kInv2Pi = 0.15915494309189533577;
    for (int b0 = 0; b0 < n; b0 += 64) {
        float sre = 0.f, sim = 0.f;
        const int e = min(n, b0 + 64);
        for (int b = b0; b < e; ++b) {
            const double phase = qx * px[b] + qy * py[b] + qz * pz[b];
            const double rev = phase * kInv2Pi;
            const double red = rev - rint(rev);
            float s, c;
            sincospif((float)(2.0 * red), &s, &c);
            sre += c;
            sim += s;
        }
        re += (double)sre;
        im += (double)sim;
    }
    A.rho[((int64_t)m * A.nq + qi) * 2 + 0] = re;
    A.rho[((int64_t)m * A.nq + qi) * 2 + 1] = im;
}

// F(q,tau) = sum_m Re rho(m) conj(rho(m+tau)).  A workgroup of 4 waves takes
// kIsfQ = 4 q-vectors, one per wave; rho_q(m) is staged twice over (m and
// m + M) so lag tau reads row[m + tau] with no wrap test: rho(m) is an LDS
// broadcast, rho(m + tau) consecutive words, each lag at a fixed offset.  Rows
// are 2M + 64 long: lanes past the last lag read the zero padding.
constexpr int kIsfQ = 4;
constexpr int kIsfLagsPerLane = 8;   // M <= 512
__global__ __launch_bounds__(256) void k_isf_from_rho(const double* __restrict__ rho, int nq, int M,
                                                     double inv_n, double* __restrict__ fqt) {
    extern __shared__ double2 sr[];  // [kIsfQ][2M + 64]
    const int R = 2 * M + 64;
    const int qb = blockIdx.x * kIsfQ;
    const int nql = min(kIsfQ, nq - qb);
    for (int t = threadIdx.x; t < M * kIsfQ; t += 256) {
        const int m = t / kIsfQ, k = t % kIsfQ;
        double2 v = make_double2(0.0, 0.0);
        if (k < nql) v = make_double2(rho[((int64_t)m * nq + qb + k) * 2], rho[((int64_t)m * nq + qb + k) * 2 + 1]);
        sr[(int64_t)k * R + m] = v;
        sr[(int64_t)k * R + m + M] = v;
    }
    for (int t = threadIdx.x; t < 64 * kIsfQ; t += 256) sr[(int64_t)(t / 64) * R + 2 * M + (t % 64)] = make_double2(0.0, 0.0);
    __syncthreads();
    const int k = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (k >= nql) return;
    const double2* row = sr + (int64_t)k * R;
    double acc[kIsfLagsPerLane];
#pragma unroll
    for (int u = 0; u < kIsfLagsPerLane; ++u) acc[u] = 0.0;
    const int nu = (M + 63) / 64;
    const double2* b = row + lane;
    for (int m = 0; m < M; ++m, ++b) {
        const double2 a = row[m];
#pragma unroll
        for (int u = 0; u < kIsfLagsPerLane; ++u) {
            if (u < nu) {
                const double2 c = b[64 * u];
                acc[u] = fma(a.x, c.x, acc[u]);
                acc[u] = fma(a.y, c.y, acc[u]);
            }
        }
    }
#pragma unroll
    for (int u = 0; u < kIsfLagsPerLane; ++u) {
        const int tau = lane + 64 * u;
        if (u < nu && tau < M) fqt[(int64_t)(qb + k) * M + tau] = acc[u] * inv_n;
    }
}

// S(q) contribution = sum_m |rho_q(m)|^2 / N, one lane per q (coalesced over q),
// slices added in order.  Eight slices' rho are read before their terms are
// added, so a lane has eight loads in flight instead of one: with a few hundred
// q the kernel is that chain of loads (257 q x 512 slices, a served S(q)
// request: 132 -> 28 us; 5377 q x 167: 48 -> 14 us; 100 401 q x 250: 134 -> 79
// us, ~5 TB/s; sixteen in flight gained nothing), and the order of the additions
// is unchanged.
__global__ void k_ssf_from_rho(const double* __restrict__ rho, int nq, int M, double inv_n,
                               double* __restrict__ sq) {
    const int qi = blockIdx.x * blockDim.x + threadIdx.x;
    if (qi >= nq) return;
    const double2* r = reinterpret_cast<const double2*>(rho) + qi;   // rho_q(m) at r[m nq]
    constexpr int U = 8;
    double acc = 0.0;
    int m = 0;
    for (; m + U <= M; m += U) {
        double2 v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = r[(int64_t)(m + u) * nq];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            acc = fma(v[u].x, v[u].x, acc);
            acc = fma(v[u].y, v[u].y, acc);
        }
    }
    for (; m < M; ++m) {
        const double2 v = r[(int64_t)m * nq];
        acc = fma(v.x, v.x, acc);
        acc = fma(v.y, v.y, acc);
    }
    sq[qi] = acc * inv_n;
}

// Direct minimum-image S(q) pair sum for general q: one block per (q, slice).
__global__ __launch_bounds__(256) void k_ssf_pairs(const double* __restrict__ pos, int64_t plane, int nalloc,
                                                  const int* __restrict__ nbs, const double* __restrict__ q,
                                                  BoxDev box, double* __restrict__ partial) {
    __shared__ double red[4];
    const int qi = blockIdx.x, m = blockIdx.y;
    const double qx = q[3 * qi], qy = q[3 * qi + 1], qz = q[3 * qi + 2];
    const double* px = pos + (int64_t)m * nalloc;
    const double* py = px + plane;
    const double* pz = py + plane;
    const int n = nbs[m];
    constexpr double kInv2Pi = 0.15915494309189533577;
    double acc = 0.0;
    const int64_t total = (int64_t)n * n;
    for (int64_t t = threadIdx.x; t < total; t += 256) {
        const int i = (int)(t / n), j = (int)(t % n);
        if (j <= i) continue;
        const double dx = min_image_exact(px[i] - px[j], box.L, box.invL);
        const double dy = min_image_exact(py[i] - py[j], box.L, box.invL);
        const double dz = min_image_exact(pz[i] - pz[j], box.L, box.invL);
        const double rev = (dx * qx + dy * qy + dz * qz) * kInv2Pi;
        const double red2 = rev - rint(rev);
        acc += 2.0 * (double)cospif((float)(2.0 * red2));
    }
    const double s = block_sum<256>(acc, red);
    if (threadIdx.x == 0) partial[(int64_t)qi * gridDim.y + m] = s + (double)n;
}

// sq[q] = sum_m partial[q][m] / N in slice order
__global__ void k_ssf_finish(const double* __restrict__ partial, int nq, int M, double inv_n,
                             double* __restrict__ sq) {
    const int qi = blockIdx.x * blockDim.x + threadIdx.x;
    if (qi >= nq) return;
    double s = 0.0;
    for (int m = 0; m < M; ++m) s += partial[(int64_t)qi * M + m];
    sq[qi] = s * inv_n;
}

hipError_t launch_rho(bool lattice, hipStream_t st, const RhoArgs& A) {
    (void)hipGetLastError();  // drop any stale error of an earlier call
    if (A.nq <= 0) return hipSuccess;
    if (lattice) {
        const size_t tables = sizeof(float2) * (2 * (A.nmax[0] + A.nmax[1] + A.nmax[2]) + 3) * kRhoBeadChunk;
        // the most q per lane that still leaves >= 2048 workgroups
        auto blocks = [&](int qpt) { return (int64_t)((A.nq + kRhoThreads * qpt - 1) / (kRhoThreads * qpt)) * A.num_slices; };
        int qpt = 8;
        while (qpt > 1 && blocks(qpt) < 2048) qpt >>= 1;
        dim3 g((A.nq + kRhoThreads * qpt - 1) / (kRhoThreads * qpt), A.num_slices), b(kRhoThreads);
        // row products when a workgroup takes 1024 or more q (up to 128 rows; a
        // q set in the reference's order has runs of n_z longer than 8, so this
        // is rarely exceeded, and a workgroup with more rows takes the per-lane
        // products); with fewer q per workgroup their set-up outweighs them.  The
        // products' space first holds the row scan (256 QPT ints).
        const int rows_cap = qpt >= 4 ? kRhoRowsMax : 0;
        const size_t lds = tables + sizeof(float2) * rows_cap * kRhoBeadChunk + sizeof(int2) * rows_cap;
        // set once per process (two threads racing here both set the same value)
        static std::atomic<bool> attr_set{false};
        if (!attr_set.load(std::memory_order_acquire)) {
            // up to 63 + 32 + 1 KB at |n_d| <= 40, above the default 64 KB of dynamic LDS
            // (the kernel's static LDS comes on top, within the 160 KB of a CU)
            const void* fns[4] = {(const void*)k_rho_lattice<1>, (const void*)k_rho_lattice<2>,
                                  (const void*)k_rho_lattice<4>, (const void*)k_rho_lattice<8>};
            for (const void* f : fns) {
                const hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, 128 * 1024);
                if (e != hipSuccess) return e;
            }
            attr_set.store(true, std::memory_order_release);
        }
        if (qpt == 8) hipLaunchKernelGGL(k_rho_lattice<8>, g, b, lds, st, A, rows_cap);
        else if (qpt == 4) hipLaunchKernelGGL(k_rho_lattice<4>, g, b, lds, st, A, rows_cap);
        else if (qpt == 2) hipLaunchKernelGGL(k_rho_lattice<2>, g, b, lds, st, A, rows_cap);
        else hipLaunchKernelGGL(k_rho_lattice<1>, g, b, lds, st, A, rows_cap);
    } else {
        dim3 g((A.nq + kRhoThreads - 1) / kRhoThreads, A.num_slices), b(kRhoThreads);
        hipLaunchKernelGGL(k_rho_general, g, b, 0, st, A);
    }
    return hipGetLastError();
}

hipError_t launch_isf(hipStream_t st, const double* rho, int nq, int M, double inv_n, double* fqt, int only_tau0) {
    (void)hipGetLastError();  // drop any stale error of an earlier call
    if (nq <= 0) return hipSuccess;
    if (only_tau0) {
        hipLaunchKernelGGL(k_ssf_from_rho, dim3((nq + 255) / 256), dim3(256), 0, st, rho, nq, M, inv_n, fqt);
    } else {
        dim3 g((nq + kIsfQ - 1) / kIsfQ), b(256);
        const size_t lds = sizeof(double2) * kIsfQ * (2 * (size_t)M + 64);
        static std::atomic<bool> attr_set{false};
        if (!attr_set.load(std::memory_order_acquire)) {
            // rows of M <= 512 slices need up to 68 KB of the 160 KB LDS
            hipError_t e = hipFuncSetAttribute((const void*)k_isf_from_rho,
                                               hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
            if (e != hipSuccess) return e;
            attr_set.store(true, std::memory_order_release);
        }
        hipLaunchKernelGGL(k_isf_from_rho, g, b, lds, st, rho, nq, M, inv_n, fqt);
    }
    return hipGetLastError();
}

hipError_t launch_ssf_pairs(hipStream_t st, const double* pos, int64_t plane, int nalloc, int M, const int* nbs,
                            const double* q, int nq, const BoxDev& box, double* partial, double inv_n, double* sq) {
    (void)hipGetLastError();  // drop any stale error of an earlier call
    if (nq <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_ssf_pairs, dim3(nq, M), dim3(256), 0, st, pos, plane, nalloc, nbs, q, box, partial);
    hipLaunchKernelGGL(k_ssf_finish, dim3((nq + 255) / 256), dim3(256), 0, st, partial, nq, M, inv_n, sq);
    return hipGetLastError();
}

int rho_max_nmax() { return kMaxNmax; }

}  // namespace pimc
