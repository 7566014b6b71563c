// pimc_virial.hip -- the virial and centroid-virial pair terms of a closed path
// (LocalAction::rDOTgradUterm1/2, deltadotgradUterm1/2, gradVSquared,
// src/action.cpp:1188-1223, :1446-1762) and the link sums of
// VirialEnergyEstimator::accumulate (src/estimator.cpp:1125-1146).
//
// One lane per bead i of slice m (workgroup (b, m) takes beads 256 b ..
// 256 b + 255, the slice staged in LDS).  The lane walks every partner j of the
// slice (no cutoff, as the reference) and accumulates
//   F_i = sum_j gradV(r_i - r_j),   gradV(r) = (lookupdVdr[k] / |r|) r
//   T_i = sum_j [ r r^T d2V / |r|^2 - r r^T dV / |r|^3 + I dV / |r| ]
// with d2V = lookupd2Vdr2[k] and dV = |gradV(r)| (a magnitude, as the
// reference takes it), then writes seven per-bead terms
//   F.r_i, (T F).r_i, F.delta_i, (T F).delta_i, |F|^2, -vel1.vel2, -vel2.vel2
// where delta_i is the bead's offset from the centroid of its worldline window
// (virialWindow links each way) and vel1 / vel2 the window and one-link
// displacements along the next links.  Each workgroup reduces its beads in a
// fixed order, so the per-slice sums the host forms are deterministic.
#include <hip/hip_runtime.h>

#include "pimc_kernels.hpp"

namespace pimc {

namespace {

// Position of bead (m, i) and the minimum-image separation of two beads
// (Path::getSeparation, include/path.h:179-184).
struct PathView {
    const double* pos;
    int64_t plane;
    int nalloc;
    __device__ __forceinline__ void at(int m, int i, double r[3]) const {
        const double* p = pos + (int64_t)m * nalloc + i;
        r[0] = p[0];
        r[1] = p[plane];
        r[2] = p[2 * plane];
    }
};

__device__ __forceinline__ void sep_of(const double a[3], const double b[3], const BoxDev& box, double s[3]) {
    for (int q = 0; q < 3; ++q) s[q] = min_image_exact(a[q] - b[q], box.L, box.invL);
}

__device__ __forceinline__ double dot3(const double a[3], const double b[3]) {
#pragma clang fp contract(off)
    double s = a[0] * b[0];
    s = s + a[1] * b[1];
    s = s + a[2] * b[2];
    return s;
}

__device__ __forceinline__ void step_link(const int* links, int M, int nalloc, int& m, int& i, int dir) {
    if (links) {
        const int64_t t = (int64_t)m * nalloc + i;
        const int nm = links[2 * t], ni = links[2 * t + 1];
        m = nm;
        i = ni;
    } else {
        m = dir > 0 ? (m + 1 == M ? 0 : m + 1) : (m == 0 ? M - 1 : m - 1);
    }
}

// The sum over partners of one bead: F (3) and, when need_t, the symmetric T
// (xx, xy, xz, yy, yz, zz).
template <int MODE, typename P>
__device__ __forceinline__ void bead_pairs(P px, P py, P pz, int j0, int j1, int i, const BoxDev& box,
                                           const AzizDev& a, bool need_t, double F[3], double T[6]) {
#pragma clang fp contract(off)
    const double xi = px[i], yi = py[i], zi = pz[i];
    const double inv_dr = 1.0 / a.dr;
    for (int j = j0; j < j1; ++j) {
        if (j == i) continue;
        double s0, s1, s2, rn, ir;
        int k;
        if (MODE == 0) {
            s0 = min_image_fast(xi - px[j], box.L, box.invL);
            s1 = min_image_fast(yi - py[j], box.L, box.invL);
            s2 = min_image_fast(zi - pz[j], box.L, box.invL);
            k = bin_rsq(norm2_exact(s0, s1, s2), inv_dr, &rn, &ir);
        } else {
            s0 = min_image_exact(xi - px[j], box.L, box.invL);
            s1 = min_image_exact(yi - py[j], box.L, box.invL);
            s2 = min_image_exact(zi - pz[j], box.L, box.invL);
            k = bin_exact(norm2_exact(s0, s1, s2), a.dr, &rn);
            // one division per pair: the reference's quotients by |r|, |r|^2 and
            // |r|^3 become products with 1/|r| (a rounding apart)
            ir = 1.0 / rn;
        }
        // outside (0, tableLength) gradV and grad2V are 0: the pair adds nothing
        if ((unsigned)(k - a.k_eps) >= (unsigned)(a.table_length - a.k_eps)) continue;
        const double d = (MODE == 2) ? a.dtable[k] : (MODE == 1 ? aziz_dvdr64(k, a) : aziz_dvdr_mix(k, a));
        // |gradV| = |dV/dr|
        const double g = d * ir;
        const double g0 = g * s0, g1 = g * s1, g2 = g * s2;
        F[0] += g0;
        F[1] += g1;
        F[2] += g2;
        if (need_t) {
            const double d2 = (MODE == 2) ? a.d2table[k] : (MODE == 1 ? aziz_d2vdr2_64(k, a) : aziz_d2vdr2_mix(k, a));
            const double dV = __builtin_fabs(d);
            const double ir2 = ir * ir;
            const double diag = dV * ir;
            // r r^T (d2V/r^2 - |gradV|/r^3) + I |gradV|/r
            const double c = d2 * ir2 - dV * ir2 * ir;
            const double s[3] = {s0, s1, s2};
            int q = 0;
#pragma unroll
            for (int u = 0; u < 3; ++u)
#pragma unroll
                for (int v = u; v < 3; ++v, ++q) {
                    T[q] += s[u] * s[v] * c;
                    if (u == v) T[q] += diag;
                }
        }
    }
}

// The per-bead terms from the bead's F and T: (T F).r, the window centroid
// offset delta and the link displacements, into o[kVirialTerms].
__device__ void bead_finish(const VirialArgs& A, int m, int i, const double F[3], const double T[6], double* o) {
#pragma clang fp contract(off)
    // T F (apply_matrix_vector_product, include/common.h:187-199)
    double TF[3];
    TF[0] = ((0.0 + T[0] * F[0]) + T[1] * F[1]) + T[2] * F[2];
    TF[1] = ((0.0 + T[1] * F[0]) + T[3] * F[1]) + T[4] * F[2];
    TF[2] = ((0.0 + T[2] * F[0]) + T[4] * F[1]) + T[5] * F[2];
    const PathView P{A.pos, A.plane, A.nalloc};
    double r1[3];
    P.at(m, i, r1);
    // window centroid (src/action.cpp:1621-1645): running separations along
    // next / prev links from the bead, the bead itself counted for gamma = 0
    double more[3] = {0.0, 0.0, 0.0}, less[3] = {0.0, 0.0, 0.0}, com[3] = {0.0, 0.0, 0.0};
    int nm = m, ni = i, pm = m, pi = i;
    double nold[3] = {r1[0], r1[1], r1[2]}, pold[3] = {r1[0], r1[1], r1[2]};
    for (int gamma = 0; gamma < A.window; ++gamma) {
        if (gamma > 0) {
            step_link(A.next, A.num_slices, A.nalloc, nm, ni, 1);
            step_link(A.prev, A.num_slices, A.nalloc, pm, pi, -1);
        }
        double rn[3], rp[3], sn[3], sp[3];
        P.at(nm, ni, rn);
        P.at(pm, pi, rp);
        sep_of(rn, nold, A.box, sn);
        sep_of(rp, pold, A.box, sp);
        for (int q = 0; q < 3; ++q) {
            more[q] += sn[q];
            less[q] += sp[q];
            com[q] += (r1[q] + more[q]) + (r1[q] + less[q]);
            nold[q] = rn[q];
            pold[q] = rp[q];
        }
    }
    double delta[3];
    for (int q = 0; q < 3; ++q) {
        com[q] /= (2.0 * A.window);
        delta[q] = min_image_exact(r1[q] - com[q], A.box.L, A.box.invL);
    }
    // link displacements (src/estimator.cpp:1125-1146)
    int km = m, ki = i;
    double vel1[3] = {0.0, 0.0, 0.0}, vel2[3] = {0.0, 0.0, 0.0}, kold[3] = {r1[0], r1[1], r1[2]};
    for (int gamma = 1; gamma <= A.window; ++gamma) {
        step_link(A.next, A.num_slices, A.nalloc, km, ki, 1);
        double rk[3], sk[3];
        P.at(km, ki, rk);
        sep_of(rk, kold, A.box, sk);
        for (int q = 0; q < 3; ++q) {
            vel1[q] += sk[q];
            kold[q] = rk[q];
        }
        if (gamma == 1)
            for (int q = 0; q < 3; ++q) vel2[q] = sk[q];
    }
    o[0] = dot3(F, r1);
    o[1] = dot3(TF, r1);
    o[2] = dot3(F, delta);
    o[3] = dot3(TF, delta);
    o[4] = dot3(F, F);
    o[5] = -dot3(vel1, vel2);
    o[6] = -dot3(vel2, vel2);
}

template <int MODE>
__global__ __launch_bounds__(256) void k_virial(VirialArgs A, AzizDev az) {
    extern __shared__ double lds[];
    __shared__ double red[4][kVirialTerms];
    const int m = blockIdx.y, b = blockIdx.x;
    const int n = A.nbs[m];
    const double* px = A.pos + (int64_t)m * A.nalloc;
    const double* py = px + A.plane;
    const double* pz = py + A.plane;
    const bool in_lds = n <= A.lds_beads;
    double* lx = lds;
    double* ly = lds + A.lds_beads;
    double* lz = lds + 2 * A.lds_beads;
    if (in_lds)
        for (int t = threadIdx.x; t < n; t += 256) {
            lx[t] = px[t];
            ly[t] = py[t];
            lz[t] = pz[t];
        }
    __syncthreads();
    const int i = b * 256 + threadIdx.x;
    double o[kVirialTerms];
#pragma unroll
    for (int q = 0; q < kVirialTerms; ++q) o[q] = 0.0;
    if (i < n) {
        const bool need_t = (m & 1) ? A.t_odd : A.t_even;
        double F[3] = {0.0, 0.0, 0.0}, T[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        if (in_lds)
            bead_pairs<MODE>(lx, ly, lz, 0, n, i, A.box, az, need_t, F, T);
        else
            bead_pairs<MODE>(px, py, pz, 0, n, i, A.box, az, need_t, F, T);
        bead_finish(A, m, i, F, T, o);
    }
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int q = 0; q < kVirialTerms; ++q) {
        const double v = wave_sum(o[q]);
        if (lane == 0) red[wave][q] = v;
    }
    __syncthreads();
    if (threadIdx.x < kVirialTerms) {
        const int q = threadIdx.x;
        A.part[((int64_t)m * gridDim.x + b) * kVirialTerms + q] = ((red[0][q] + red[1][q]) + red[2][q]) + red[3][q];
    }
}

// Partner chunks when the slices alone leave CUs idle: block (b C + c, m)
// writes the partial F and T of rows 256 b ... against partners [c Nc, (c+1) Nc).
template <int MODE>
__global__ __launch_bounds__(256) void k_virial_chunk(VirialArgs A, AzizDev az) {
    extern __shared__ double lds[];
    const int C = A.chunks;
    const int m = blockIdx.y, b = blockIdx.x / C, c = blockIdx.x - b * C;
    const int n = A.nbs[m];
    const double* px = A.pos + (int64_t)m * A.nalloc;
    const double* py = px + A.plane;
    const double* pz = py + A.plane;
    const bool in_lds = n <= A.lds_beads;
    double* lx = lds;
    double* ly = lds + A.lds_beads;
    double* lz = lds + 2 * A.lds_beads;
    if (in_lds)
        for (int t = threadIdx.x; t < n; t += 256) {
            lx[t] = px[t];
            ly[t] = py[t];
            lz[t] = pz[t];
        }
    __syncthreads();
    const int i = b * 256 + threadIdx.x;
    if (i >= n) return;
    const int nc = (n + C - 1) / C;
    const int j0 = c * nc, j1 = min(n, j0 + nc);
    const bool need_t = (m & 1) ? A.t_odd : A.t_even;
    double F[3] = {0.0, 0.0, 0.0}, T[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (in_lds)
        bead_pairs<MODE>(lx, ly, lz, j0, j1, i, A.box, az, need_t, F, T);
    else
        bead_pairs<MODE>(px, py, pz, j0, j1, i, A.box, az, need_t, F, T);
    double* o = A.fpart + (((int64_t)m * C + c) * A.nalloc + i) * 9;
    for (int q = 0; q < 3; ++q) o[q] = F[q];
    for (int q = 0; q < 6; ++q) o[3 + q] = T[q];
}

// The chunks of each row summed in chunk order, then the per-bead terms.
__global__ __launch_bounds__(256) void k_virial_sum(VirialArgs A) {
    __shared__ double red[4][kVirialTerms];
    const int C = A.chunks;
    const int m = blockIdx.y, b = blockIdx.x;
    const int n = A.nbs[m];
    const int i = b * 256 + threadIdx.x;
    double o[kVirialTerms];
#pragma unroll
    for (int q = 0; q < kVirialTerms; ++q) o[q] = 0.0;
    if (i < n) {
        double F[3] = {0.0, 0.0, 0.0}, T[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        for (int c = 0; c < C; ++c) {
            const double* p = A.fpart + (((int64_t)m * C + c) * A.nalloc + i) * 9;
            for (int q = 0; q < 3; ++q) F[q] += p[q];
            for (int q = 0; q < 6; ++q) T[q] += p[3 + q];
        }
        bead_finish(A, m, i, F, T, o);
    }
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int q = 0; q < kVirialTerms; ++q) {
        const double v = wave_sum(o[q]);
        if (lane == 0) red[wave][q] = v;
    }
    __syncthreads();
    if (threadIdx.x < kVirialTerms) {
        const int q = threadIdx.x;
        A.part[((int64_t)m * gridDim.x + b) * kVirialTerms + q] = ((red[0][q] + red[1][q]) + red[2][q]) + red[3][q];
    }
}

}  // namespace

hipError_t launch_virial(int mode, hipStream_t st, const VirialArgs& A, const AzizDev& az, int* blocks_per_slice) {
    (void)hipGetLastError();
    const int B = (A.nalloc + 255) / 256;
    *blocks_per_slice = B;
    dim3 g(B, A.num_slices), b(256);
    const size_t lds = sizeof(double) * 3 * (size_t)A.lds_beads;
    if (A.chunks > 1 && A.fpart) {
        dim3 gc(B * A.chunks, A.num_slices);
        if (mode == 2)
            hipLaunchKernelGGL(k_virial_chunk<2>, gc, b, lds, st, A, az);
        else if (mode == 1)
            hipLaunchKernelGGL(k_virial_chunk<1>, gc, b, lds, st, A, az);
        else
            hipLaunchKernelGGL(k_virial_chunk<0>, gc, b, lds, st, A, az);
        hipLaunchKernelGGL(k_virial_sum, g, b, 0, st, A);
        return hipGetLastError();
    }
    if (mode == 2)
        hipLaunchKernelGGL(k_virial<2>, g, b, lds, st, A, az);
    else if (mode == 1)
        hipLaunchKernelGGL(k_virial<1>, g, b, lds, st, A, az);
    else
        hipLaunchKernelGGL(k_virial<0>, g, b, lds, st, A, az);
    return hipGetLastError();
}

}  // namespace pimc
