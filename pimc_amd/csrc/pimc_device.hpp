// pimc_device.hpp -- device-side pair-term arithmetic for gfx950.
//
// A bead pair's potential follows the reference's step table:
//   V = lookupV[k],  k = int(|r| / dr),  dr = 1e-6 rm,  V = 0 if k <= 0 or
//   k >= tableLength   (include/potential.h:249-260, src/potential.cpp:1792-1796)
// where lookupV[k] = valueV(r_k) at the k-fold accumulated radius r_k
// (include/potential.h:176-183).  The kernels therefore
//   1. form the minimum-image separation in FP64 (include/container.h:50-59),
//   2. compute the bin k exactly as the reference does (FP64 sqrt and division
//      in the exact modes; a refined FP32 root in the FP32 mode),
//   3. evaluate valueV at the bin radius x_k = k dr / rm, either in FP32
//      (PIMC_V_MIXED), FP64 (PIMC_V_FP64), or by gathering the FP64 table
//      built on the host exactly like the reference's (PIMC_V_TABLE).
// The three valueV branches (x < EPS, x < 0.01, x < D) are resolved by bin
// thresholds replayed on the host from the accumulated radius, so they switch at
// exactly the reference's bins.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pimc {

constexpr int kMaxAziz = 3;

// Aziz constants as the kernels consume them (filled on the host).
struct AzizDev {
    int table_length;   // tableLength = int(maxSep/dr)
    int k_eps;          // first bin with x_k >= EPS (below: V = 0)
    int k_hc;           // first bin with x_k >= 0.01 (below: hard core eps*A*exp(..))
    int k_d;            // first bin with x_k >= D (at and above: F = 1)
    int k_far;          // first bin from which on the repulsion is below 2^-53 of the dispersion
    double dr;          // 1e-6 rm
    double s;           // dr / rm, so x_k = k s
    // FP32 closed form, epsilon folded in, base-2 exponent of the repulsion:
    //   eps*A*exp(-alpha x + beta x^2) = 2^(c0 + c1 k + c2 k^2)
    double rep_c0, rep_c1, rep_c2;
    double eC6d, eC8d, eC10d;   // epsilon * C6, C8, C10
    // the dispersion as a polynomial in 1/k^2: epsilon C_n / s^n, and D / s
    double eC6k, eC8k, eC10k, Ds;
    // FP64 closed form
    double eps, A, alpha, beta, D, C6, C8, C10;
    double eps_rm;         // epsilon / rm (dV/dr = (epsilon/rm) dV/dx)
    double inv_rm;         // 1 / rm
    const double* table;   // device lookupV (PIMC_V_TABLE)
    const double* dtable;  // device lookupdVdr (PIMC_V_TABLE, gradient terms)
    const double* d2table; // device lookupd2Vdr2 (PIMC_V_TABLE, virial T-matrix)
};

struct BoxDev {
    double L, invL;     // cube side and its inverse (sideInv)
    double rc2;         // cutoff^2 (Vnn / interaction list)
    int ngrid;          // numNNGrid per dimension (cube)
    double grid_size;   // sizeNNGrid
};

// ---------------------------------------------------------------------------
// Minimum image, exactly as putInBC: r -= L * floor(r * invL + 0.5).
// ---------------------------------------------------------------------------
__device__ __forceinline__ double min_image_exact(double d, double L, double invL) {
#pragma clang fp contract(off)
    return d - L * floor(d * invL + 0.5);
}

__device__ __forceinline__ double min_image_fast(double d, double L, double invL) {
    return fma(-L, floor(fma(d, invL, 0.5)), d);
}

// |r|^2 in the reference's order ((0 + x*x) + y*y) + z*z without contraction.
__device__ __forceinline__ double norm2_exact(double x, double y, double z) {
#pragma clang fp contract(off)
    double s = x * x;
    s = s + y * y;
    s = s + z * z;
    return s;
}

// Exact-mode bin: int(sqrt(r2)/dr) with IEEE sqrt and division, as the host does.
__device__ __forceinline__ int bin_exact(double r2, double dr, double* r_out) {
    double r = __builtin_sqrt(r2);
    *r_out = r;
    return (int)(r / dr);
}

__device__ __forceinline__ float fast_exp2(float x) { return __builtin_amdgcn_exp2f(x); }

// valueV at the bin radius with the chip's FP32 transcendental units (exp2, rcp)
// and FP64 arithmetic around them.  On gfx950 an unpacked FP32 op and an FP64
// FMA issue at the same rate, so only the transcendentals are worth taking in
// FP32; each exponent is split hi + lo so the FP32 exp2 sees an exactly
// representable argument and the lo part is applied in FP64.  Per-pair error
// is set by v_exp_f32 (about 1 FP32 ulp).  Requires k_eps <= k < table_length.
__device__ __forceinline__ double split_exp2(double e) {
    const float eh = (float)e;
    const double el = e - (double)eh;
    return (double)fast_exp2(eh) * fma(el, 0.69314718055994530942, 1.0);
}

// BETA0: the caller knows beta = 0 (rep_c2 = 0; the 1979 potential), so the
// exponent's quadratic term (an exact no-op then) is left out.
template <bool BETA0 = false>
__device__ __forceinline__ double aziz_vmix(int k, const AzizDev& a) {
    const double kd = (double)k;
    const double rep = split_exp2(BETA0 ? fma(a.rep_c1, kd, a.rep_c0) : fma(fma(a.rep_c2, kd, a.rep_c1), kd, a.rep_c0));
    // 1/x^2 = s^-2 / k^2: the dispersion is a polynomial in 1/k^2 with the powers
    // of s folded into its coefficients.  k^2 is exact in FP64; 1/k^2 is the
    // hardware FP64 reciprocal estimate (relative error ~2^-23) refined by one
    // Newton step (~2^-46).
    const double q = kd * kd;
    const double r0 = __builtin_amdgcn_rcp(q);
    const double iq = fma(r0, fma(-q, r0, 1.0), r0);
    const double disp = (iq * iq * iq) * fma(fma(a.eC10k, iq, a.eC8k), iq, a.eC6k);
    // F = 1 at and above D: rep - disp (== fma(-disp, 1, rep)); a wave whose
    // pairs all lie beyond D skips the damping block
    double v = rep - disp;
    if (k < a.k_d) {
        // F(x) = exp(-(D/x - 1)^2) for x < D, with D/x - 1 = (D/s) k / k^2 - 1
        const double t = fma(a.Ds * kd, iq, -1.0);
        const double F = split_exp2(-1.4426950408889634074 * t * t);
        // Below x = 0.01 (k < k_hc) valueV is the bare repulsion; there t >= D/0.01 - 1
        // > 120, so exp2(-t^2 log2 e) underflows to exactly 0 in FP32, F = 0 and
        // fma(-disp, 0, rep) == rep: the hard-core branch needs no select.
        v = fma(-disp, F, rep);
    }
    return v;
}

// valueV of a far bin, k_far <= k < table_length: the repulsion there is below
// 2^-53 of the dispersion, under the rounding of the FP64 difference, so V is
// minus the dispersion.  1/k^2 comes from y, the FP64 reciprocal-root estimate
// of |w| for a separation of |w| bins (k <= |w| < k + 1, y within ~2^-23): y^2
// is within 2/k + 2^-22 of 1/k^2, under 1e-6 for k >= k_far (x >= 3.1 for the
// 1979, 1987 and 1995 parameters), and one Newton step takes that to ~1e-12 with
// no FP64 reciprocal.
__device__ __forceinline__ double aziz_disp_y(double kd, double y, const AzizDev& a) {
    const double q = kd * kd;
    const double r0 = y * y;
    const double iq = fma(r0, fma(-q, r0, 1.0), r0);
    return (iq * iq * iq) * fma(fma(a.eC10k, iq, a.eC8k), iq, a.eC6k);
}
__device__ __forceinline__ double aziz_vfar(double kd, double y, const AzizDev& a) { return -aziz_disp_y(kd, y, a); }

// valueV of a bin at or beyond D, k_d <= k < table_length (F = 1): repulsion
// minus dispersion, 1/k^2 as in aziz_vfar (2/k + 2^-22 <= 2e-6 before the Newton
// step at x >= D >= 1.24, ~4e-12 after).
template <bool BETA0 = false>
__device__ __forceinline__ double aziz_vmid(double kd, double y, const AzizDev& a) {
    const double rep = split_exp2(BETA0 ? fma(a.rep_c1, kd, a.rep_c0) : fma(fma(a.rep_c2, kd, a.rep_c1), kd, a.rep_c0));
    return rep - aziz_disp_y(kd, y, a);
}

// valueV at the bin radius, FP64 arithmetic.  Requires k_eps <= k < table_length.
__device__ __forceinline__ double aziz_v64(int k, const AzizDev& a) {
    const double x = (double)k * a.s;
    const double rep = a.A * exp(-a.alpha * x + a.beta * x * x);
    if (k < a.k_hc) return a.eps * rep;
    const double ix2 = 1.0 / (x * x);
    const double ix6 = ix2 * ix2 * ix2;
    const double ix8 = ix6 * ix2;
    const double ix10 = ix8 * ix2;
    double F = 1.0;
    if (k < a.k_d) {
        const double u = a.D / x - 1.0;
        F = exp(-u * u);
    }
    return a.eps * (rep - (a.C6 * ix6 + a.C8 * ix8 + a.C10 * ix10) * F);
}

// valuedVdr at the bin radius (src/potential.cpp:1847-1873): FP64 closed form.
// Requires k_eps <= k < table_length.
__device__ __forceinline__ double aziz_dvdr64(int k, const AzizDev& a) {
    const double x = (double)k * a.s;
    const double T1 = a.A * (-a.alpha + 2.0 * a.beta * x) * exp(-a.alpha * x + a.beta * x * x);
    if (k < a.k_hc) return a.eps_rm * T1;
    const double ix = 1.0 / x;
    const double ix2 = ix * ix;
    const double ix6 = ix2 * ix2 * ix2;
    const double ix7 = ix6 * ix, ix8 = ix6 * ix2, ix9 = ix8 * ix, ix10 = ix8 * ix2, ix11 = ix10 * ix;
    double F = 1.0, dF = 0.0;
    if (k < a.k_d) {
        const double t = a.D * ix - 1.0;
        const double e = exp(-t * t);
        F = e;
        dF = 2.0 * a.D * ix2 * t * e;
    }
    const double T2 = (6.0 * a.C6 * ix7 + 8.0 * a.C8 * ix9 + 10.0 * a.C10 * ix11) * F;
    const double T3 = -(a.C6 * ix6 + a.C8 * ix8 + a.C10 * ix10) * dF;
    return a.eps_rm * (T1 + T2 + T3);
}

// valued2Vdr2 at the bin radius (src/potential.cpp:1880-1909): FP64 closed
// form, including the reference's epsilon/rm (not epsilon/rm^2) factor of the
// hard-core branch.  Requires k_eps <= k < table_length.
__device__ __forceinline__ double aziz_d2vdr2_64(int k, const AzizDev& a) {
    const double x = (double)k * a.s;
    const double ab = (a.alpha - 2.0 * a.beta * x) * (a.alpha - 2.0 * a.beta * x);
    const double T1 = a.A * (2.0 * a.beta + ab) * exp(-a.alpha * x + a.beta * x * x);
    if (k < a.k_hc) return a.eps_rm * T1;
    const double ix = 1.0 / x;
    const double ix2 = ix * ix;
    const double ix6 = ix2 * ix2 * ix2;
    const double ix7 = ix6 * ix, ix8 = ix6 * ix2, ix9 = ix8 * ix, ix10 = ix8 * ix2, ix11 = ix10 * ix;
    const double ix12 = ix11 * ix;
    double F = 1.0, dF = 0.0, d2F = 0.0;
    if (k < a.k_d) {
        const double t = a.D * ix - 1.0;
        const double e = exp(-t * t);
        F = e;
        dF = 2.0 * a.D * ix2 * t * e;
        const double D = a.D;
        d2F = 2.0 * D * ix2 * ix * (2.0 * D * D * D * ix2 * ix - 4.0 * D * D * ix2 - D * ix + 2.0) * e;
    }
    const double T2 = -(42.0 * a.C6 * ix8 + 72.0 * a.C8 * ix10 + 110.0 * a.C10 * ix12) * F;
    const double T3 = 2.0 * (6.0 * a.C6 * ix7 + 8.0 * a.C8 * ix9 + 10.0 * a.C10 * ix11) * dF;
    const double T4 = -(a.C6 * ix6 + a.C8 * ix8 + a.C10 * ix10) * d2F;
    return a.eps_rm * a.inv_rm * (T1 + T2 + T3 + T4);
}

// valuedVdr at the bin radius with the FP32 transcendental units (split
// exponents, FP64 around them), the counterpart of aziz_vmix.
__device__ __forceinline__ double aziz_dvdr_mix(int k, const AzizDev& a) {
    const double kd = (double)k;
    const double x = kd * a.s;
    // eps*A*exp(-alpha x + beta x^2) = 2^(c0 + c1 k + c2 k^2)
    const double rep = split_exp2(fma(fma(a.rep_c2, kd, a.rep_c1), kd, a.rep_c0));
    const double T1 = rep * (-a.alpha + 2.0 * a.beta * x);   // epsilon folded into rep
    double ix = (double)__builtin_amdgcn_rcpf((float)x);
    ix = ix * fma(-x, ix, 2.0);
    ix = ix * fma(-x, ix, 2.0);                  // two Newton steps: FP64 accurate
    const double ix2 = ix * ix;
    const double ix6 = ix2 * ix2 * ix2;
    double F = 1.0, dF = 0.0;
    if (k < a.k_d) {
        const double t = fma(a.D, ix, -1.0);
        const double e = split_exp2(-1.4426950408889634074 * t * t);
        F = e;
        dF = 2.0 * a.D * ix2 * t * e;
    }
    // epsilon * [ (6 C6 x^-7 + 8 C8 x^-9 + 10 C10 x^-11) F - (C6 x^-6 + C8 x^-8 + C10 x^-10) dF ]
    const double pd = ix6 * ix * fma(fma(10.0 * a.eC10d, ix2, 8.0 * a.eC8d), ix2, 6.0 * a.eC6d);
    const double pv = ix6 * fma(fma(a.eC10d, ix2, a.eC8d), ix2, a.eC6d);
    const double att = fma(pd, F, -pv * dF);
    // below x = 0.01 the damping factor and its derivative underflow to 0 (see aziz_vmix)
    return (T1 + att) * a.inv_rm;
}

// valued2Vdr2 at the bin radius with the FP32 transcendental units, the
// counterpart of aziz_d2vdr2_64 (the same hard-core epsilon/rm factor).
__device__ __forceinline__ double aziz_d2vdr2_mix(int k, const AzizDev& a) {
    const double kd = (double)k;
    const double x = kd * a.s;
    const double rep = split_exp2(fma(fma(a.rep_c2, kd, a.rep_c1), kd, a.rep_c0));   // epsilon folded in
    const double ab = (a.alpha - 2.0 * a.beta * x) * (a.alpha - 2.0 * a.beta * x);
    const double T1 = rep * (2.0 * a.beta + ab);
    if (k < a.k_hc) return T1 * a.inv_rm;
    double ix = (double)__builtin_amdgcn_rcpf((float)x);
    ix = ix * fma(-x, ix, 2.0);
    ix = ix * fma(-x, ix, 2.0);                  // two Newton steps: FP64 accurate
    const double ix2 = ix * ix;
    const double ix6 = ix2 * ix2 * ix2;
    double F = 1.0, dF = 0.0, d2F = 0.0;
    if (k < a.k_d) {
        const double D = a.D;
        const double t = fma(D, ix, -1.0);
        const double e = split_exp2(-1.4426950408889634074 * t * t);
        F = e;
        dF = 2.0 * D * ix2 * t * e;
        d2F = 2.0 * D * ix2 * ix * (2.0 * D * D * D * ix2 * ix - 4.0 * D * D * ix2 - D * ix + 2.0) * e;
    }
    // epsilon * [ -(42 C6 x^-8 + 72 C8 x^-10 + 110 C10 x^-12) F
    //             + 2 (6 C6 x^-7 + 8 C8 x^-9 + 10 C10 x^-11) dF - (C6 x^-6 + C8 x^-8 + C10 x^-10) d2F ]
    const double p2 = ix6 * ix2 * fma(fma(110.0 * a.eC10d, ix2, 72.0 * a.eC8d), ix2, 42.0 * a.eC6d);
    const double p1 = ix6 * ix * fma(fma(10.0 * a.eC10d, ix2, 8.0 * a.eC8d), ix2, 6.0 * a.eC6d);
    const double p0 = ix6 * fma(fma(a.eC10d, ix2, a.eC8d), ix2, a.eC6d);
    const double att = -p2 * F + 2.0 * p1 * dF - p0 * d2F;
    return (T1 + att) * (a.inv_rm * a.inv_rm);
}

// The gradient factor g of gradV(sep) = g * sep, g = lookupdVdr[k] / |sep|
// (include/potential.h:996-1002 with direct(), :249-260); 0 outside the table.
// Mixed-mode bin with both |r| and 1/|r|: the hardware FP64 reciprocal-root
// estimate (relative error ~2^-23) refined by one Newton step (~2^-46), then
// |r| = r2 / |r|; the bin differs from the exact one only when |r|/dr lies
// within ~1e-13 relative of an integer.
__device__ __forceinline__ int bin_rsq(double r2, double inv_dr, double* r_out, double* ir_out) {
    const double y0 = __builtin_amdgcn_rsq(r2);
    const double h = fma(-(r2 * y0), y0, 1.0);
    const double y = fma(0.5 * y0, h, y0);
    const double r = r2 * y;
    *r_out = r;
    *ir_out = y;
    return (int)(r * inv_dr);
}

template <int MODE>
__device__ __forceinline__ double grad_factor_from_r2(double r2, const AzizDev& a, double inv_dr) {
    int k;
    double rn, ir = 0.0;
    if (MODE == 0) {
        k = bin_rsq(r2, inv_dr, &rn, &ir);
    } else {
        k = bin_exact(r2, a.dr, &rn);
    }
    const bool valid = (unsigned)(k - a.k_eps) < (unsigned)(a.table_length - a.k_eps);
    const int kk = valid ? k : a.k_hc;   // keep the arithmetic finite on masked lanes
    double d;
    if (MODE == 0)
        d = aziz_dvdr_mix(kk, a);
    else if (MODE == 1)
        d = aziz_dvdr64(kk, a);
    else
        d = valid ? a.dtable[k] : 0.0;
    if (MODE == 0) return valid ? d * ir : 0.0;
    return valid ? d / rn : 0.0;
}

// Pair value for separation components (dx,dy,dz) already minimum-imaged.
// MODE: 0 mixed (FP32 transcendentals), 1 FP64, 2 table.  Returns 0 outside the table.
template <int MODE>
__device__ __forceinline__ double pair_value_from_r2(double r2, const AzizDev& a, double inv_dr) {
    int k;
    if (MODE == 0) {
        double r, ir;
        k = bin_rsq(r2, inv_dr, &r, &ir);
    } else {
        double r;
        k = bin_exact(r2, a.dr, &r);
    }
    // k_eps <= k < tableLength (k_eps >= 1 always: bin 0 has x = 0 < EPS), one compare
    const bool valid = (unsigned)(k - a.k_eps) < (unsigned)(a.table_length - a.k_eps);
    double v = 0.0;
    if (MODE == 0) {
        const int kk = valid ? k : a.k_hc;  // keep the arithmetic finite on masked lanes
        v = aziz_vmix(kk, a);
    } else if (MODE == 1) {
        const int kk = valid ? k : a.k_hc;
        v = aziz_v64(kk, a);
    } else {
        v = valid ? a.table[k] : 0.0;
    }
    return valid ? v : 0.0;
}

template <int MODE>
__device__ __forceinline__ double sep_r2(double xi, double yi, double zi, double xj, double yj, double zj,
                                         const BoxDev& b) {
    if (MODE == 0) {
        const double dx = min_image_fast(xj - xi, b.L, b.invL);
        const double dy = min_image_fast(yj - yi, b.L, b.invL);
        const double dz = min_image_fast(zj - zi, b.L, b.invL);
        return fma(dz, dz, fma(dy, dy, dx * dx));
    } else {
        const double dx = min_image_exact(xj - xi, b.L, b.invL);
        const double dy = min_image_exact(yj - yi, b.L, b.invL);
        const double dz = min_image_exact(zj - zi, b.L, b.invL);
        return norm2_exact(dx, dy, dz);
    }
}

// LookupTable::gridIndex for one coordinate (include/lookuptable.h:139-147).
__device__ __host__ __forceinline__ int grid_coord(double p, double L, double size) {
#pragma clang fp contract(off)
    return (int)(fabs(p + 0.5 * L - 1.0e-7) / (size + 1.0e-7));
}

// Neighbouring cells of the 27-cell stencil with periodic wrap (the set
// LookupTable::updateInteractionList walks, src/lookuptable.cpp:480-486).
__device__ __forceinline__ bool cells_adjacent(int c1, int c2, int ng) {
    const int x1 = c1 / (ng * ng), y1 = (c1 / ng) % ng, z1 = c1 % ng;
    const int x2 = c2 / (ng * ng), y2 = (c2 / ng) % ng, z2 = c2 % ng;
    int dx = x1 - x2, dy = y1 - y2, dz = z1 - z2;
    dx = dx < 0 ? -dx : dx;
    dy = dy < 0 ? -dy : dy;
    dz = dz < 0 ? -dz : dz;
    return (dx <= 1 || dx == ng - 1) && (dy <= 1 || dy == ng - 1) && (dz <= 1 || dz == ng - 1);
}

__device__ __forceinline__ int cell_of(double x, double y, double z, const BoxDev& b) {
    // positions inside the box map into [0, ngrid); the clamp only guards memory
    const int gx = min(max(grid_coord(x, b.L, b.grid_size), 0), b.ngrid - 1);
    const int gy = min(max(grid_coord(y, b.L, b.grid_size), 0), b.ngrid - 1);
    const int gz = min(max(grid_coord(z, b.L, b.grid_size), 0), b.ngrid - 1);
    return (gx * b.ngrid + gy) * b.ngrid + gz;
}

// Worm::getState (include/worm.h:110-117): 0 HEADTAIL, 1 SPECIAL, 2 NONE.
struct WormDev {
    int head[2], tail[2], sp1[2], sp2[2];
};

__device__ __forceinline__ int bead_state(const WormDev& w, int slice, int ptcl) {
    if ((slice == w.head[0] && ptcl == w.head[1]) || (slice == w.tail[0] && ptcl == w.tail[1])) return 0;
    if ((slice == w.sp1[0] && ptcl == w.sp1[1]) || (slice == w.sp2[0] && ptcl == w.sp2[1])) return 1;
    return 2;
}

// Worm::factor(state1, bead2) (src/worm.cpp:108-118)
__device__ __forceinline__ double pair_factor(int s1, int s2) {
    if (s1 == 2 && s2 == 2) return 1.0;
    if (s1 == 0 && s2 == 0) return 0.0;
    return 0.5;
}

__device__ __forceinline__ bool worm_touches_slice(const WormDev& w, int slice) {
    return w.head[0] == slice || w.tail[0] == slice || w.sp1[0] == slice || w.sp2[0] == slice;
}

// Path::updateBead -> LookupTable::updateBead (src/lookuptable.cpp:355-393, :430-458):
// bead (m, i) moves to cell c_new; the label table keeps the reference's
// swap-with-last order.
__device__ __forceinline__ void cell_update_bead(int m, int i, int c_new, int nalloc, int num_slices, int* cell, int* label,
                                 int* numlab, int* hash) {
    const int bi = m * nalloc + i;
    const int c_old = cell[bi];
    if (c_new == c_old || c_old < 0) return;  // unchanged cell, or bead not in the table
    const int64_t hb_old = ((int64_t)c_old * num_slices + m) * nalloc;
    const int lab = label[bi];
    const int last = numlab[c_old * num_slices + m] - 1;
    if (lab < last) {
        const int moved = hash[hb_old + last];
        hash[hb_old + last] = -1;
        hash[hb_old + lab] = moved;
        label[m * nalloc + moved] = lab;
    } else {
        hash[hb_old + last] = -1;
    }
    numlab[c_old * num_slices + m] = last;
    const int64_t hb_new = ((int64_t)c_new * num_slices + m) * nalloc;
    const int nl = numlab[c_new * num_slices + m];
    hash[hb_new + nl] = i;
    label[bi] = nl;
    numlab[c_new * num_slices + m] = nl + 1;
    cell[bi] = c_new;
}

// ---------------------------------------------------------------------------
// Deterministic wave64 / workgroup FP64 reductions (fixed shuffle tree).
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

template <int NT>
__device__ __forceinline__ double block_sum(double v, double* lds) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) lds[wid] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int w = 0; w < NT / 64; ++w) r += lds[w];
    }
    return r;  // valid in thread 0
}

}  // namespace pimc
