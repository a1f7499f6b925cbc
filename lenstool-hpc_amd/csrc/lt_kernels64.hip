// lt_kernels64.hip -- FP64 lens-map kernels for gfx950: gradient, Hessian (+ fused map epilogue),
// potential. One thread owns PX pixels of one row (columns lane, lane+64, ...); a 64x4 workgroup
// covers 64*PX columns of 4 rows. Halo records are walked in wavefront-uniform order, run by run.
//
// Formulas (restated from the mathematics; reference equivalents cited):
//   SIS gradient  src/gradient.cpp:260-299      PIEMD gradient  src/gradient.hpp:64-76,
//   dPIE gradient src/gradient.cpp:744-809      PIEMD Hessian   src/gradient2.cpp:54-82, 92-205
//   potential     src/potential_GPU.cu:47-116   epilogues       src/grid_map_*_GPU.cu
#include <hip/hip_runtime.h>

#include "lenstool_hip/lt_status.h"
#include "lt_device.h"
#include "lt_kernels.h"
#include "lt_math.h"

namespace lt {

// ------------------------------------------------------------------------------------------------
// per-halo Hessian contributions, accumulated in the lab frame as
//   T = a + c (trace, rotation invariant), D = a - c, B = b.
// Halo frame -> lab frame: D' = cos2t D - 2 sin2t B, B' = sin2t D / 2 + cos2t B; the halo amplitude is
// folded into hc2 = amph cos2t and hs2 = amph sin2t.
// ------------------------------------------------------------------------------------------------

struct PixelFrame {
    double x, y, x2, y2;
};

__device__ __forceinline__ PixelFrame halo_frame(const HaloRec& h, double px, double py)
{
    PixelFrame f;
    double tx = px - h.xh;
    double ty = py - h.yh;
    f.x = __builtin_fma(tx, h.cs, ty * h.sn);
    f.y = __builtin_fma(ty, h.cs, -(tx * h.sn));
    f.x2 = f.x * f.x;
    f.y2 = f.y * f.y;
    return f;
}

// One I_{0.5} term of the PIEMD Hessian in the halo frame (Kassiola & Kovner 1993), split into
// trace / difference / off-diagonal parts. The den2 pieces cancel in the trace.
struct PiemdH {
    double T, D, B;
};

__device__ __forceinline__ PiemdH piemd_hess_term(const HaloRec& h, const PixelFrame& f, double rem2,
                                                  double yi, double cxx, double X1, double Y6,
                                                  double Y5, double XY3, double r2, double e2r)
{
    double q = r2 + rem2;
    double iw = rsq64(q);
    double w = q * iw;
    double t = __builtin_fma(h.e2, w, -yi);
    double den1 = __builtin_fma(t, t, cxx);
    double n2 = e2r - f.y;
    double den2 = __builtin_fma(n2, n2, f.x2);
    double ip = rcp64(den1 * den2);
    double id1 = den2 * ip;
    double id2 = den1 * ip;
    double bxx = __builtin_fma(X1, iw, __builtin_fma(-h.hk2, w, f.y));
    double byy = __builtin_fma(h.hk4, w, __builtin_fma(Y6, iw, -Y5));
    double bxy = __builtin_fma(XY3, iw, -f.x);
    PiemdH o;
    o.T = (bxx + byy) * id1;
    o.D = __builtin_fma(bxx - byy, id1, (n2 + n2) * id2);
    o.B = __builtin_fma(bxy, id1, f.x * id2);
    return o;
}

template <bool CUT>
__device__ __forceinline__ void hess_piemd(const HaloRec& h, double px, double py, double& aT, double& aD,
                                           double& aB)
{
    PixelFrame f = halo_frame(h, px, py);
    double rem2 = __builtin_fma(f.x2, h.kx, f.y2 * h.ky);
    double yi = f.y * h.icx1;
    double cxx = h.cx1sq * f.x2;
    double X1 = h.hk1 * f.x2;
    double Y6 = h.hk6 * f.y2;
    double Y5 = h.hk5 * f.y;
    double XY3 = h.hk3 * f.x * f.y;
    PiemdH a = piemd_hess_term(h, f, rem2, yi, cxx, X1, Y6, Y5, XY3, h.ra2, h.e2ra);
    double dT = a.T, dD = a.D, dB = a.B;
    if (CUT) {
        PiemdH b = piemd_hess_term(h, f, rem2, yi, cxx, X1, Y6, Y5, XY3, h.rb2, h.e2rb);
        dT -= b.T;
        dD -= b.D;
        dB -= b.B;
    }
    aT = __builtin_fma(h.amph, dT, aT);
    aD = __builtin_fma(h.hc2, dD, aD);
    aD = __builtin_fma(-2.0 * h.hs2, dB, aD);
    aB = __builtin_fma(0.5 * h.hs2, dD, aB);
    aB = __builtin_fma(h.hc2, dB, aB);
}

// SIS / SIE: phi = b0 R, R^2 = (1-e) x^2 + (1+e) y^2; halo-frame Hessian
// amph/R^3 (y^2, -xy, x^2) with amph = b0 (1-e)(1+e).
__device__ __forceinline__ void hess_sis(const HaloRec& h, double px, double py, double& aT, double& aD,
                                         double& aB)
{
    PixelFrame f = halo_frame(h, px, py);
    double R2 = __builtin_fma(h.sa, f.x2, h.sb * f.y2);
    double iR = rsq64(R2);
    double g = iR * iR * iR;
    double dyx = f.y2 - f.x2;
    double xy = f.x * f.y;
    aT = __builtin_fma(g * h.amph, f.x2 + f.y2, aT);
    aD = __builtin_fma(g, __builtin_fma(h.hc2, dyx, 2.0 * h.hs2 * xy), aD);
    aB = __builtin_fma(g, __builtin_fma(0.5 * h.hs2, dyx, -(h.hc2 * xy)), aB);
}

// ------------------------------------------------------------------------------------------------
// gradient terms, accumulated in the lab frame
// ------------------------------------------------------------------------------------------------

// N conj(D) of the I_{0.5} ratio Z = N/D (N = cx1 x + i t, D = x + i n2); its real part is
// non-negative everywhere, so arg Z lies in [-pi/2, pi/2].
struct NconjD {
    double re, im, den1, den2;
};

__device__ __forceinline__ NconjD piemd_ratio(const HaloRec& h, const PixelFrame& f, double rem2, double yi,
                                              double cxx, double r2, double e2r)
{
    double q = r2 + rem2;
    double w = q * rsq64(q);
    double t = __builtin_fma(h.e2, w, -yi);
    double n2 = e2r - f.y;
    NconjD o;
    o.re = __builtin_fma(t, n2, h.cx1 * f.x2);
    o.im = f.x * __builtin_fma(-h.cx1, n2, t);
    o.den1 = __builtin_fma(t, t, cxx);
    o.den2 = __builtin_fma(n2, n2, f.x2);
    return o;
}

// PIEMD (CUT = false) or dPIE (CUT = true): (ampg arg, -ampg ln|.|) of Z(rc) or Z(rc)/Z(rcut).
// For the dPIE the two logarithms and the two arguments are merged into one of each:
// ln|Za| - ln|Zb| = ln(den1a den2b / (den2a den1b)) / 2, arg Za - arg Zb = arg(Na conj(Da) Nb' Db)
// (both arguments in [-pi/2, pi/2], so the difference needs no unwrapping).
template <bool CUT>
__device__ __forceinline__ void grad_piemd(const HaloRec& h, double px, double py, double& gx, double& gy)
{
    PixelFrame f = halo_frame(h, px, py);
    double rem2 = __builtin_fma(f.x2, h.kx, f.y2 * h.ky);
    double yi = f.y * h.icx1;
    double cxx = h.cx1sq * f.x2;
    NconjD a = piemd_ratio(h, f, rem2, yi, cxx, h.ra2, h.e2ra);
    double pre, pim, num, den;
    if (CUT) {
        NconjD b = piemd_ratio(h, f, rem2, yi, cxx, h.rb2, h.e2rb);
        pre = __builtin_fma(a.re, b.re, a.im * b.im);
        pim = __builtin_fma(a.im, b.re, -(a.re * b.im));
        num = a.den1 * b.den2;
        den = a.den2 * b.den1;
    } else {
        pre = a.re;
        pim = a.im;
        num = a.den1;
        den = a.den2;
    }
    double arg = atan2(pim, pre);
    double lnr = 0.5 * log(num * rcp64(den));
    double re = h.ampg * arg;
    double im = -h.ampg * lnr;
    gx += __builtin_fma(re, h.cs, -(im * h.sn));
    gy += __builtin_fma(im, h.cs, re * h.sn);
}

__device__ __forceinline__ void grad_sis(const HaloRec& h, double px, double py, double& gx, double& gy)
{
    PixelFrame f = halo_frame(h, px, py);
    double R2 = __builtin_fma(h.sa, f.x2, h.sb * f.y2);
    double s = h.ampg * rsq64(R2);
    double rx = h.sa * f.x * s;
    double ry = h.sb * f.y * s;
    gx += __builtin_fma(rx, h.cs, -(ry * h.sn));
    gy += __builtin_fma(ry, h.cs, rx * h.sn);
}

// ------------------------------------------------------------------------------------------------
// potential (complex I_{0.5} potential of Kassiola & Kovner, src/potential_GPU.cu:47-72)
// ------------------------------------------------------------------------------------------------
struct cplx {
    double re, im;
};
__device__ __forceinline__ cplx cmul(cplx a, cplx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ cplx cdiv(cplx a, cplx b)
{
    double n = 1.0 / (b.re * b.re + b.im * b.im);
    return {(a.re * b.re + a.im * b.im) * n, (a.im * b.re - a.re * b.im) * n};
}
__device__ __forceinline__ cplx ccosh(cplx a)
{
    double ch = cosh(a.re), sh = sinh(a.re), s, c;
    sincos(a.im, &s, &c);
    return {ch * c, sh * s};
}
__device__ __forceinline__ cplx csinh(cplx a)
{
    double ch = cosh(a.re), sh = sinh(a.re), s, c;
    sincos(a.im, &s, &c);
    return {sh * c, ch * s};
}
__device__ __forceinline__ cplx clog(cplx a) { return {0.5 * log(a.re * a.re + a.im * a.im), atan2(a.im, a.re)}; }

__device__ double pi05(double x, double y, double eps, double rc, double b0)
{
    double sqe = sqrt(eps);
    double ci = .5 * (1. - eps * eps) / sqe;
    double rem2 = x * x / ((1. + eps) * (1. + eps)) + y * y / ((1. - eps) * (1. - eps));
    double e1 = 2. * sqe / (1 - eps);
    double e2 = 2. * sqe / (1 + eps);
    double z = sqrt(x * x + y * y);
    cplx eta = {-.5 * asinh(e1 * y / z), .5 * asin(e2 * x / z)};
    double srem = sqrt(rem2);
    cplx zeta = {0.5 * log((srem + sqrt(rc * rc + rem2)) / rc), 0.};
    cplx b1 = ccosh({eta.re + zeta.re, eta.im});
    cplx b2 = ccosh({eta.re - zeta.re, eta.im});
    cplx ce = ccosh(eta);
    cplx a1 = clog(cdiv(cmul(ce, ce), cmul(b1, b2)));
    cplx a2 = clog(cdiv(b1, b2));
    cplx c1 = cmul(csinh({2. * eta.re, 2. * eta.im}), a1);
    cplx c2 = cmul(csinh({2. * zeta.re, 0.}), a2);
    return b0 * ci * rc / srem * ((c1.im + c2.im) * x - (c1.re + c2.re) * y);
}

__device__ __forceinline__ double pot_halo(const HaloRec& h, int type, double px, double py)
{
    PixelFrame f = halo_frame(h, px, py);
    if (type == TYPE_SIS) return h.b0 * sqrt(__builtin_fma(h.sa, f.x2, h.sb * f.y2));
    double pa = pi05(f.x, f.y, h.eps, h.rc, h.b0);
    if (type == TYPE_PIEMD) return pa;
    double ps = pi05(f.x, f.y, h.eps, h.rcut, h.b0);
    return (h.rcut / (h.rcut - h.rc)) * (pa - ps);
}

// ------------------------------------------------------------------------------------------------
// epilogues (src/grid_map_ampli_GPU.cu:329-572, grid_map_shear_GPU.cu:208-342,
// grid_map_mass_GPU.cu:307-322); T, D, B are the Hessian sums a+c, a-c, b.
// The map kind is a template parameter, and within a kind every mode's value is formed and one of
// them selected, so each kernel has a single store per output plane with a fixed base pointer.
// ------------------------------------------------------------------------------------------------
__device__ __forceinline__ void eigen(double A, double B, double C, double& lam, double& mu)
{
    double e = sqrt((A - C) * (A - C) + 4. * B * B);
    lam = .5 * (A + C + e);
    mu = .5 * (A + C - e);
}

template <int KIND>
__device__ __forceinline__ void hess_epilogue(const OutDesc& o, int64_t idx, double T, double D, double B)
{
    const double a = 0.5 * (T + D), c = 0.5 * (T - D), b = B;
    const double r = o.r;
    if constexpr (KIND == LT_MAP_HESSIAN) {
        if (o.p0) o.p0[idx] = a;
        if (o.p1) o.p1[idx] = b;
        if (o.p2) o.p2[idx] = c;
        if (o.p3) o.p3[idx] = b;
    } else if constexpr (KIND == LT_MAP_FIELDS) {
        double kappa = 0.5 * T * r, g1 = 0.5 * D * r, g2 = B * r;
        double omk = 1. - kappa;
        double det = __builtin_fma(omk, omk, -__builtin_fma(g1, g1, g2 * g2));
        if (o.p0) o.p0[idx] = kappa;
        if (o.p1) o.p1[idx] = g1;
        if (o.p2) o.p2[idx] = g2;
        if (o.p3) o.p3[idx] = det;
        if (o.p4) o.p4[idx] = 1. / det;
    } else if constexpr (KIND == LT_MAP_AMPLI) {
        double lam, mu;
        eigen(1. - a * r, -b * r, 1. - c * r, lam, mu);
        const double lm = lam * mu;
        const double kappa = (a + c) * r / 2., ga1 = (a - c) * r / 2., ga2 = b * r;
        const double gp = sqrt(ga1 * ga1 + ga2 * ga2) / (1 - kappa);
        const double h1 = (a - c) / 2.;
        const int m = o.mode;
        double v = sqrt(h1 * h1 + b * b);                                       // 6: gamma, no dlsds
        v = m == 5 ? (a + c) / 2. : v;                                          // 5: kappa, no dlsds
        v = m == 4 ? (1 - kappa) * (1 + gp * gp) / (1 - gp * gp) : v;           // 4
        v = m == 3 ? -2.5 * log10(fabs(lm)) : v;                                // 3
        v = m == 2 ? 1. / fabs(lm) : v;                                         // 2
        v = m == 1 ? 1. / lm : v;                                               // 1
        o.p0[idx] = v;
    } else if constexpr (KIND == LT_MAP_SHEAR) {
        double lam, mu;
        eigen(1. - a * r, -b * r, 1. - c * r, lam, mu);
        const double q = lam / mu;
        const int m = o.mode;
        double v = b * r;                                                       // 4: gamma2
        v = m == 3 ? (a - c) * r / 2. : v;                                      // 3: gamma1
        v = m == 2 ? fabs((q * q - 1.) / (q * q + 1.)) : v;                     // 2
        v = m == 1 ? (lam - mu) / 2. : v;                                       // 1
        o.p0[idx] = v;
    } else {  // LT_MAP_MASS
        o.p0[idx] = (a + c) * 0.5 * o.mult;
    }
}

// ------------------------------------------------------------------------------------------------
// kernels
// ------------------------------------------------------------------------------------------------
constexpr int WAVE = 64;
constexpr int ROWS = 4;

// Sample s of a batch (blockIdx.z) uses its own halo records and output planes.
__device__ __forceinline__ const HaloRec* select_sample(const GridDesc& g, const HaloRec* halos, OutDesc& o)
{
    const int64_t s = blockIdx.z;
    const int64_t off = s * g.plane_stride;
    if (o.p0) o.p0 += off;
    if (o.p1) o.p1 += off;
    if (o.p2) o.p2 += off;
    if (o.p3) o.p3 += off;
    if (o.p4) o.p4 += off;
    return halos + s * g.halo_stride;
}

template <int PX, int KIND>
__global__ __launch_bounds__(WAVE* ROWS) void k_hessian64(const HaloRec* __restrict__ halos,
                                                          const Run* __restrict__ runs, int nruns,
                                                          GridDesc g, OutDesc o)
{
    const int row = blockIdx.y * ROWS + threadIdx.y;
    if (row >= g.ny) return;
    halos = select_sample(g, halos, o);
    const int col0 = blockIdx.x * (WAVE * PX) + threadIdx.x;
    const double py = pixel_coord(g.ymin, row + g.jstart, g.dy);
    double px[PX];
#pragma unroll
    for (int k = 0; k < PX; ++k) px[k] = pixel_coord(g.xmin, col0 + k * WAVE + g.istart, g.dx);
    double tT[PX], tD[PX], tB[PX];
#pragma unroll
    for (int k = 0; k < PX; ++k) tT[k] = tD[k] = tB[k] = 0.;
    for (int r = 0; r < nruns; ++r) {
        const Run run = runs[r];
        double sT[PX], sD[PX], sB[PX];
#pragma unroll
        for (int k = 0; k < PX; ++k) sT[k] = sD[k] = sB[k] = 0.;
        const int end = run.start + run.count;
        switch (run.type) {
        case TYPE_DPIE:
            for (int i = run.start; i < end; ++i) {
                const HaloRec& h = halos[i];
#pragma unroll
                for (int k = 0; k < PX; ++k) hess_piemd<true>(h, px[k], py, sT[k], sD[k], sB[k]);
            }
            break;
        case TYPE_PIEMD:
            for (int i = run.start; i < end; ++i) {
                const HaloRec& h = halos[i];
#pragma unroll
                for (int k = 0; k < PX; ++k) hess_piemd<false>(h, px[k], py, sT[k], sD[k], sB[k]);
            }
            break;
        case TYPE_SIS:
            for (int i = run.start; i < end; ++i) {
                const HaloRec& h = halos[i];
#pragma unroll
                for (int k = 0; k < PX; ++k) hess_sis(h, px[k], py, sT[k], sD[k], sB[k]);
            }
            break;
        default:  // TYPE_SHEAR: constant lab-frame Hessian
            for (int i = run.start; i < end; ++i) {
                const HaloRec& h = halos[i];
#pragma unroll
                for (int k = 0; k < PX; ++k) {
                    sT[k] += h.h14a + h.h14c;
                    sD[k] += h.h14a - h.h14c;
                    sB[k] += h.h14b;
                }
            }
            break;
        }
        const double w = run.weight;
#pragma unroll
        for (int k = 0; k < PX; ++k) {
            tT[k] = __builtin_fma(w, sT[k], tT[k]);
            tD[k] = __builtin_fma(w, sD[k], tD[k]);
            tB[k] = __builtin_fma(w, sB[k], tB[k]);
        }
    }
#pragma unroll
    for (int k = 0; k < PX; ++k) {
        const int col = col0 + k * WAVE;
        if (col < g.nx) hess_epilogue<KIND>(o, (int64_t)row * g.pitch + col, tT[k], tD[k], tB[k]);
    }
}

template <int PX>
__global__ __launch_bounds__(WAVE* ROWS) void k_gradient64(const HaloRec* __restrict__ halos,
                                                           const Run* __restrict__ runs, int nruns,
                                                           GridDesc g, OutDesc o)
{
    const int row = blockIdx.y * ROWS + threadIdx.y;
    if (row >= g.ny) return;
    halos = select_sample(g, halos, o);
    const int col0 = blockIdx.x * (WAVE * PX) + threadIdx.x;
    const double py = pixel_coord(g.ymin, row + g.jstart, g.dy);
    double px[PX], tx[PX], ty[PX];
#pragma unroll
    for (int k = 0; k < PX; ++k) {
        px[k] = pixel_coord(g.xmin, col0 + k * WAVE + g.istart, g.dx);
        tx[k] = ty[k] = 0.;
    }
    for (int r = 0; r < nruns; ++r) {
        const Run run = runs[r];
        double sx[PX], sy[PX];
#pragma unroll
        for (int k = 0; k < PX; ++k) sx[k] = sy[k] = 0.;
        const int end = run.start + run.count;
        switch (run.type) {
        case TYPE_DPIE:
            for (int i = run.start; i < end; ++i) {
                const HaloRec& h = halos[i];
#pragma unroll
                for (int k = 0; k < PX; ++k) grad_piemd<true>(h, px[k], py, sx[k], sy[k]);
            }
            break;
        case TYPE_PIEMD:
            for (int i = run.start; i < end; ++i) {
                const HaloRec& h = halos[i];
#pragma unroll
                for (int k = 0; k < PX; ++k) grad_piemd<false>(h, px[k], py, sx[k], sy[k]);
            }
            break;
        default:  // TYPE_SIS
            for (int i = run.start; i < end; ++i) {
                const HaloRec& h = halos[i];
#pragma unroll
                for (int k = 0; k < PX; ++k) grad_sis(h, px[k], py, sx[k], sy[k]);
            }
            break;
        }
#pragma unroll
        for (int k = 0; k < PX; ++k) {
            tx[k] += sx[k];
            ty[k] += sy[k];
        }
    }
    const double m = o.kind == LT_MAP_DPL ? o.r : 1.0;
#pragma unroll
    for (int k = 0; k < PX; ++k) {
        const int col = col0 + k * WAVE;
        if (col < g.nx) {
            const int64_t idx = (int64_t)row * g.pitch + col;
            o.p0[idx] = tx[k] * m;
            o.p1[idx] = ty[k] * m;
        }
    }
}

__global__ __launch_bounds__(WAVE* ROWS) void k_potential64(const HaloRec* __restrict__ halos,
                                                            const Run* __restrict__ runs, int nruns,
                                                            GridDesc g, OutDesc o)
{
    const int row = blockIdx.y * ROWS + threadIdx.y;
    const int col = blockIdx.x * WAVE + threadIdx.x;
    if (row >= g.ny || col >= g.nx) return;
    halos = select_sample(g, halos, o);
    const double px = pixel_coord(g.xmin, col + g.istart, g.dx);
    const double py = pixel_coord(g.ymin, row + g.jstart, g.dy);
    double tot = 0.;
    for (int r = 0; r < nruns; ++r) {
        const Run run = runs[r];
        double s = 0.;
        for (int i = run.start; i < run.start + run.count; ++i) s += pot_halo(halos[i], run.type, px, py);
        tot += s;
    }
    o.p0[(int64_t)row * g.pitch + col] = tot * o.mult;
}

// Gradient at arbitrary points for a batch of lens models (blockIdx.y = sample), e.g. at the
// observed images for the source barycentres. Same per-run accumulation as the grid kernel.
__global__ __launch_bounds__(256) void k_points_gradient64(const HaloRec* __restrict__ halos,
                                                           const Run* __restrict__ runs, int nruns,
                                                           const double* __restrict__ px,
                                                           const double* __restrict__ py, int npts,
                                                           int64_t halo_stride, double* __restrict__ gx,
                                                           double* __restrict__ gy)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t s = blockIdx.y;
    if (i >= npts) return;
    halos += s * halo_stride;
    const double x = px[i], y = py[i];
    double tx = 0., ty = 0.;
    for (int r = 0; r < nruns; ++r) {
        const Run run = runs[r];
        double sx = 0., sy = 0.;
        const int end = run.start + run.count;
        switch (run.type) {
        case TYPE_DPIE:
            for (int h = run.start; h < end; ++h) grad_piemd<true>(halos[h], x, y, sx, sy);
            break;
        case TYPE_PIEMD:
            for (int h = run.start; h < end; ++h) grad_piemd<false>(halos[h], x, y, sx, sy);
            break;
        default:
            for (int h = run.start; h < end; ++h) grad_sis(halos[h], x, y, sx, sy);
            break;
        }
        tx += sx;
        ty += sy;
    }
    gx[s * npts + i] = tx;
    gy[s * npts + i] = ty;
}

// Epilogue alone, on Hessian planes already in device memory (the reference's map functions
// amplif_k_grid_CPU_GPU & co. work this way, src/grid_map_ampli_GPU.cu:305-547).
template <int KIND>
__global__ __launch_bounds__(256) void k_epilogue(const double* __restrict__ a, const double* __restrict__ b,
                                                  const double* __restrict__ c, int64_t n, OutDesc o)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const double av = a[i], cv = c[i];
        hess_epilogue<KIND>(o, i, av + cv, av - cv, b[i]);
    }
}

// ------------------------------------------------------------------------------------------------
// launchers
// ------------------------------------------------------------------------------------------------
static dim3 grid_for(const GridDesc& g, int px)
{
    return dim3((unsigned)((g.nx + WAVE * px - 1) / (WAVE * px)), (unsigned)((g.ny + ROWS - 1) / ROWS),
                (unsigned)(g.nsamples > 0 ? g.nsamples : 1));
}

template <int PX>
static hipError_t launch_hessian64_px(const HaloRec* halos, const Run* runs, int nruns, const GridDesc& g,
                                      const OutDesc& o, hipStream_t s)
{
    dim3 blk(WAVE, ROWS, 1), grd = grid_for(g, PX);
    switch (o.kind) {
    case LT_MAP_FIELDS: hipLaunchKernelGGL((k_hessian64<PX, LT_MAP_FIELDS>), grd, blk, 0, s, halos, runs, nruns, g, o); break;
    case LT_MAP_AMPLI: hipLaunchKernelGGL((k_hessian64<PX, LT_MAP_AMPLI>), grd, blk, 0, s, halos, runs, nruns, g, o); break;
    case LT_MAP_SHEAR: hipLaunchKernelGGL((k_hessian64<PX, LT_MAP_SHEAR>), grd, blk, 0, s, halos, runs, nruns, g, o); break;
    case LT_MAP_MASS: hipLaunchKernelGGL((k_hessian64<PX, LT_MAP_MASS>), grd, blk, 0, s, halos, runs, nruns, g, o); break;
    default: hipLaunchKernelGGL((k_hessian64<PX, LT_MAP_HESSIAN>), grd, blk, 0, s, halos, runs, nruns, g, o); break;
    }
    return hipGetLastError();
}

hipError_t launch_hessian64(const HaloRec* halos, const Run* runs, int nruns, const GridDesc& g,
                            const OutDesc& o, int px, hipStream_t s)
{
    switch (px) {
    case 1: return launch_hessian64_px<1>(halos, runs, nruns, g, o, s);
    case 4: return launch_hessian64_px<4>(halos, runs, nruns, g, o, s);
    default: return launch_hessian64_px<2>(halos, runs, nruns, g, o, s);
    }
}

hipError_t launch_epilogue64(const double* a, const double* b, const double* c, int64_t n, const OutDesc& o,
                             hipStream_t s)
{
    int64_t nb = (n + 255) / 256;
    unsigned blocks = (unsigned)(nb > 16384 ? 16384 : (nb ? nb : 1));
    dim3 blk(256), grd(blocks);
    switch (o.kind) {
    case LT_MAP_FIELDS: hipLaunchKernelGGL((k_epilogue<LT_MAP_FIELDS>), grd, blk, 0, s, a, b, c, n, o); break;
    case LT_MAP_AMPLI: hipLaunchKernelGGL((k_epilogue<LT_MAP_AMPLI>), grd, blk, 0, s, a, b, c, n, o); break;
    case LT_MAP_SHEAR: hipLaunchKernelGGL((k_epilogue<LT_MAP_SHEAR>), grd, blk, 0, s, a, b, c, n, o); break;
    case LT_MAP_MASS: hipLaunchKernelGGL((k_epilogue<LT_MAP_MASS>), grd, blk, 0, s, a, b, c, n, o); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_points_gradient64(const HaloRec* halos, const Run* runs, int nruns, const double* px,
                                   const double* py, int npts, int nsamples, int64_t halo_stride, double* gx,
                                   double* gy, hipStream_t s)
{
    hipLaunchKernelGGL(k_points_gradient64, dim3((npts + 255) / 256, nsamples), dim3(256), 0, s, halos, runs, nruns,
                       px, py, npts, halo_stride, gx, gy);
    return hipGetLastError();
}

hipError_t launch_gradient64(const HaloRec* halos, const Run* runs, int nruns, const GridDesc& g,
                             const OutDesc& o, int px, hipStream_t s)
{
    dim3 blk(WAVE, ROWS, 1);
    switch (px) {
    case 1: hipLaunchKernelGGL(k_gradient64<1>, grid_for(g, 1), blk, 0, s, halos, runs, nruns, g, o); break;
    case 4: hipLaunchKernelGGL(k_gradient64<4>, grid_for(g, 4), blk, 0, s, halos, runs, nruns, g, o); break;
    default: hipLaunchKernelGGL(k_gradient64<2>, grid_for(g, 2), blk, 0, s, halos, runs, nruns, g, o); break;
    }
    return hipGetLastError();
}

hipError_t launch_potential64(const HaloRec* halos, const Run* runs, int nruns, const GridDesc& g,
                              const OutDesc& o, hipStream_t s)
{
    dim3 blk(WAVE, ROWS, 1);
    hipLaunchKernelGGL(k_potential64, grid_for(g, 1), blk, 0, s, halos, runs, nruns, g, o);
    return hipGetLastError();
}

}  // namespace lt
