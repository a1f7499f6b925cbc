// lt_terms64.h -- FP64 per-(halo, pixel) terms, map epilogues and launch geometry shared by the
// FP64 kernels (lt_kernels64.hip) and the FP32-compensated kernels (lt_kernels32.hip), which fall
// back to these terms near halo cores.
//
// Formulas (restated from the mathematics; reference equivalents cited):
//   SIS gradient  src/gradient.cpp:260-299      PIEMD gradient  src/gradient.hpp:64-76,
//   dPIE gradient src/gradient.cpp:744-809      PIEMD Hessian   src/gradient2.cpp:54-82, 92-205
//   potential     src/potential_GPU.cu:47-116   epilogues       src/grid_map_*_GPU.cu
#pragma once

#include <hip/hip_runtime.h>

#include "lenstool_hip/lt_status.h"
#include "lt_device.h"
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

// One I_{0.5} term of the PIEMD Hessian in the halo frame (Kassiola & Kovner 1993) in the
// division-free form: with s = sqrt(e), a = 1/(1+e), b = 1/(1-e), cx1 = a/b and
// w = sqrt(r^2 + a^2 x^2 + b^2 y^2), the reference's N / D (src/gradient2.cpp:54-82) equals Q / S with
// Q = r + cx1 w + 2 i s a^2 x and S = w + r cx1 + 2 s a b y, and with U = b^2 y / w + 2 s a b,
// t = 2 s w - y / cx1 the second derivatives are, up to the factor 2 s a^2 ci:
//   phi_xx ~ (Re Q - cx1 a^2 x^2 / w) / |Q|^2,  phi_xy ~ x (2 s a^2 - cx1 U) / |Q|^2,
//   phi_yy ~ (Re Q + U t) / |Q|^2.
// Re Q >= r > 0: unlike the closed form, whose two pieces are both singular at halo-frame
// (0, 2 r s) and cancel there, nothing cancels anywhere. The trace needs no quotient: (xx + yy) / |Q|^2
// = 1 / W exactly (the term's convergence, elliptical in w; with the scaled W below), so only xx - yy
// and xy are divided by |Q|^2. Returned: the numerators of xx - yy and xy, |Q|^2 and 1 / W (the
// caller forms the reciprocals of |Q|^2: one for both radii of a dPIE).
struct PiemdQ {
    double dd, xy, q2, iW;
};

// Evaluated in scaled halo-frame coordinates X = cx1 a x, Y = cx1 b y (the scales are folded into
// the rotation constants csx .. sny) with W = cx1 w: then W^2 = cx1^2 r^2 + X^2 + Y^2,
// Re Q = W + r, and with U' = U / b = Y / W + 2 s a,
//   xx = Re Q - X^2 / W,  xy = k1 X - k2 X U',  yy = Re Q + U' (k3 W - k4 Y),  Im Q = k1 X,
// k1 = 2 s a / cx1, k2 = b / a, k3 = 2 s b / cx1, k4 = 1 / cx1^2, k5 = 2 s a: every pixel-dependent
// quantity is one operation from X, Y and W. xx - yy = -(X^2 / W + U' (k3 W - k4 Y)): Re Q drops out.
struct PiemdShared {
    double wq0, Qi2, A, Y, Qi, kx2, ky4;
};

__device__ __forceinline__ PiemdShared piemd_shared(const HaloRec& h, double px, double py)
{
    const double tx = px - h.xh, ty = py - h.yh;
    const double X = __builtin_fma(tx, h.csx, ty * h.snx);
    PiemdShared s;
    s.Y = __builtin_fma(ty, h.csy, -(tx * h.sny));
    s.A = X * X;
    s.wq0 = __builtin_fma(s.Y, s.Y, s.A);
    s.Qi = h.k1 * X;
    s.Qi2 = s.Qi * s.Qi;
    s.kx2 = h.k2 * X;
    s.ky4 = h.k4 * s.Y;
    return s;
}

// one core radius r (r2c = cx1^2 r^2)
__device__ __forceinline__ PiemdQ piemd_q_term(const HaloRec& h, const PiemdShared& s, double r, double r2c)
{
    const double W2 = s.wq0 + r2c;
    const double iW = rsq64_n2(W2);
    const double W = W2 * iW;
    const double Qr = W + r;
    const double U = __builtin_fma(s.Y, iW, h.k5);
    PiemdQ o;
    o.dd = -__builtin_fma(s.A, iW, U * __builtin_fma(h.k3, W, -s.ky4));
    o.xy = __builtin_fma(-s.kx2, U, s.Qi);
    o.q2 = __builtin_fma(Qr, Qr, s.Qi2);
    o.iW = iW;
    return o;
}

template <bool CUT>
__device__ __forceinline__ void hess_piemd(const HaloRec& h, double px, double py, double& aT, double& aD,
                                           double& aB)
{
    const PiemdShared s = piemd_shared(h, px, py);
    const PiemdQ a = piemd_q_term(h, s, h.rc, h.ra2c);
    double dT, dD, dB;
    if (CUT) {
        const PiemdQ b = piemd_q_term(h, s, h.rcut, h.rb2c);
        // one reciprocal for both radii: 1/|Qa|^2 = |Qb|^2 / (|Qa|^2 |Qb|^2), and the same for b
        const double ip = rcp64_n2(a.q2 * b.q2);
        const double ia = b.q2 * ip, ib = a.q2 * ip;
        dT = a.iW - b.iW;
        dD = __builtin_fma(a.dd, ia, -(b.dd * ib));
        dB = __builtin_fma(a.xy, ia, -(b.xy * ib));
    } else {
        const double ia = rcp64_n2(a.q2);
        dT = a.iW;
        dD = a.dd * ia;
        dB = a.xy * ia;
    }
    // dT, dD, dB: xx + yy, xx - yy and xy of the unscaled halo-frame derivatives (the scaling cancels)
    aT = __builtin_fma(h.qT, dT, aT);
    aD = __builtin_fma(h.qc2, dD, __builtin_fma(h.qs2m2, dB, aD));
    aB = __builtin_fma(h.qs2h, dD, __builtin_fma(h.qc2, dB, aB));
}

// SIS / SIE: phi = b0 R, R^2 = (1-e) x^2 + (1+e) y^2; halo-frame Hessian
// amph/R^3 (y^2, -xy, x^2) with amph = b0 (1-e)(1+e).
__device__ __forceinline__ void hess_sis(const HaloRec& h, double px, double py, double& aT, double& aD,
                                         double& aB)
{
    PixelFrame f = halo_frame(h, px, py);
    double R2 = __builtin_fma(h.sa, f.x2, h.sb * f.y2);
    double iR = rsq64_n2(R2);
    double g = iR * iR * iR;
    double dyx = f.y2 - f.x2;
    double xy = f.x * f.y;
    aT = __builtin_fma(g * h.amph, f.x2 + f.y2, aT);
    aD = __builtin_fma(g, __builtin_fma(h.hc2, dyx, 2.0 * h.hs2 * xy), aD);
    aB = __builtin_fma(g, __builtin_fma(0.5 * h.hs2, dyx, -(h.hc2 * xy)), aB);
}

// Circular SIS (e = 0): the Hessian b0/r^3 (y^2, -xy, x^2) holds in any frame, so it is formed from
// the lab-frame offsets, with no rotation to the halo frame and back (T = b0 / r).
__device__ __forceinline__ void hess_sis_circ(const HaloRec& h, double px, double py, double& aT, double& aD,
                                              double& aB)
{
    const double tx = px - h.xh, ty = py - h.yh;
    const double x2 = tx * tx, y2 = ty * ty;
    const double iR = rsq64_n2(x2 + y2);
    const double ga = h.amph * (iR * iR * iR);
    aT = __builtin_fma(h.amph, iR, aT);
    aD = __builtin_fma(ga, y2 - x2, aD);
    aB = __builtin_fma(-ga, tx * ty, aB);
}

// ------------------------------------------------------------------------------------------------
// gradient terms, accumulated in the lab frame
// ------------------------------------------------------------------------------------------------

// N conj(D) of the I_{0.5} ratio Z = N/D (N = cx1 x + i t, D = x + i n2); its real part is
// non-negative everywhere, so arg Z lies in [-pi/2, pi/2].
// Z = N / D of the reference's I_{0.5} ratio (src/gradient.hpp:64-76), N = cx1 x + i (2 s w - y / cx1),
// D = x + i (2 r s - y) with s = sqrt(e), w = sqrt(r^2 + a^2 x^2 + b^2 y^2), a = 1/(1+e), b = 1/(1-e),
// taken in the form of the Hessian's division-free derivation (lt_kernels32.hip):
//   N / D = Q / S,   Q = r + cx1 w + i 2 s a^2 x,   S = w + r cx1 + 2 s a b y > 0,
// so arg Z = arg Q (Re Q >= r > 0) and ln|Z| = ln(|Q|^2 / S^2) / 2. N and D both vanish at the
// halo-frame point (0, 2 r s) and are differences of nearly equal numbers next to it, where the closed
// form loses up to 1e-7 relative; Q and S have no such cancellation.
struct QS64 {
    double R, Q2, S;  // Re Q, |Q|^2, S
};

__device__ __forceinline__ QS64 qs64(const HaloRec& h, double rem2, double I2, double vy, double r, double r2,
                                     double rcx)
{
    const double q = r2 + rem2;
    const double w = q * rsq64_n2(q);
    QS64 o;
    o.R = __builtin_fma(h.cx1, w, r);
    o.Q2 = __builtin_fma(o.R, o.R, I2);
    o.S = (w + rcx) + vy;
    return o;
}

// PIEMD (CUT = false) or dPIE (CUT = true): (ampg arg, -ampg ln|.|) of Z(rc) or Z(rc) / Z(rcut). For
// the dPIE the two logarithms and the two arguments are merged into one of each, with I = Im Q the same
// for both radii: arg(Za / Zb) = atan2(I (Rb - Ra), Ra Rb + I^2) (Ra Rb + I^2 > 0: no unwrapping) and
// ln|Za / Zb| = ln(|Qa|^2 Sb^2 / (|Qb|^2 Sa^2)) / 2.
template <bool CUT>
__device__ __forceinline__ void grad_piemd(const HaloRec& h, double px, double py, double& gx, double& gy)
{
    PixelFrame f = halo_frame(h, px, py);
    const double rem2 = __builtin_fma(f.x2, h.kx, f.y2 * h.ky);
    const double I = h.gki * f.x;
    const double I2 = I * I;
    const double vy = h.gkv * f.y;
    const QS64 a = qs64(h, rem2, I2, vy, h.rc, h.ra2, h.rca);
    double arg, ratio;
    if (CUT) {
        const QS64 b = qs64(h, rem2, I2, vy, h.rcut, h.rb2, h.rcb);
        arg = atan2p_lt(I * (b.R - a.R), __builtin_fma(a.R, b.R, I2));
        ratio = (a.Q2 * (b.S * b.S)) * rcp64_n2(b.Q2 * (a.S * a.S));
    } else {
        arg = atan2p_lt(I, a.R);
        ratio = a.Q2 * rcp64_n2(a.S * a.S);
    }
    const double lnr = 0.5 * log_lt(ratio);
    const double re = h.ampg * arg;
    const double im = -h.ampg * lnr;
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

// Circular SIS: b0 (x, y) / r in any frame, formed from the lab-frame offsets.
__device__ __forceinline__ void grad_sis_circ(const HaloRec& h, double px, double py, double& gx, double& gy)
{
    const double tx = px - h.xh, ty = py - h.yh;
    const double s = h.ampg * rsq64(__builtin_fma(tx, tx, ty * ty));
    gx = __builtin_fma(s, tx, gx);
    gy = __builtin_fma(s, ty, gy);
}

// ------------------------------------------------------------------------------------------------
// potential (the I_{0.5} potential of Kassiola & Kovner, src/potential_GPU.cu:47-72)
// ------------------------------------------------------------------------------------------------
// The reference evaluates the potential of a core radius r with complex cosh / sinh / log of
// eta = -asinh(e1 y / z) / 2 + i asin(e2 x / z) / 2 and zeta = ln((rem + sqrt(r^2 + rem^2)) / r) / 2
// (two complex logarithms and their arguments per radius). The same function is
//   phi_r = x alpha_x + y alpha_y - 2 sqrt(e) r ln((1 + e) |Q| / (2 r))            (halo frame)
// with (alpha_x, alpha_y) = (arg Q, ln(S / |Q|)) the I_{0.5} deflection of the gradient (grad_piemd,
// without its amplitude), Q = r + cx1 w + i 2 s a^2 x, S = w + r cx1 + 2 s a b y, s = sqrt(e),
// a = 1/(1+e), b = 1/(1-e), w = sqrt(r^2 + a^2 x^2 + b^2 y^2). It is the cored isothermal-ellipsoid
// potential (x . grad phi minus the logarithmic term that makes its Laplacian 2 kappa): both forms have
// the same gradient, and both vanish in the limit at the centre, where Q = c r with c = 1 + cx1 =
// 2 / (1 + e), so they are one function (they agree to a few 1e-15 relative, checked at 40 digits
// against the extended-precision reference form by tests/test_potential_form.py). With
// d = w - r = rem^2 / (w + r), Re Q = c r + cx1 d and S = c r + d + v (v = 2 s a b y), the logarithms
//   lq = ln(|Q|^2 / (c r)^2) = log1p((cx1 d (2 c r + cx1 d) + (Im Q)^2) / (c r)^2),
//   ls = ln(S / (c r))       = log1p((d + v) / (c r))
// have arguments formed without cancellation, and phi_r = x arg Q + y (ls - lq / 2) - s r lq (PIEMD).
// For the dPIE (core radius a minus cut radius b) the deflection parts are merged, with no difference
// of nearly equal numbers near the centre or far outside the cut radius:
//   * Rb - Ra = D (c - cx1 sigma), D = rcut - rcore, sigma = (da + db) / (wa + wb) in [0, 1)
//     (db - da = -D sigma), so arg(Qa / Qb) = atan2(I (Rb - Ra), Ra Rb + I^2);
//   * alpha_y,a - alpha_y,b = ln(Sa^2 |Qb|^2 / (Sb^2 |Qa|^2)) / 2 = log1p(N / (Sb^2 |Qa|^2)) / 2 with
//     N = Pa (|Qb|^2 - |Qa|^2) + |Qa|^2 (Pa - Pb), P = S^2 - |Q|^2 = g (S + Re Q) - I^2,
//     g = S - Re Q = (1 - cx1) d + v, |Qb|^2 - |Qa|^2 = (Rb - Ra)(Rb + Ra) and
//     Pa - Pb = -ga c D (2 - sigma) + (Sb + Rb)(1 - cx1) D sigma;
// leaving per dPIE halo x pixel one arctangent and three logarithms, against three of each in the
// reference's form. Measured in a float64 emulation against 30-digit values, over halos of
// e <= 0.63 (mass ellipticity up to 0.9) at distances 1e-3 to 1e3 cut radii: within 2.1e-14 of the
// potential (5e-13 at e = 0.85, where S = w + v + cx1 r is itself a difference next to the -y axis;
// the PIEMD form within 2.6e-14 there).
// At the halo centre itself the reference's potential is 0/0 (NaN); this form gives its limit 0 there,
// so the kernels track the smallest rem^2 of a pixel over its PIEMD halos and return NaN when it is 0.
struct PotRad {
    double w, d, R, lq;  // w, w - r, Re Q, ln(|Q|^2 / (c r)^2)
};

__device__ __forceinline__ PotRad pot_rad(const HaloRec& h, double rem2, double I2, double r, double r2, double cr,
                                          double icr2)
{
    const double q = r2 + rem2;
    PotRad o;
    o.w = q * rsq64_n2(q);
    o.d = rem2 * rcp64_n2(o.w + r);
    const double cd = h.cx1 * o.d;
    o.R = cd + cr;
    o.lq = log1p_lt(__builtin_fma(cd, __builtin_fma(2., cr, cd), I2) * icr2);
    return o;
}

// pi05 of a PIEMD (type 8) or dPIE (type 81) halo without its amplitude; mn = min(mn, rem^2).
template <bool CUT>
__device__ __forceinline__ double pot_piemd(const HaloRec& h, const PixelFrame& f, double& mn)
{
    const double rem2 = __builtin_fma(f.x2, h.kx, f.y2 * h.ky);
    mn = min64(mn, rem2);
    const double I = h.gki * f.x;
    const double I2 = I * I;
    const double v = h.gkv * f.y;
    const PotRad a = pot_rad(h, rem2, I2, h.rc, h.ra2, h.pcra, h.picra2);
    if (!CUT) {
        const double ls = log1p_lt((a.d + v) * h.picra);
        const double arg = atan2p_lt(I, a.R);
        const double ay = __builtin_fma(-0.5, a.lq, ls);
        return __builtin_fma(f.x, arg, __builtin_fma(f.y, ay, -(h.psra * a.lq)));
    }
    const PotRad b = pot_rad(h, rem2, I2, h.rcut, h.rb2, h.pcrb, h.picrb2);
    const double sig = (a.d + b.d) * rcp64_n2(a.w + b.w);
    const double dR = h.pdr * __builtin_fma(-h.cx1, sig, h.pc);
    const double arg = atan2p_lt(I * dR, __builtin_fma(a.R, b.R, I2));
    const double Q2a = __builtin_fma(a.R, a.R, I2);
    const double Sa = (h.pcra + a.d) + v, Sb = (h.pcrb + b.d) + v;
    const double ga = __builtin_fma(h.pomc, a.d, v);
    const double Pa = __builtin_fma(ga, Sa + a.R, -I2);
    const double dP = __builtin_fma(Sb + b.R, h.pomcdr * sig, -(ga * (h.pcdr * (2. - sig))));
    const double N = __builtin_fma(Pa, dR * (a.R + b.R), Q2a * dP);
    const double ay = 0.5 * log1p_lt(N * rcp64_n2((Sb * Sb) * Q2a));
    const double L = __builtin_fma(h.psra, a.lq, -(h.psrb * b.lq));
    return __builtin_fma(f.x, arg, __builtin_fma(f.y, ay, -L));
}

// One halo's potential at (px, py) (SIS: b0 R, R^2 = (1-e) x^2 + (1+e) y^2).
__device__ __forceinline__ double pot_halo(const HaloRec& h, int type, double px, double py, double& mn)
{
    const PixelFrame f = halo_frame(h, px, py);
    if (type == TYPE_SIS) return h.ampg * sqrt(__builtin_fma(h.sa, f.x2, h.sb * f.y2));
    const double v = (type == TYPE_PIEMD) ? pot_piemd<false>(h, f, mn) : pot_piemd<true>(h, f, mn);
    return h.ampg * v;
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

// Launch geometry: a 64 x 4 work-group, PX pixels per thread along a row, blockIdx.z = sample.
inline dim3 grid_for(const GridDesc& g, int px)
{
    return dim3((unsigned)((g.nx + WAVE * px - 1) / (WAVE * px)), (unsigned)((g.ny + ROWS - 1) / ROWS),
                (unsigned)(g.nsamples > 0 ? g.nsamples : 1));
}

}  // namespace lt
