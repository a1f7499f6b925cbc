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
// (0, 2 r s) and cancel there, nothing cancels anywhere. Returned: the numerators of xx, xy, yy
// and |Q|^2 (the caller forms the reciprocals: one for both radii of a dPIE).
struct PiemdQ {
    double xx, xy, yy, q2;
};

// Evaluated in scaled halo-frame coordinates X = cx1 a x, Y = cx1 b y (the scales are folded into
// the rotation constants csx .. sny) with W = cx1 w: then W^2 = cx1^2 r^2 + X^2 + Y^2,
// Re Q = W + r, and with U' = U / b = Y / W + 2 s a,
//   xx = Re Q - X^2 / W,  xy = k1 X - k2 X U',  yy = Re Q + U' (k3 W - k4 Y),  Im Q = k1 X,
// k1 = 2 s a / cx1, k2 = b / a, k3 = 2 s b / cx1, k4 = 1 / cx1^2, k5 = 2 s a: every pixel-dependent
// quantity is one operation from X, Y and W.
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
    o.xx = __builtin_fma(-s.A, iW, Qr);
    o.xy = __builtin_fma(-s.kx2, U, s.Qi);
    o.yy = __builtin_fma(U, __builtin_fma(h.k3, W, -s.ky4), Qr);
    o.q2 = __builtin_fma(Qr, Qr, s.Qi2);
    return o;
}

template <bool CUT>
__device__ __forceinline__ void hess_piemd(const HaloRec& h, double px, double py, double& aT, double& aD,
                                           double& aB)
{
    const PiemdShared s = piemd_shared(h, px, py);
    const PiemdQ a = piemd_q_term(h, s, h.rc, h.ra2c);
    double xx, xy, yy;
    if (CUT) {
        const PiemdQ b = piemd_q_term(h, s, h.rcut, h.rb2c);
        // one reciprocal for both radii: 1/|Qa|^2 = |Qb|^2 / (|Qa|^2 |Qb|^2), and the same for b
        const double ip = rcp64_n2(a.q2 * b.q2);
        const double ia = b.q2 * ip, ib = a.q2 * ip;
        xx = __builtin_fma(a.xx, ia, -(b.xx * ib));
        xy = __builtin_fma(a.xy, ia, -(b.xy * ib));
        yy = __builtin_fma(a.yy, ia, -(b.yy * ib));
    } else {
        const double ia = rcp64_n2(a.q2);
        xx = a.xx * ia;
        xy = a.xy * ia;
        yy = a.yy * ia;
    }
    // xx, xy, yy are the numerators of the unscaled halo-frame derivatives (the scaling cancels)
    const double dT = xx + yy, dD = xx - yy, dB = xy;
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
    const double w = q * rsq64(q);
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
        arg = atan2_lt(I * (b.R - a.R), __builtin_fma(a.R, b.R, I2));
        ratio = (a.Q2 * (b.S * b.S)) * rcp64(b.Q2 * (a.S * a.S));
    } else {
        arg = atan2_lt(I, a.R);
        ratio = a.Q2 * rcp64(a.S * a.S);
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
// potential (complex I_{0.5} potential of Kassiola & Kovner, src/potential_GPU.cu:47-72)
// ------------------------------------------------------------------------------------------------
// The reference's I_{0.5} potential (src/potential_GPU.cu:47-72) is written with complex
// cosh / sinh / log of eta = -asinh(e1 y / z) / 2 + i asin(e2 x / z) / 2 and the real
// zeta = ln((rem + sqrt(rc^2 + rem^2)) / rc) / 2. Every hyperbolic and circular function of eta and
// zeta it uses has an algebraic value: with p = e1 y / z, q = e2 x / z, P = sqrt(1 + p^2),
// C = sqrt(1 - q^2) and W = sqrt(rc^2 + rem^2),
//   cosh 2 eta = P C - i p q,   sinh 2 eta = -p C + i P q,   tanh eta = sinh 2 eta / (1 + cosh 2 eta),
//   sinh 2 zeta = rem / rc,     sinh^2 zeta = (W / rc - 1) / 2,   tanh zeta = rem / (rc + W),
// so that (identities cosh(a+b) cosh(a-b) = cosh^2 a + sinh^2 b and cosh(a+b) / cosh(a-b) =
// (1 + tanh a tanh b) / (1 - tanh a tanh b))
//   ln(cosh^2 eta / (cosh(eta+zeta) cosh(eta-zeta))) = ln((1 + P C - i p q) / (P C + W / rc - i p q))
//   ln(cosh(eta+zeta) / cosh(eta-zeta))             = ln((D + t N) / (D - t N)),
// N = -p C + i P q, D = 1 + P C - i p q, t = tanh zeta. Two more steps make it shorter and
// accurate next to the centre:
//  * the first logarithm is -ln(1 + z) with z = (W / rc - 1) / D and W / rc - 1 = rem^2 / (rc (rc + W))
//    = (rem / rc) t, formed without cancellation; ln|1 + z| = log1p(2 Re z + |z|^2) / 2. (Written as
//    the logarithm of a quotient of two nearly equal numbers, its relative error grows as rc^2 / r^2
//    at a distance r from the centre: about 1e-11 of the potential at r = rc / 100.)
//  * the second logarithm a2 = ln((D + t N) / (D - t N)) = ln(1 + z2), z2 = 2 t N / (D - t N), is
//    also taken as a log1p (its quotient tends to 1 at the centre, where t -> 0);
//  * for the dPIE, the second logarithms of the core and cut radii enter the difference with the same
//    weight, so they are one logarithm: ln((D + ta N)(D - tb N) / ((D - ta N)(D + tb N))) = ln(1 + z2)
//    with z2 = 2 (ta - tb) D N / ((D - ta N)(D + tb N)). Its argument is the difference of the two
//    arguments (no branch cut is crossed: each arg cosh(eta +- zeta) = atan(tanh(a +- zeta) tan b)
//    lies in (-pi/4, pi/4), eta = a + i b, |b| < pi/4, so each argument difference is in
//    (-pi/2, pi/2) and theirs in (-pi, pi)).
// Per dPIE halo x pixel: three logarithms and three arguments instead of four of each.
struct PotShared {
    double p, q, PC, pq, Nr, Ni, rem2, isrem, srem, onePC, kr, ki;
};

// pixel quantities shared by both core radii (per-halo constants from the record: the divisions
// and square roots of eps are the packer's)
__device__ __forceinline__ PotShared pot_shared(const HaloRec& h, const PixelFrame& f)
{
    const double iz = rsq64(f.x2 + f.y2);
    PotShared s;
    s.p = h.pkp * f.y * iz;
    s.q = h.pkq * f.x * iz;
    const double P2 = __builtin_fma(s.p, s.p, 1.), C2 = __builtin_fma(-s.q, s.q, 1.);
    const double P = P2 * rsq64(P2), C = C2 * rsq64(C2);
    s.PC = P * C;
    s.pq = s.p * s.q;
    s.Nr = -s.p * C;
    s.Ni = P * s.q;
    s.rem2 = __builtin_fma(f.x2, h.kx, f.y2 * h.ky);
    s.isrem = rsq64(s.rem2);
    s.srem = s.rem2 * s.isrem;
    // 1 / D = conj(D) / |D|^2 = (kr + i ki) with D = 1 + PC - i pq
    s.onePC = 1. + s.PC;
    const double iD2 = rcp64(__builtin_fma(s.onePC, s.onePC, s.pq * s.pq));
    s.kr = s.onePC * iD2;
    s.ki = s.pq * iD2;
    return s;
}

// The core-radius part of pi05: a1 = ln(cosh^2 eta / (cosh(eta+zeta) cosh(eta-zeta))) = -ln(1 + z),
// z = (rem / rc) t / D, and t = tanh zeta = rem / (rc + W).
struct PotCore {
    double a1r, a1i, t;
};

__device__ __forceinline__ PotCore pot_core(const PotShared& s, double rc, double irc)
{
    const double W2 = __builtin_fma(rc, rc, s.rem2);
    const double W = W2 * rsq64(W2);
    PotCore c;
    c.t = s.srem * rcp64(rc + W);
    const double w1 = (s.srem * irc) * c.t;
    const double zr = w1 * s.kr, zi = w1 * s.ki;
    c.a1r = -0.5 * log1p_lt(__builtin_fma(zi, zi, zr * (2. + zr)));
    c.a1i = -atan2_lt(zi, 1. + zr);
    return c;
}

// pi05 without the amplitude b0 ci, as rc / rem L(N a1) + L(a2) with L(c) = c.im x - c.re y:
// the reference's rc / rem ((c1 + c2).im x - (c1 + c2).re y), c1 = sinh(2 eta) a1 = N a1,
// c2 = (rem / rc) a2; for the dPIE the difference of the core and cut radii's values.
__device__ __forceinline__ double pot_halo(const HaloRec& h, int type, double px, double py)
{
    PixelFrame f = halo_frame(h, px, py);
    if (type == TYPE_SIS) return h.b0 * sqrt(__builtin_fma(h.sa, f.x2, h.sb * f.y2));
    const PotShared s = pot_shared(h, f);
    const PotCore a = pot_core(s, h.rc, h.irc);
    // a2 = ln(1 + z2), z2 = U / V: PIEMD U = 2 t N, V = D - t N; dPIE U = 2 (ta - tb) D N,
    // V = (D - ta N)(D + tb N)
    const double amr = __builtin_fma(-a.t, s.Nr, s.onePC), ami = __builtin_fma(-a.t, s.Ni, -s.pq);
    double cr, ci, Ur, Ui, Vr, Vi;
    if (type == TYPE_PIEMD) {
        Ur = 2. * a.t * s.Nr;
        Ui = 2. * a.t * s.Ni;
        Vr = amr;
        Vi = ami;
        cr = h.rc * a.a1r;
        ci = h.rc * a.a1i;
    } else {
        const PotCore b = pot_core(s, h.rcut, h.ircut);
        const double bpr = __builtin_fma(b.t, s.Nr, s.onePC), bpi = __builtin_fma(b.t, s.Ni, -s.pq);
        Vr = __builtin_fma(amr, bpr, -(ami * bpi));
        Vi = __builtin_fma(amr, bpi, ami * bpr);
        const double dt = 2. * (a.t - b.t);
        const double dnr = __builtin_fma(s.onePC, s.Nr, s.pq * s.Ni), dni = __builtin_fma(s.onePC, s.Ni, -(s.pq * s.Nr));
        Ur = dt * dnr;
        Ui = dt * dni;
        cr = __builtin_fma(h.rc, a.a1r, -(h.rcut * b.a1r));
        ci = __builtin_fma(h.rc, a.a1i, -(h.rcut * b.a1i));
    }
    const double iV2 = rcp64(__builtin_fma(Vr, Vr, Vi * Vi));
    const double z2r = __builtin_fma(Ur, Vr, Ui * Vi) * iV2, z2i = __builtin_fma(Ui, Vr, -(Ur * Vi)) * iV2;
    const double a2r = 0.5 * log1p_lt(__builtin_fma(z2i, z2i, z2r * (2. + z2r)));
    const double a2i = atan2_lt(z2i, 1. + z2r);
    // N (cr + i ci)
    const double nr = __builtin_fma(s.Nr, cr, -(s.Ni * ci)), ni = __builtin_fma(s.Nr, ci, s.Ni * cr);
    const double l1 = __builtin_fma(ni, f.x, -(nr * f.y));
    const double l2 = __builtin_fma(a2i, f.x, -(a2r * f.y));
    return h.ampg * __builtin_fma(s.isrem, l1, l2);
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
