// lt_delense_geom.h -- triangle geometry of the delensing scan, for the device kernels and for host
// checks of the same code.
//
// cell_hit() is the reference's test (src/delense_CPU.cpp:60-112, src/delense_CPU_utils.hpp:81-214):
// FP64, contraction off, operation for operation. far_from_cell() is a cheaper test that, when it
// returns true, proves that both of cell_hit()'s triangle tests fail for that source, so the scan can
// skip them without changing a single result.
//
// Why far_from_cell() is exact. Let a, b, c be the source-plane corners of a triangle as cell_hit()
// computes them, P the source, d = cross(b - a, c - a) and e_i the three edge determinants, all exact
// reals, and M a bound on every coordinate involved. tri_det() evaluates six products and five sums
// in FP64, so each computed value is within E = 6 g6 M^2 < 4.0e-15 M^2 of the exact one (g6 = 6u/(1-6u),
// u = 2^-53). Let R bound the distance from corner g of the cell to its other corners (both triangles
// lie in the disk of radius R about it) and let P be farther than (1 + eps) R from that corner. P's
// distance m to either triangle then exceeds eps R while the triangle's diameter D is at most 2R.
// With barycentric coordinates al_i = e_i / d (sum 1), the negative ones sum to at least
// m / D > eps / 2 (P - Q = s (Q - Q') for two points Q, Q' of the triangle, s that sum), so one e_i has
// the sign opposite to d and |e_i| > eps |d| / 4. Hence, whenever |d| > 8E / eps, |e_i| exceeds 2E:
// the computed d and e_i carry the exact signs, their product is negative and far from underflow
// (M >= 1e-50), and the reference's tri_inside() returns false. |d| is bounded from below by a
// quadratic in D_ls/D_s whose coefficients are set up once per cell: d(dr) = c0 + dr (c1 + dr c2) for
// the corners p_i - dr g_i; its FP64 evaluation, and the rounding of the corners themselves, stay
// within 3.2e-14 M^2 of the exact d. Two settings are used:
//   near: eps = 1e-3, distance above 1.0011 R, |d(dr)| > 1e-10 M^2 (8E/eps = 3.2e-11 M^2);
//   far:  eps = 1,    distance above 2.0001 R, |d(dr)| > 1e-13 M^2 (8E/eps = 3.2e-14 M^2);
// the extra 1e-4 R in each distance (R >= 1e-9 M) covers the rounding of the distance itself. The
// scan's fast path uses the near setting, which holds for every cell that is not almost singular;
// the far one catches tiny or nearly singular cells before the reference test. Non-finite or huge
// gradients, sources or D_ls/D_s switch both off (the full reference test runs), so NaN and infinity
// behave exactly as in the reference.
#pragma once

#include <hip/hip_runtime.h>

namespace lt {

struct P2 {
    double x, y;
};

// scalar triple product of (A,1), (B,1), (C,1) (src/delense_CPU_utils.hpp:110-117)
__host__ __device__ inline double tri_det(P2 A, P2 B, P2 C)
{
#pragma clang fp contract(off)
    return (B.x * C.y - B.y * C.x + A.x * B.y - A.y * B.x + A.y * C.x - A.x * C.y);
}

// inside-or-on-border test (src/delense_CPU_utils.hpp:125-142; :180-193 is the same test)
__host__ __device__ inline bool tri_inside(P2 P, P2 a, P2 b, P2 c)
{
#pragma clang fp contract(off)
    double d = tri_det(a, b, c);
    if (tri_det(a, b, P) * d < 0.) return false;
    if (tri_det(b, c, P) * d < 0.) return false;
    if (tri_det(c, a, P) * d < 0.) return false;
    return true;
}

__host__ __device__ inline P2 to_source(double x, double y, double dr, double gx, double gy)
{
#pragma clang fp contract(off)
    P2 s;
    s.x = x - dr * gx;
    s.y = y - dr * gy;
    return s;
}

// The four corners of a cell: image-plane coordinates as the reference forms them and the gradient
// at g (x, y), g+n (x, y+dy), g+1 (x+dx, y), g+n+1 (x+dx, y+dy).
struct CellCorners {
    double xp, yp, xq, yq;
    double g0x, g0y, g1x, g1y, g2x, g2y, g3x, g3y;
};

// Which triangle of the cell holds source P after delensing: 0 none, 1 upper, 2 lower
// (src/delense_CPU.cpp:60-112: the upper triangle is tested first; the lower only if it misses).
__host__ __device__ inline int cell_hit(const CellCorners& c, double px, double py, double dr)
{
#pragma clang fp contract(off)
    const P2 P = {px, py};
    // upper triangle a=(x,y) b=(x,y+dy) c=(x+dx,y)
    const P2 ua = to_source(c.xp, c.yp, dr, c.g0x, c.g0y);
    const P2 ub = to_source(c.xp, c.yq, dr, c.g1x, c.g1y);
    const P2 uc = to_source(c.xq, c.yp, dr, c.g2x, c.g2y);
    if (tri_inside(P, ua, ub, uc)) return 1;
    // lower triangle a=(x+dx,y+dy) b=(x+dx,y) c=(x,y+dy)
    const P2 la = to_source(c.xq, c.yq, dr, c.g3x, c.g3y);
    const P2 lb = to_source(c.xq, c.yp, dr, c.g2x, c.g2y);
    const P2 lc = to_source(c.xp, c.yq, dr, c.g1x, c.g1y);
    return tri_inside(P, la, lb, lc) ? 2 : 0;
}

// Per-cell set-up of the cull for the sources of one sample (|x|, |y| <= pmax, drlo <= dr <= drhi):
// determinant polynomials of both triangles, the squared radius (1.0011 R)^2 (the far setting's
// (2.0001 R)^2 is rn2 (2.0001 / 1.0011)^2, rounded up) and the near threshold tau = 1e-10 M^2. R bounds |v_i - v_g| = |(p_i - p_g) - dr (g_i - g_g)| over the
// sample's D_ls/D_s range: each distance is convex in dr, so its largest value on [drlo, drhi] is at
// an end.
struct CellFilter {
    double c0, c1u, c2u, c1l, c2l, rn2, tau;
    double rn;  // an upper bound of sqrt(rn2)
    double m;   // the coordinate bound M
    bool on;
    bool d_ok;  // |d| > tau for every dr of the range (both triangles): the margin needs only rho
};

__host__ __device__ inline double abs_d(double v) { return __builtin_fabs(v); }
// fmax drops NaN; the filter checks finiteness separately before it relies on these maxima
__host__ __device__ inline double max_d(double a, double b) { return __builtin_fmax(a, b); }

// An upper bound of sqrt(q) for q >= 0 from the FP32 square root (relative error below 2^-22, the
// rounding of q to FP32 within 2^-24), enlarged by 1e-6 relative; q below 1e-30 counts as 1e-30 (the
// bound stays a normal FP32 number; a bound of 1e-15 is far below any length these bounds compare
// with). Four instructions instead of the FP64 square root's dozen; +infinity when q exceeds the FP32
// range, which switches the tests that use it off.
__host__ __device__ inline double sqrt_up(double q)
{
    return (double)__builtin_sqrtf((float)__builtin_fmax(q, 1e-30)) * (1. + 1e-6);
}

__host__ __device__ inline double corner_dist2(double px, double py, double gx, double gy, double dr)
{
    const double x = __builtin_fma(-dr, gx, px), y = __builtin_fma(-dr, gy, py);
    return __builtin_fma(x, x, y * y);
}

__host__ __device__ inline CellFilter cell_filter(const CellCorners& c, double pmax, double drlo, double drhi)
{
    CellFilter f;
    const double hx = c.xq - c.xp, hy = c.yq - c.yp;
    // corner gradient differences: G_i = g_i - g0 (upper), H_i = g_i - g3 (lower)
    const double G1x = c.g1x - c.g0x, G1y = c.g1y - c.g0y, G2x = c.g2x - c.g0x, G2y = c.g2y - c.g0y;
    const double G3x = c.g3x - c.g0x, G3y = c.g3y - c.g0y;
    const double H1x = c.g1x - c.g3x, H1y = c.g1y - c.g3y, H2x = c.g2x - c.g3x, H2y = c.g2y - c.g3y;
    // upper: cross((0,hy) - dr G1, (hx,0) - dr G2); lower: cross((0,-hy) - dr H2, (-hx,0) - dr H1)
    f.c0 = -(hx * hy);
    f.c1u = hy * G2x + hx * G1y;
    f.c2u = G1x * G2y - G1y * G2x;
    f.c1l = -(hy * H1x + hx * H2y);
    f.c2l = H2x * H1y - H2y * H1x;
    // finiteness (NaN and infinity propagate through the sums) and the coordinate bound M
    const double sg = abs_d(c.g0x) + abs_d(c.g0y) + abs_d(c.g1x) + abs_d(c.g1y) + abs_d(c.g2x) + abs_d(c.g2y) +
                      abs_d(c.g3x) + abs_d(c.g3y);
    const double sp = abs_d(c.xp) + abs_d(c.yp) + abs_d(c.xq) + abs_d(c.yq);
    const double mc1 = max_d(max_d(max_d(abs_d(c.g0x), abs_d(c.g0y)), max_d(abs_d(c.g1x), abs_d(c.g1y))),
                             max_d(max_d(abs_d(c.g2x), abs_d(c.g2y)), max_d(abs_d(c.g3x), abs_d(c.g3y))));
    const double mc0 = max_d(max_d(abs_d(c.xp), abs_d(c.yp)), max_d(abs_d(c.xq), abs_d(c.yq)));
    const double drmax = max_d(abs_d(drlo), abs_d(drhi));
    const double m = max_d(mc0 + drmax * mc1, pmax) * 1.000001;
    f.m = m;
    f.on = sg <= 1e100 && sp <= 1e100 && pmax <= 1e100 && drmax <= 1e100 && drlo <= drhi && m >= 1e-50 &&
           m <= 1e100;
    f.tau = 1e-10 * m * m;
    double r2 = 0.;
    for (int e = 0; e < 2; ++e) {
        const double dr = e ? drhi : drlo;
        r2 = max_d(r2, corner_dist2(0., hy, G1x, G1y, dr));
        r2 = max_d(r2, corner_dist2(hx, 0., G2x, G2y, dr));
        r2 = max_d(r2, corner_dist2(hx, hy, G3x, G3y, dr));
    }
    const double r = sqrt_up(r2) * 1.000001 + 1e-9 * m;
    f.rn2 = (1.0011 * r) * (1.0011 * r);
    f.rn = (1.0011 * r) * (1. + 1e-15);
    // |d| over the whole range: with dr = mid + t, |t| <= w, the quadratic is exactly
    // d(mid) + d'(mid) t + c2 t^2, so |d| >= |d(mid)| - |d'(mid)| w - |c2| w^2 (the factors 1 + 1e-12
    // and the slack 1e-13 M^2 cover the rounding of these few operations)
    const double mid = 0.5 * (drlo + drhi);
    const double w = max_d(drhi - mid, mid - drlo) * (1. + 1e-12);
    const double qu = __builtin_fma(mid, __builtin_fma(mid, f.c2u, f.c1u), f.c0);
    const double ql = __builtin_fma(mid, __builtin_fma(mid, f.c2l, f.c1l), f.c0);
    const double su = __builtin_fma(abs_d(__builtin_fma(2. * mid, f.c2u, f.c1u)), w, abs_d(f.c2u) * w * w);
    const double sl = __builtin_fma(abs_d(__builtin_fma(2. * mid, f.c2l, f.c1l)), w, abs_d(f.c2l) * w * w);
    const double slack = f.tau + 1e-13 * m * m;
    f.d_ok = f.on && abs_d(qu) > su * (1. + 1e-12) + slack && abs_d(ql) > sl * (1. + 1e-12) + slack;
    return f;
}

// Cull margin of the near setting, min(rho^2 - (1.0011 R)^2, min(|d_upper|, |d_lower|) - tau):
// positive when it applies. With f.on every input is finite and bounded (|values| <= 1e100, so every
// product stays finite), so the margin is never NaN and "margin > 0" is exactly "rho^2 > (1.0011 R)^2
// and both |d| > tau".
__host__ __device__ inline double far_margin(const CellFilter& f, const CellCorners& c, double px, double py, double dr)
{
    const double ax = __builtin_fma(-dr, c.g0x, c.xp);
    const double ay = __builtin_fma(-dr, c.g0y, c.yp);
    const double ex = px - ax, ey = py - ay;
    const double rho2 = __builtin_fma(ex, ex, ey * ey);
    const double du = __builtin_fma(dr, __builtin_fma(dr, f.c2u, f.c1u), f.c0);
    const double dl = __builtin_fma(dr, __builtin_fma(dr, f.c2l, f.c1l), f.c0);
    return __builtin_fmin(rho2 - f.rn2, __builtin_fmin(abs_d(du), abs_d(dl)) - f.tau);
}

// The same margin for a cell with d_ok: rho^2 - (1.0011 R)^2 alone (the |d| condition holds for the
// whole D_ls/D_s range).
__host__ __device__ inline double far_margin_rho(const CellFilter& f, const CellCorners& c, double px, double py,
                                                 double dr)
{
    const double ex = px - __builtin_fma(-dr, c.g0x, c.xp);
    const double ey = py - __builtin_fma(-dr, c.g0y, c.yp);
    return __builtin_fma(ex, ex, ey * ey) - f.rn2;
}

// FP32 pre-test of far_margin_rho for cells with d_ok. Every input rounded to FP32 (relative error
// u = 2^-24 of at most M: the coordinates, the sources and dr |g| are bounded by M), then
// ax = px - (xp - dr g0x) formed with one FMA and one subtraction: the FP32 offset differs from the
// FP64 one by at most 9 u M per axis (input roundings u M for xp, 2 u M for dr g0x, u M for px, two
// operation roundings within 2 u M and 3 u M), so rho differs by at most 13 u M < 7.8e-7 M. The squared
// radius rf2 is that of sqrt(rn2) + 1e-6 M, enlarged by a further 1e-6 relative for the rounding of
// rho^2 and of rf2 itself: a positive FP32 margin implies rho^2 > rn2 in exact arithmetic, which is
// what the FP64 margin establishes. Off (rf2 = +infinity) unless d_ok and M <= 1e30.
struct CellFilter32 {
    float xp, yp, g0x, g0y, rf2;
};

__host__ __device__ inline CellFilter32 cell_filter32(const CellFilter& f, const CellCorners& c)
{
    CellFilter32 g;
    g.xp = (float)c.xp;
    g.yp = (float)c.yp;
    g.g0x = (float)c.g0x;
    g.g0y = (float)c.g0y;
    const bool ok = f.d_ok && f.m <= 1e30;
    const double rf = f.rn + 1e-6 * f.m;
    g.rf2 = ok ? (float)(rf * rf * (1. + 1e-6)) : __builtin_inff();
    return g;
}

__host__ __device__ inline float far_margin_rho32(const CellFilter32& g, float px, float py, float dr)
{
#pragma clang fp contract(off)
    const float ex = px - __builtin_fmaf(-dr, g.g0x, g.xp);
    const float ey = py - __builtin_fmaf(-dr, g.g0y, g.yp);
    return __builtin_fmaf(ex, ex, ey * ey) - g.rf2;
}

// Tile-level pre-test: one FP32 test per (tile of cells, source) instead of one per (cell, source).
// For the cells i of a tile, all with d_ok, a reference cell r of the tile, and the exact differences
// delta_i = p_i - p_r, gamma_i = g0_i - g0_r of their corner g (position and gradient), the exact
// distance from a source P to cell i's mapped corner v_i(dr) = p_i - dr g0_i is at least
// |P - v_r(dr)| - |delta_i| - |dr| |gamma_i|. So if
//   |P - v_r(dr)| > T >= T_i = sqrt(rn2_i) + |delta_i| + drmax |gamma_i|   for every cell i,
// with drmax >= |dr| for every source of the sample, every cell's distance margin holds in exact
// arithmetic (rho_i^2 > rn2_i) and, with d_ok, both triangle tests of every cell of the tile fail.
// tile_reach() forms T_i from upper bounds of the two norms (sqrt_up) and of sqrt(rn2) (rn), with a few
// roundings (within 8 u of the sum of those bounds); tile_filter32() enlarges
// the tile's maximum T by 1e-12 relative and builds the FP32 test of far_margin_rho32() about the
// reference cell's corner, with that test's own slack (1e-6 M_r, M_r the reference cell's M, which
// bounds every FP32 input: its corner, dr |g0| and the sources) and relative enlargement: a positive
// margin implies |P - v_r(dr)| > T exactly. Off (rf2 = +infinity) unless ok and M_r <= 1e30.
__host__ __device__ inline double tile_reach(const CellFilter& f, const CellCorners& c, double xr, double yr,
                                             double gxr, double gyr, double drmax)
{
    const double ax = c.xp - xr, ay = c.yp - yr, bx = c.g0x - gxr, by = c.g0y - gyr;
    const double dx = sqrt_up(__builtin_fma(ax, ax, ay * ay));
    const double dg = sqrt_up(__builtin_fma(bx, bx, by * by));
    return __builtin_fma(drmax, dg, f.rn + dx);
}

__host__ __device__ inline CellFilter32 tile_filter32(double xr, double yr, double gxr, double gyr, double T, double Mr,
                                                      bool ok)
{
    CellFilter32 g;
    g.xp = (float)xr;
    g.yp = (float)yr;
    g.g0x = (float)gxr;
    g.g0y = (float)gyr;
    const double rf = T * (1. + 1e-12) + 1e-6 * Mr;
    g.rf2 = (ok && Mr <= 1e30) ? (float)(rf * rf * (1. + 1e-6)) : __builtin_inff();
    return g;
}

// true: both of cell_hit()'s triangle tests provably fail for source (px, py, dr), by the near or the
// far setting (dr must lie in the range the filter was set up for)
__host__ __device__ inline bool far_from_cell(const CellFilter& f, const CellCorners& c, double px, double py,
                                              double dr)
{
    const double ax = __builtin_fma(-dr, c.g0x, c.xp);
    const double ay = __builtin_fma(-dr, c.g0y, c.yp);
    const double ex = px - ax, ey = py - ay;
    const double rho2 = __builtin_fma(ex, ex, ey * ey);
    const double du = __builtin_fma(dr, __builtin_fma(dr, f.c2u, f.c1u), f.c0);
    const double dl = __builtin_fma(dr, __builtin_fma(dr, f.c2l, f.c1l), f.c0);
    const double dmin = __builtin_fmin(abs_d(du), abs_d(dl));
    const bool near_ok = (rho2 > f.rn2) & (dmin > f.tau);
    const bool far_ok = (rho2 > 3.9919 * f.rn2) & (dmin > 1e-3 * f.tau);  // (2.0001 / 1.0011)^2 = 3.99161..
    return f.on & (near_ok | far_ok);
}

}  // namespace lt
