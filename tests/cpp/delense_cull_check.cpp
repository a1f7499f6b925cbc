// Host check of the delensing scan's cull (lenstool-hpc_amd/csrc/lt_delense_geom.h), the same header
// the device kernel compiles:
//   1. soundness: whenever far_from_cell() (near or far setting) or the fast path's margins
//      (on && far_margin() > 0, d_ok && far_margin_rho() > 0) say "skip", the reference test
//      cell_hit() finds no hit, and d_ok implies |exact det| > 3.2e-11 M^2 for both triangles;
//   2. the two error bounds its proof rests on, against __float128 arithmetic:
//      |tri_det computed - exact| <= 4.0e-15 M^2 and |d(dr) polynomial - exact det of the computed
//      corners| <= 3.2e-14 M^2 (M = largest coordinate involved);
//   3. efficiency: on lens-like cells (0.05" or larger) far from the source the fast path's margin
//      decides almost every pair;
//   4. the FP32 pre-test (d_ok && far_margin_rho32() > 0) is sound on its own: whenever it skips,
//      cell_hit() finds no hit and the exact distance from the source to the cell's corner g exceeds
//      sqrt(rn2), which is what the FP64 margin establishes; on lens-like far pairs it decides about as
//      often as the FP64 margin;
//   5. the tile pre-test (tile_reach / tile_filter32 over an 8 x 8 tile of cells, all with d_ok): whenever
//      it skips a source, no cell of the tile has a hit and every cell's exact corner distance exceeds
//      sqrt(rn2); on lens-like tiles it skips most sources that no cell of the tile is near;
//   6. sqrt_up(q) is not below sqrt(q) (long double) over 1e-300 .. 1e300.
// Cells come from smooth lens-like gradient fields, near-collinear and collapsed mapped triangles,
// wild and huge gradients, sources on corners and edges, and NaN / infinity in gradients, sources and
// D_ls/D_s. Exit status 0 = all checks hold; the counts are printed.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <limits>
#include <random>

#include "lt_delense_geom.h"

using lt::CellCorners;
using lt::CellFilter;
using lt::P2;

typedef __float128 q;

static q qdet(P2 a, P2 b, P2 c)
{
    return ((q)b.x - a.x) * ((q)c.y - a.y) - ((q)b.y - a.y) * ((q)c.x - a.x);
}

static q q__sqrt(double v) { return (q)std::sqrt((long double)v); }

static double amax(double a, double b) { return std::fabs(a) > std::fabs(b) ? std::fabs(a) : std::fabs(b); }

struct Stats {
    long trials = 0, culled = 0, hits = 0, unsound = 0, det_bad = 0, poly_bad = 0, far_pairs = 0, far_culled = 0,
         dok = 0, dok_bad = 0, fast32 = 0, fp32_bad = 0, far_rho = 0, far_culled32 = 0;
    double det_worst = 0, poly_worst = 0;
};

static void check(Stats& st, const CellCorners& c, double px, double py, double dr, double pmax, double drlo,
                  double drhi, bool lens_like_far)
{
    const CellFilter f = lt::cell_filter(c, pmax, drlo, drhi);
    const bool far = lt::far_from_cell(f, c, px, py, dr);
    const bool fast = (f.on && lt::far_margin(f, c, px, py, dr) > 0.) ||
                      (f.d_ok && lt::far_margin_rho(f, c, px, py, dr) > 0.);
    const lt::CellFilter32 f32 = lt::cell_filter32(f, c);
    const bool fast32 = f.d_ok && lt::far_margin_rho32(f32, (float)px, (float)py, (float)dr) > 0.f;
    st.fast32 += fast32;
    if (fast32) {
        // the exact distance from the source to the exact corner g = (xp - dr g0x, yp - dr g0y)
        const q ex = (q)px - ((q)c.xp - (q)dr * c.g0x), ey = (q)py - ((q)c.yp - (q)dr * c.g0y);
        if (!(ex * ex + ey * ey > (q)f.rn2)) ++st.fp32_bad;
    }
    const int hit = lt::cell_hit(c, px, py, dr);
    ++st.trials;
    st.culled += far;
    st.dok += f.d_ok;
    st.hits += hit != 0;
    if ((far || fast || fast32) && hit) {
        if (st.unsound < 5)
            std::printf("UNSOUND: corners (%.17g %.17g %.17g %.17g) g (%.17g %.17g | %.17g %.17g | %.17g %.17g | %.17g "
                        "%.17g) P (%.17g %.17g) dr %.17g\n",
                        c.xp, c.yp, c.xq, c.yq, c.g0x, c.g0y, c.g1x, c.g1y, c.g2x, c.g2y, c.g3x, c.g3y, px, py, dr);
        ++st.unsound;
    }
    if (lens_like_far) {
        ++st.far_pairs;
        st.far_culled += fast;
        // the pairs the kernel's FP32 pre-test is for: d_ok cells whose FP64 distance margin decides
        if (f.d_ok && lt::far_margin_rho(f, c, px, py, dr) > 0.) {
            ++st.far_rho;
            st.far_culled32 += fast32;
        }
    }
    if (!f.on || !std::isfinite(px) || !std::isfinite(py) || !std::isfinite(dr)) return;
    // the error bounds behind the proof, on the corners the reference computes
    const P2 v0 = lt::to_source(c.xp, c.yp, dr, c.g0x, c.g0y);
    const P2 v1 = lt::to_source(c.xp, c.yq, dr, c.g1x, c.g1y);
    const P2 v2 = lt::to_source(c.xq, c.yp, dr, c.g2x, c.g2y);
    const P2 v3 = lt::to_source(c.xq, c.yq, dr, c.g3x, c.g3y);
    const P2 P = {px, py};
    double m = amax(amax(amax(v0.x, v0.y), amax(v1.x, v1.y)), amax(amax(v2.x, v2.y), amax(v3.x, v3.y)));
    m = amax(m, amax(px, py));
    const q m2 = (q)m * m;
    const P2 tris[2][3] = {{v0, v1, v2}, {v3, v2, v1}};
    for (const auto& t : tris) {
        const P2 probes[4][3] = {{t[0], t[1], t[2]}, {t[0], t[1], P}, {t[1], t[2], P}, {t[2], t[0], P}};
        for (const auto& pr : probes) {
            const q err = (q)lt::tri_det(pr[0], pr[1], pr[2]) - qdet(pr[0], pr[1], pr[2]);
            const double rel = m2 > 0 ? (double)((err < 0 ? -err : err) / m2) : 0.;
            if (rel > st.det_worst) st.det_worst = rel;
            if (rel > 4.0e-15) ++st.det_bad;
        }
    }
    // the polynomial against the exact determinant of the computed corners; the bound uses the
    // coordinate bound of the proof: max(|p| + |dr| |g|, |P|)
    const double mc = amax(amax(amax(c.xp, c.yp), c.xq), c.yq) +
                      std::fabs(dr) * amax(amax(amax(amax(c.g0x, c.g0y), amax(c.g1x, c.g1y)), amax(c.g2x, c.g2y)),
                                           amax(c.g3x, c.g3y));
    const double mp = mc > amax(px, py) ? mc : amax(px, py);
    const q mp2 = (q)mp * mp;
    const double du = __builtin_fma(dr, __builtin_fma(dr, f.c2u, f.c1u), f.c0);
    const double dl = __builtin_fma(dr, __builtin_fma(dr, f.c2l, f.c1l), f.c0);
    const q eu = (q)du - qdet(v0, v1, v2), el = (q)dl - qdet(v3, v2, v1);
    for (q e : {eu, el}) {
        const double rel = mp2 > 0 ? (double)((e < 0 ? -e : e) / mp2) : 0.;
        if (rel > st.poly_worst) st.poly_worst = rel;
        if (rel > 3.2e-14) ++st.poly_bad;
    }
    if (f.d_ok) {
        const q du_x = qdet(v0, v1, v2), dl_x = qdet(v3, v2, v1);
        const q need = (q)3.2e-11 * mp2;
        if ((du_x < 0 ? -du_x : du_x) <= need || (dl_x < 0 ? -dl_x : dl_x) <= need) ++st.dok_bad;
    }
}

struct TileStats {
    long tiles = 0, tiles_on = 0, pairs = 0, skipped = 0, unsound = 0, dist_bad = 0, far_pairs = 0, far_skipped = 0;
};

// Check 5: an 8 x 8 tile of cells on a (9 x 9)-corner grid of spacing h in a lens-like deflection field
// (regime 0: smooth; 1: with a strong local perturbation; 2: some corners collapsed onto a point),
// 12 sources of one sample; the tile test against every cell's reference test and exact distance.
static void check_tile(TileStats& ts, std::mt19937_64& rng)
{
    std::uniform_real_distribution<double> U(0.0, 1.0);
    std::normal_distribution<double> N(0.0, 1.0);
    const int regime = (int)(U(rng) * 3);
    const double h = std::pow(10.0, -1.5 + 1.8 * U(rng));
    const double x0 = (U(rng) - 0.5) * 200.0, y0 = (U(rng) - 0.5) * 200.0;
    const double a = 0.9 * (2 * U(rng) - 1), b = 0.5 * (2 * U(rng) - 1), cc = 0.5 * (2 * U(rng) - 1),
                 e = 0.9 * (2 * U(rng) - 1);
    const double bx = 20 * N(rng), by = 20 * N(rng), kx = 2e-3 * N(rng), ky = 2e-3 * N(rng);
    double GX[9][9], GY[9][9], X[9], Y[9];
    for (int i = 0; i < 9; ++i) {
        X[i] = x0 + i * h;
        Y[i] = y0 + i * h;
    }
    const double vx = (U(rng) - 0.5) * 50, vy = (U(rng) - 0.5) * 50;
    for (int j = 0; j < 9; ++j)
        for (int i = 0; i < 9; ++i) {
            const double u = X[i] - x0, v = Y[j] - y0;
            GX[j][i] = bx + a * u + b * v + kx * u * u + 0.2 * x0;
            GY[j][i] = by + cc * u + e * v + ky * v * v + 0.2 * y0;
            if (regime == 1) {
                const double r2 = (u - 4 * h) * (u - 4 * h) + (v - 4 * h) * (v - 4 * h) + 0.3 * h * h;
                GX[j][i] += 3 * h * h * (u - 4 * h) / r2;
                GY[j][i] += 3 * h * h * (v - 4 * h) / r2;
            }
            if (regime == 2 && U(rng) < 0.1) {  // this corner maps onto (vx, vy) at dr = 1
                GX[j][i] = X[i] - vx;
                GY[j][i] = Y[j] - vy;
            }
        }
    // sources of the sample: near the mapped tile, or anywhere
    const int ns = 12;
    double sx[ns], sy[ns], sd[ns];
    double pmax = 0, lo = 0, hi = 0;
    for (int k = 0; k < ns; ++k) {
        sd[k] = 0.2 + 1.2 * U(rng);
        const int ci = (int)(U(rng) * 9), cj = (int)(U(rng) * 9);
        const double mx = X[ci] - sd[k] * GX[cj][ci], my = Y[cj] - sd[k] * GY[cj][ci];
        const double r = U(rng);
        const double span = 9 * h * (1 + std::fabs(a) + std::fabs(e));
        if (r < 0.3) {
            sx[k] = mx + span * (U(rng) - 0.5) * 0.5;
            sy[k] = my + span * (U(rng) - 0.5) * 0.5;
        } else if (r < 0.7) {
            sx[k] = mx + span * (U(rng) - 0.5) * 6;
            sy[k] = my + span * (U(rng) - 0.5) * 6;
        } else {
            sx[k] = (U(rng) - 0.5) * 200;
            sy[k] = (U(rng) - 0.5) * 200;
        }
        pmax = std::fmax(pmax, std::fmax(std::fabs(sx[k]), std::fabs(sy[k])));
        lo = k ? std::fmin(lo, sd[k]) : sd[k];
        hi = k ? std::fmax(hi, sd[k]) : sd[k];
    }
    CellCorners cs[64];
    CellFilter fs[64];
    bool all_ok = true;
    for (int l = 0; l < 64; ++l) {
        const int i = l & 7, j = l >> 3;
        CellCorners& c = cs[l];
        c.xp = X[i];
        c.yp = Y[j];
        c.xq = X[i + 1];
        c.yq = Y[j + 1];
        c.g0x = GX[j][i];
        c.g0y = GY[j][i];
        c.g1x = GX[j + 1][i];
        c.g1y = GY[j + 1][i];
        c.g2x = GX[j][i + 1];
        c.g2y = GY[j][i + 1];
        c.g3x = GX[j + 1][i + 1];
        c.g3y = GY[j + 1][i + 1];
        fs[l] = lt::cell_filter(c, pmax, lo, hi);
        all_ok = all_ok && fs[l].d_ok;
    }
    ++ts.tiles;
    // reference cell: a random one (the kernel takes the first active lane)
    const int r = (int)(U(rng) * 64);
    const double drmax = std::fmax(std::fabs(lo), std::fabs(hi));
    double T = 0;
    for (int l = 0; l < 64; ++l)
        T = std::fmax(T, lt::tile_reach(fs[l], cs[l], cs[r].xp, cs[r].yp, cs[r].g0x, cs[r].g0y, drmax));
    const lt::CellFilter32 tf = lt::tile_filter32(cs[r].xp, cs[r].yp, cs[r].g0x, cs[r].g0y, T, fs[r].m, all_ok);
    ts.tiles_on += all_ok;
    for (int k = 0; k < ns; ++k) {
        ++ts.pairs;
        const bool skip = lt::far_margin_rho32(tf, (float)sx[k], (float)sy[k], (float)sd[k]) > 0.f;
        bool any_hit = false, any_near = false;
        for (int l = 0; l < 64; ++l) {
            const CellCorners& c = cs[l];
            if (lt::cell_hit(c, sx[k], sy[k], sd[k])) any_hit = true;
            if (!(fs[l].d_ok && lt::far_margin_rho(fs[l], c, sx[k], sy[k], sd[k]) > 0.)) any_near = true;
            if (skip) {
                const q ex = (q)sx[k] - ((q)c.xp - (q)sd[k] * c.g0x), ey = (q)sy[k] - ((q)c.yp - (q)sd[k] * c.g0y);
                if (!(ex * ex + ey * ey > (q)fs[l].rn2)) ++ts.dist_bad;
            }
        }
        ts.skipped += skip;
        if (skip && any_hit) {
            if (ts.unsound < 5) std::printf("UNSOUND TILE: regime %d h %.3g source (%.17g %.17g) dr %.17g\n", regime, h, sx[k], sy[k], sd[k]);
            ++ts.unsound;
        }
        // a source that every cell's own FP64 margin rejects by a wide distance: the tile test should too
        if (all_ok && regime == 0 && !any_near) {
            double mind = 1e300;
            for (int l = 0; l < 64; ++l)
                mind = std::fmin(mind, std::hypot(sx[k] - (cs[l].xp - sd[k] * cs[l].g0x), sy[k] - (cs[l].yp - sd[k] * cs[l].g0y)));
            if (mind > 2 * T) {
                ++ts.far_pairs;
                ts.far_skipped += skip;
            }
        }
    }
}

// pmax and the D_ls/D_s range of a one-source sample, as the kernel's per-sample bound forms them
static void bounds(double px, double py, double dr, double* pmax, double* lo, double* hi)
{
    const bool fin = std::fabs(px) <= 1e100 && std::fabs(py) <= 1e100 && std::fabs(dr) <= 1e100;
    const double nan = std::numeric_limits<double>::quiet_NaN();
    *pmax = fin ? std::fmax(std::fabs(px), std::fabs(py)) : nan;
    *lo = fin ? dr : nan;
    *hi = fin ? dr : nan;
}

int main(int argc, char** argv)
{
    const long n = argc > 1 ? std::atol(argv[1]) : 2000000;
    std::mt19937_64 rng(20240601);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    std::normal_distribution<double> N(0.0, 1.0);
    Stats st;
    const double nan = std::numeric_limits<double>::quiet_NaN(), inf = std::numeric_limits<double>::infinity();
    for (long t = 0; t < n; ++t) {
        CellCorners c;
        const int regime = (int)(t % 6);
        const double h = std::pow(10.0, -3.0 + 4.0 * U(rng));  // cell size 1e-3 .. 10
        const double x0 = (U(rng) - 0.5) * 200.0, y0 = (U(rng) - 0.5) * 200.0;
        c.xp = x0;
        c.yp = y0;
        c.xq = c.xp + h;
        c.yq = c.yp + h * (0.5 + U(rng));
        double dr = 0.05 + 1.5 * U(rng);
        const double px_[4] = {c.xp, c.xp, c.xq, c.xq}, py_[4] = {c.yp, c.yq, c.yp, c.yq};
        double gxv[4], gyv[4];
        bool lens_far = false;
        if (regime == 0 || regime == 1) {
            // smooth lens-like field: affine + curvature, deflections up to ~50"
            const double a = 2 * U(rng) - 1, b = 2 * U(rng) - 1, cc = 2 * U(rng) - 1, e = 2 * U(rng) - 1;
            const double kx = 1e-3 * N(rng), ky = 1e-3 * N(rng);
            const double bx = 20 * N(rng), by = 20 * N(rng);
            for (int i = 0; i < 4; ++i) {
                const double X = px_[i] - x0, Y = py_[i] - y0;
                gxv[i] = bx + a * X + b * Y + kx * X * X + 0.2 * x0;
                gyv[i] = by + cc * X + e * Y + ky * Y * Y + 0.2 * y0;
            }
        } else if (regime == 2) {
            // mapped corners chosen first: collinear (or nearly) triangles, then g = (p - v) / dr
            const double vx = (U(rng) - 0.5) * 100, vy = (U(rng) - 0.5) * 100;
            const double ux = N(rng), uy = N(rng);
            const double eps = U(rng) < 0.2 ? 0.0 : std::pow(10.0, -17.0 + 10.0 * U(rng));
            for (int i = 0; i < 4; ++i) {
                const double s = h * (U(rng) - 0.5) * 3;
                const double wx = vx + s * ux + eps * N(rng), wy = vy + s * uy + eps * N(rng);
                gxv[i] = (px_[i] - wx) / dr;
                gyv[i] = (py_[i] - wy) / dr;
            }
        } else if (regime == 3) {
            // collapsed cell: every corner maps to (nearly) the same point
            const double vx = (U(rng) - 0.5) * 50, vy = (U(rng) - 0.5) * 50;
            for (int i = 0; i < 4; ++i) {
                gxv[i] = (px_[i] - vx) / dr;
                gyv[i] = (py_[i] - vy) / dr;
            }
        } else if (regime == 4) {
            // wild gradients, magnitudes up to 1e8
            const double sc = std::pow(10.0, 8.0 * U(rng));
            for (int i = 0; i < 4; ++i) {
                gxv[i] = sc * N(rng);
                gyv[i] = sc * N(rng);
            }
        } else {
            // lens-like with a non-finite value somewhere (gradient, source or dr)
            for (int i = 0; i < 4; ++i) {
                gxv[i] = 10 * N(rng) + 0.3 * (px_[i] - x0);
                gyv[i] = 10 * N(rng) + 0.3 * (py_[i] - y0);
            }
            const int where = (int)(U(rng) * 8);
            const double bad = U(rng) < 0.5 ? nan : (U(rng) < 0.5 ? inf : -inf);
            if (where < 4) gxv[where] = bad;
            else if (where < 6) gyv[where - 4] = bad;
            else if (where == 6) dr = bad;
        }
        c.g0x = gxv[0];
        c.g0y = gyv[0];
        c.g1x = gxv[1];
        c.g1y = gyv[1];
        c.g2x = gxv[2];
        c.g2y = gyv[2];
        c.g3x = gxv[3];
        c.g3y = gyv[3];
        // the source: on a mapped corner, on an edge midpoint, near the cell, or anywhere
        const P2 v0 = lt::to_source(c.xp, c.yp, dr, c.g0x, c.g0y);
        const P2 v1 = lt::to_source(c.xp, c.yq, dr, c.g1x, c.g1y);
        const P2 v2 = lt::to_source(c.xq, c.yp, dr, c.g2x, c.g2y);
        const P2 v3 = lt::to_source(c.xq, c.yq, dr, c.g3x, c.g3y);
        const P2 vs[4] = {v0, v1, v2, v3};
        double px, py;
        const double r = U(rng);
        if (r < 0.15) {
            const P2 v = vs[(int)(U(rng) * 4)];
            px = v.x;
            py = v.y;
        } else if (r < 0.3) {
            const P2 a = vs[(int)(U(rng) * 4)], b = vs[(int)(U(rng) * 4)];
            px = 0.5 * (a.x + b.x);
            py = 0.5 * (a.y + b.y);
        } else if (r < 0.7) {
            const double span = std::fabs(v3.x - v0.x) + std::fabs(v3.y - v0.y) + std::fabs(v1.x - v2.x) + 1e-300;
            px = v0.x + span * 3 * (U(rng) - 0.3);
            py = v0.y + span * 3 * (U(rng) - 0.3);
        } else {
            px = (U(rng) - 0.5) * 200;
            py = (U(rng) - 0.5) * 200;
            lens_far = regime <= 1 && h >= 0.05 && std::hypot(px - v0.x, py - v0.y) > 20 * (h + dr * 2 * h);
        }
        if (regime == 5 && (int)(U(rng) * 8) == 7) px = U(rng) < 0.5 ? nan : inf;
        double pmax, lo, hi;
        bounds(px, py, dr, &pmax, &lo, &hi);
        // sometimes looser bounds, as other sources of the sample would give
        if (U(rng) < 0.3) {
            pmax *= 1 + 10 * U(rng);
            lo -= U(rng) * std::fabs(lo);
            hi += U(rng) * std::fabs(hi);
        }
        check(st, c, px, py, dr, pmax, lo, hi, lens_far);
    }
    TileStats ts;
    for (long t = 0; t < n / 64; ++t) check_tile(ts, rng);
    // sqrt_up() bounds the square root from above, over magnitudes 1e-300 .. 1e300 (and 0)
    long sqrt_bad = 0;
    for (long t = 0; t < n / 8; ++t) {
        const double qv = t == 0 ? 0.0 : std::pow(10.0, -300.0 + 600.0 * std::uniform_real_distribution<double>(0, 1)(rng));
        if (!(q__sqrt(qv) <= (q)lt::sqrt_up(qv))) ++sqrt_bad;
    }
    std::printf("trials %ld culled %ld hits %ld unsound %ld\n", st.trials, st.culled, st.hits, st.unsound);
    std::printf("tri_det error worst %.3g M^2 (bound 4.0e-15, over: %ld); d(dr) error worst %.3g M^2 (bound 3.2e-14, over: %ld)\n",
                st.det_worst, st.det_bad, st.poly_worst, st.poly_bad);
    std::printf("cells with d_ok %ld (violations %ld)\n", st.dok, st.dok_bad);
    std::printf("lens-like far pairs %ld culled %ld (%.4f); of the %ld with d_ok decided by the FP64 distance "
                "margin, the FP32 pre-test decides %ld (%.4f)\n", st.far_pairs, st.far_culled,
                st.far_pairs ? (double)st.far_culled / st.far_pairs : 0.0, st.far_rho, st.far_culled32,
                st.far_rho ? (double)st.far_culled32 / st.far_rho : 0.0);
    std::printf("FP32 pre-test skips %ld (exact distance not above sqrt(rn2): %ld)\n", st.fast32, st.fp32_bad);
    std::printf("tiles %ld (all cells d_ok: %ld), (tile, source) pairs %ld, skipped %ld, unsound %ld, exact distance "
                "not above sqrt(rn2): %ld; far pairs %ld skipped %ld (%.4f)\n", ts.tiles, ts.tiles_on, ts.pairs, ts.skipped,
                ts.unsound, ts.dist_bad, ts.far_pairs, ts.far_skipped, ts.far_pairs ? (double)ts.far_skipped / ts.far_pairs : 0.0);
    std::printf("sqrt_up below the square root: %ld of %ld\n", sqrt_bad, n / 8);
    const bool ok = st.unsound == 0 && st.det_bad == 0 && st.poly_bad == 0 && st.dok_bad == 0 && st.dok > 0 &&
                    sqrt_bad == 0 &&
                    ts.unsound == 0 && ts.dist_bad == 0 && ts.skipped > 0 && ts.far_pairs > 0 &&
                    ts.far_skipped == ts.far_pairs &&
                    st.hits > 0 && st.fp32_bad == 0 && st.fast32 > 0 &&
                    st.far_pairs > 0 && st.far_culled >= 0.999 * st.far_pairs &&
                    st.far_rho > 0 && st.far_culled32 >= 0.99 * st.far_rho;
    std::printf("%s\n", ok ? "OK" : "FAIL");
    return ok ? 0 : 1;
}
