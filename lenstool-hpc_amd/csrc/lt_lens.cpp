// lt_lens.cpp -- packs a caller's SOA lens model (the reference's Potential_SOA layout) into the
// device halo records and runs. All per-halo constants of the closed forms are formed here, once
// per model, in double precision.
#include <cmath>
#include <cstring>
#include <cstdio>

#include "lenstool_hip/lt_status.h"
#include "lt_kernels.h"

namespace lt {

static bool finite(double v) { return std::isfinite(v); }

// Radius (in units of the halo's b0, times rcut/(rcut - rcore) for the dPIE) inside which the
// FP32-compensated Hessian takes a halo's term in FP64. Measured FP32 rounding of one term is about
// 1e-6 b0 t05 / r at halo-frame radius r, and 1.2e-7 b0 / r for the SIS: these radii keep every FP32
// term under 5e-8 (PIEMD) and 1.2e-8 (SIS).
constexpr double kFp32NearScalePiemd = 20.0;
constexpr double kFp32NearScaleSis = 10.0;

static int bad(std::string& err, int i, const char* what)
{
    char buf[160];
    std::snprintf(buf, sizeof buf, "halo %d: %s", i, what);
    err = buf;
    return LT_ERR_INVALID;
}

int pack_model(const lt_potential_soa* L, int n, Use use, std::vector<HaloRec>& recs,
               std::vector<HaloRec32>& recs32, std::vector<Run>& runs, std::string& err)
{
    recs.clear();
    recs32.clear();
    runs.clear();
    if (n < 0) {
        err = "negative halo count";
        return LT_ERR_INVALID;
    }
    if (n == 0) return LT_OK;
    if (!L || !L->type || !L->position_x || !L->position_y || !L->b0 || !L->ellipticity_potential ||
        !L->anglecos || !L->anglesin) {
        err = "lens model: a required array (type, position_x/y, b0, ellipticity_potential, anglecos/sin) is NULL";
        return LT_ERR_INVALID;
    }
    if (use == Use::Kmap && !L->dlsds) {
        err = "lens model: dlsds is NULL (needed by the Kmap / mass map)";
        return LT_ERR_INVALID;
    }
    recs.resize(n);
    recs32.resize(n);
    for (int i = 0; i < n; ++i) {
        const int t = L->type[i];
        const bool piemd = (t == TYPE_PIEMD || t == TYPE_DPIE);
        bool ok = (t == TYPE_SIS || piemd) || (t == TYPE_SHEAR && (use == Use::Hessian || use == Use::Kmap));
        if (!ok) {
            char buf[160];
            std::snprintf(buf, sizeof buf, "halo %d: type %d has no %s implementation", i, t,
                          use == Use::Gradient ? "gradient" : use == Use::Potential ? "potential" : "Hessian");
            err = buf;
            return LT_ERR_TYPE;
        }
        HaloRec h{};
        const double x = L->position_x[i], y = L->position_y[i], b0 = L->b0[i];
        const double eps = L->ellipticity_potential[i];
        const double cs = L->anglecos[i], sn = L->anglesin[i];
        if (!finite(x) || !finite(y) || !finite(b0) || !finite(eps) || !finite(cs) || !finite(sn))
            return bad(err, i, "non-finite parameter");
        h.xh = x;
        h.yh = y;
        h.cs = cs;
        h.sn = sn;
        const double cos2 = cs * cs - sn * sn, sin2 = 2.0 * cs * sn;
        if (t == TYPE_SIS) {
            if (!(std::fabs(eps) < 1.0)) return bad(err, i, "SIS ellipticity_potential must be in (-1, 1)");
            h.sa = 1.0 - eps;
            h.sb = 1.0 + eps;
            h.ampg = b0;
            h.amph = b0 * h.sa * h.sb;
            h.circ = (eps == 0.0) ? 1.0 : 0.0;
        } else if (piemd) {
            if (!L->rcore || (t == TYPE_DPIE && !L->rcut)) {
                err = "lens model: rcore / rcut is NULL";
                return LT_ERR_INVALID;
            }
            const double rc = L->rcore[i];
            const double rcut = (t == TYPE_DPIE) ? L->rcut[i] : 0.0;
            if (!(eps > 0.0 && eps < 1.0)) return bad(err, i, "PIEMD ellipticity_potential must be in (0, 1)");
            if (!(rc >= 0.0) || !finite(rc)) return bad(err, i, "PIEMD rcore must be >= 0");
            if (t == TYPE_DPIE && (!finite(rcut) || !(rcut > 0.0) || rcut == rc))
                return bad(err, i, "dPIE rcut must be > 0 and differ from rcore");
            const double sqe = std::sqrt(eps);
            const double cx1 = (1.0 - eps) / (1.0 + eps);
            const double ci = 0.5 * (1.0 - eps * eps) / sqe;
            h.rc = rc;
            h.rcut = rcut;
            h.kx = 1.0 / ((1.0 + eps) * (1.0 + eps));
            h.ky = 1.0 / ((1.0 - eps) * (1.0 - eps));
            h.e2 = 2.0 * sqe;
            h.cx1 = cx1;
            h.gki = 2.0 * sqe / ((1.0 + eps) * (1.0 + eps));
            h.gkv = 2.0 * sqe / ((1.0 + eps) * (1.0 - eps));
            h.ra2 = rc * rc;
            h.rb2 = rcut * rcut;
            h.rca = cx1 * rc;
            h.rcb = cx1 * rcut;
            const double t05 = (t == TYPE_DPIE) ? rcut / (rcut - rc) : 1.0;
            h.ampg = b0 * t05 * ci;
            h.amph = b0 * t05 * ci;
            // Hessian in scaled halo-frame coordinates X = cx1 a x, Y = cx1 b y (lt_terms64.h)
            {
                const double a = 1.0 / (1.0 + eps), b = 1.0 / (1.0 - eps);
                h.csx = cx1 * a * cs;
                h.snx = cx1 * a * sn;
                h.csy = cx1 * b * cs;
                h.sny = cx1 * b * sn;
                h.k1 = h.e2 * a / cx1;
                h.k2 = b / a;
                h.k3 = h.e2 * b / cx1;
                h.k4 = 1.0 / (cx1 * cx1);
                h.k5 = h.e2 * a;
                h.ra2c = cx1 * cx1 * rc * rc;
                h.rb2c = cx1 * cx1 * rcut * rcut;
            }
            // potential constants (lt_terms64.h pot_piemd): c = 1 + cx1 = 2 / (1 + e)
            const double c = 2.0 / (1.0 + eps);
            h.psra = sqe * rc;
            h.psrb = sqe * rcut;
            h.pcra = c * rc;
            h.pcrb = c * rcut;
            h.picra = 1.0 / (c * rc);
            h.picra2 = h.picra * h.picra;
            h.picrb2 = (t == TYPE_DPIE) ? 1.0 / ((c * rcut) * (c * rcut)) : 0.0;
            h.pc = c;
            h.pomc = 1.0 - cx1;
            h.pdr = rcut - rc;
            h.pcdr = c * (rcut - rc);
            h.pomcdr = (1.0 - cx1) * (rcut - rc);
        } else {  // external shear: 3 eps (cos 2t, sin 2t; sin 2t, -cos 2t)
            h.h14a = 3.0 * eps * cos2;
            h.h14b = 3.0 * eps * sin2;
            h.h14c = -3.0 * eps * cos2;
        }
        h.hc2 = h.amph * cos2;
        h.hs2 = h.amph * sin2;
        if (piemd) {
            const double a = 1.0 / (1.0 + eps);
            const double e2a2 = h.e2 * a * a;
            h.qT = h.amph * e2a2;
            h.qc2 = h.hc2 * e2a2;
            h.qs2m2 = -2.0 * h.hs2 * e2a2;
            h.qs2h = 0.5 * h.hs2 * e2a2;
        }
        recs[i] = h;

        HaloRec32& f = recs32[i];
        f = HaloRec32{};
        if (piemd) {
            const double a = 1.0 / (1.0 + eps), e2 = h.e2;
            // The FP32 terms are the I_{0.5} Hessian times 1 / (2 sqrt(e) a^2): fold that factor back in.
            const double amp = h.amph * e2 * a * a;
            f.csx = (float)h.csx;
            f.snx = (float)h.snx;
            f.csy = (float)h.csy;
            f.sny = (float)h.sny;
            f.k1 = (float)h.k1;
            f.k2 = (float)h.k2;
            f.k3 = (float)h.k3;
            f.k4 = (float)h.k4;
            f.k5 = (float)h.k5;
            f.ra = (float)h.rc;
            f.ra2c = (float)h.ra2c;
            f.rb = (float)h.rcut;
            f.rb2c = (float)h.rb2c;
            f.hT = (float)amp;
            f.hc2 = (float)(amp * cos2);
            f.hs2m2 = (float)(-2.0 * amp * sin2);
            f.hs2h = (float)(0.5 * amp * sin2);
            // FP32 rounding of one term is about 1e-6 b0 t05 / r (r = halo-frame radius) away from the
            // core; inside R = 20 b0 t05 it would exceed 5e-8, so the FP64 term is used there.
            const double t05 = (t == TYPE_DPIE) ? h.rcut / (h.rcut - h.rc) : 1.0;
            const double R = kFp32NearScalePiemd * std::fabs(b0 * t05);
            f.near2 = (float)std::fmin(R * R, 3.0e38);
            // X^2 + Y^2 = cx1^2 (a^2 x^2 + b^2 y^2) <= (cx1 b)^2 d^2 = a^2 d^2 (b > a): the region
            // X^2 + Y^2 < a^2 R^2 holds the disk d < R and lies inside d < R / cx1
            f.near2s = (float)std::fmin(a * a * R * R, 3.0e38);
            f.nearr = (float)std::fmin(R * (1.0 + eps) / (1.0 - eps), 1.0e19);
        } else if (t == TYPE_SIS) {
            f.sa = (float)h.sa;
            f.sb = (float)h.sb;
            f.hT = (float)h.amph;
            f.csx = f.csy = (float)cs;
            f.snx = f.sny = (float)sn;
            f.hc2 = (float)h.hc2;
            f.hs2m2 = (float)(-2.0 * h.hs2);
            f.hs2h = (float)(0.5 * h.hs2);
            f.circ = (float)h.circ;
            const double R = kFp32NearScaleSis * std::fabs(b0);
            f.near2 = (float)(R * R);
            f.near2s = f.near2;  // plain rotation (and the lab frame for a circular SIS): X^2 + Y^2 = d^2
            f.nearr = (float)R;
        } else {
            f.near2 = 3.0e38f;  // type 14 stays FP64 (a constant per halo)
            f.near2s = 3.0e38f;
            f.nearr = 1.0e19f;
        }
        if (use == Use::Gradient) {
            // gradient models carry the gradient's FP32 constants instead (same near radius)
            const float near2 = f.near2;
            GradRec32 g{};
            g.cs = (float)cs;
            g.sn = (float)sn;
            g.near2 = near2;
            if (piemd) {
                const double a = 1.0 / (1.0 + eps), b = 1.0 / (1.0 - eps), sqe = std::sqrt(eps);
                g.kx = (float)h.kx;
                g.ky = (float)h.ky;
                g.cx1 = (float)h.cx1;
                g.ki = (float)(2.0 * sqe * a * a);
                g.kv = (float)(2.0 * sqe * a * b);
                g.ra = (float)h.rc;
                g.rb = (float)h.rcut;
                g.ra2 = (float)h.ra2;
                g.rb2 = (float)h.rb2;
                g.rca = (float)(h.cx1 * h.rc);
                g.rcb = (float)(h.cx1 * h.rcut);
                g.ampg = (float)h.ampg;
            } else if (t == TYPE_SIS) {
                g.sa = (float)h.sa;
                g.sb = (float)h.sb;
                g.ampg = (float)h.ampg;
                g.circ = (float)h.circ;
            }
            std::memcpy(&f, &g, sizeof(g));
        } else if (use == Use::Potential) {
            // potential models carry the potential's FP32 constants
            PotRec32 q{};
            q.cs = (float)cs;
            q.sn = (float)sn;
            if (piemd) {
                q.kx = (float)h.kx;
                q.ky = (float)h.ky;
                q.cx1 = (float)h.cx1;
                q.gki = (float)h.gki;
                q.gkv = (float)h.gkv;
                q.ra = (float)h.rc;
                q.ra2 = (float)h.ra2;
                q.cra = (float)h.pcra;
                q.icra2 = (float)h.picra2;
                q.icra = (float)h.picra;
                q.sra = (float)h.psra;
                q.rb = (float)h.rcut;
                q.rb2 = (float)h.rb2;
                q.crb = (float)h.pcrb;
                q.icrb2 = (float)h.picrb2;
                q.srb = (float)h.psrb;
                q.pc = (float)h.pc;
                q.pomc = (float)h.pomc;
                q.pdr = (float)h.pdr;
                q.pcdr = (float)h.pcdr;
                q.pomcdr = (float)h.pomcdr;
                q.amp = (float)h.ampg;
            } else if (t == TYPE_SIS) {
                q.sa = (float)h.sa;
                q.sb = (float)h.sb;
                q.amp = (float)h.ampg;
            }
            std::memcpy(&f, &q, sizeof(q));
        }
    }
    // runs: equal type (and, for the Kmap, equal dlsds); the scan never reads past n - 1
    int s = 0;
    while (s < n) {
        const int t = L->type[s];
        const double w = (use == Use::Kmap) ? L->dlsds[s] : 1.0;
        if (!finite(w)) return bad(err, s, "non-finite dlsds");
        int c = 1;
        while (s + c < n && L->type[s + c] == t && (use != Use::Kmap || L->dlsds[s + c] == w)) ++c;
        Run r{};
        r.type = t;
        r.start = s;
        r.count = c;
        r.weight = w;
        r.circ = (t == TYPE_SIS) ? 1 : 0;
        for (int i = s; i < s + c && r.circ; ++i) r.circ = recs[i].circ != 0.0;
        runs.push_back(r);
        s += c;
    }
    return LT_OK;
}

}  // namespace lt
