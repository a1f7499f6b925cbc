// lt_par.cpp -- reader of Lenstool parameter files (.par) and the catalogues they name (potfiles,
// multiple-image files), producing the lens model, frame, cosmology and run settings of the map
// path. Follows the reference's readers section by section:
//   cosmology            src/module_readParameters.cpp:278-384
//   runmode / sections   src/module_readParameters.cpp:385-700 (read_runmode, read_runmode_*)
//   potentials           src/module_readParameters.cpp:2234-2489 (..._PotentialSOA_direct)
//   b0 / epot            src/module_readParameters.cpp:2808-2944
//   potfiles             src/module_readParameters.cpp:1102-1297, 1398-1656
//   frame                src/module_readParameters.cpp:2954-3030
//   images               src/module_readParameters.cpp:852-935
//   dlsds                src/module_readParameters.cpp:230-257
//   WCS round trip       src/convert_coordinates.cpp
// Reference behaviour kept on purpose: potentials are slotted by ascending type (file order within
// a type) and potfile clumps follow them in catalogue order; a flat universe (omegaK = 0) takes
// omegaX = 1 - omegaM whatever `lambda` says; a PIEMD with zero ellipticity gets epot = 1e-5; potfile
// positions go through the relative -> absolute -> relative coordinate round trip; `profil` inside a
// potfile block is ignored (its `type` line sets the profile). Unlike the reference, a missing
// catalogue or redshift is an error status instead of exit(-1).
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "lenstool_hip/lt_capi.h"
#include "lt_internal.h"
#include "lt_kernels.h"

using namespace lt;

namespace {

struct Halo {
    int type = 0;
    double x = 0, y = 0, b0 = 0, vdisp = 0, angle = 0, ell = 0, epot = 0, rcore = 0, rcut = 0, z = 0;
    double mag = 0, lum = 0, theta = 0;
};

struct Potfile {
    int ftype = 0, type = 0, ircut = 0, islope = 0, isigma = 0, ivdslope = 0;
    std::string file;
    double zlens = 0, mag0 = 0, sigma1 = 0, core = -1.0, corekpc = -1.0;
    double cut1 = DBL_MAX, cutkpc1 = DBL_MAX, slope1 = 0, vdslope1 = 0;
};

std::string tok(const std::string& line, int k)
{
    std::istringstream s(line);
    std::string w;
    for (int i = 0; i <= k; ++i)
        if (!(s >> w)) return "";
    return w;
}

bool starts(const std::string& a, const char* b) { return a.compare(0, std::strlen(b), b) == 0; }

bool exists(const std::string& p)
{
    std::ifstream f(p.c_str());
    return (bool)f;
}

// a catalogue path as written, else relative to the parameter file's directory, else its base name there
std::string resolve(const std::string& name, const std::string& dir)
{
    if (exists(name)) return name;
    if (!name.empty() && name[0] != '/' && exists(dir + "/" + name)) return dir + "/" + name;
    const size_t k = name.find_last_of('/');
    const std::string base = k == std::string::npos ? name : name.substr(k + 1);
    if (exists(dir + "/" + base)) return dir + "/" + base;
    return "";
}

// b0 and potential ellipticity of a potential block (src/module_readParameters.cpp:2808-2871)
void potential_parameters(Halo& h)
{
    switch (h.type) {
    case 5:
        h.b0 = 4. * LT_PI_C2 * h.vdisp * h.vdisp;
        h.epot = h.ell / 3.;
        break;
    case 8:
    case 81:
        h.b0 = 6. * LT_PI_C2 * h.vdisp * h.vdisp;
        if (h.ell == 0. && h.epot != 0.) h.ell = 2. * h.epot / (1. + h.epot * h.epot);
        else if (h.ell == 0. && h.epot == 0.) h.epot = 0.00001;
        else h.epot = (1. - std::sqrt(1 - h.ell * h.ell)) / h.ell;
        break;
    case 14:
        h.epot = h.ell / 3.;
        break;
    default:
        break;
    }
}

// the same for a potfile clump (src/module_readParameters.cpp:2873-2944: the SIS branch differs)
void clump_parameters(Halo& h)
{
    switch (h.type) {
    case 5:
        h.b0 = 4. * LT_PI_C2 * h.vdisp * h.vdisp;
        if (h.ell == 0. && h.epot != 0.) h.epot = h.ell / 3.;
        else h.ell = h.epot * 3.;
        break;
    case 8:
    case 81:
        potential_parameters(h);
        break;
    default:
        if (h.ell == 0. && h.epot != 0.) h.ell = h.epot * 3.;
        else h.epot = h.ell / 3.;
        break;
    }
}

// relative arcsec -> absolute degrees -> relative arcsec (src/convert_coordinates.cpp)
void wcs_round_trip(double& x, double& y, int mode, double ra, double dec)
{
    const double dtr = std::acos(-1.) / 180.;
    if (mode == 1 || mode == 3) {
        x /= -3600. * std::cos(dec * dtr);
        x += ra;
        y /= 3600.;
        y += dec;
        x -= ra;
        x *= -3600 * std::cos(dec * dtr);
        y -= dec;
        y *= 3600;
    } else if (mode == 2) {
        x += ra;
        y += dec;
        x -= ra;
        y -= dec;
    }
}

}  // namespace

struct lt_par_s {
    lt_par_info info{};
    lt_grid_param frame{};
    lt_cosmo_param cosmo{};
    std::vector<int> type;
    std::vector<double> x, y, b0, vdisp, angle, ell, epot, rcore, rcut, z, mag, lum, theta, cs, sn, dlsds;
    std::vector<int> nimages;
    std::vector<lt_galaxy> images;
    std::string imagefile;

    void push(const Halo& h)
    {
        type.push_back(h.type);
        x.push_back(h.x);
        y.push_back(h.y);
        b0.push_back(h.b0);
        vdisp.push_back(h.vdisp);
        angle.push_back(h.angle);
        ell.push_back(h.ell);
        epot.push_back(h.epot);
        rcore.push_back(h.rcore);
        rcut.push_back(h.rcut);
        z.push_back(h.z);
        mag.push_back(h.mag);
        lum.push_back(h.lum);
        theta.push_back(h.theta);
        cs.push_back(std::cos(h.angle));
        sn.push_back(std::sin(h.angle));
    }
};

namespace {

// kpc per arcsec at redshift z: d0 / h * D_A(z)
double kpc_per_arcsec(double z, const lt_cosmo_param& c) { return LT_D0 / c.h * cosmo_observer_object(z, c); }

int parse(lt_par_s& P, const std::string& path)
{
    std::ifstream in(path.c_str());
    if (!in) return set_error(LT_ERR_NOT_FOUND, "parameter file " + path + " not found");
    std::vector<std::string> L;
    for (std::string s; std::getline(in, s);) L.push_back(s);
    const size_t slash = path.find_last_of('/');
    const std::string dir = slash == std::string::npos ? "." : path.substr(0, slash);
    const double dtr = std::acos(-1.) / 180.;

    // defaults (src/module_readParameters.cpp:284-291, 700-741)
    lt_cosmo_param& C = P.cosmo;
    C.model = 1;
    C.H0 = 50;
    C.h = 1.;
    C.omegaM = 1.;
    C.omegaX = 0;
    C.wX = -1.;
    C.wa = 0.;
    C.curvature = 0.;
    lt_par_info& R = P.info;
    R.nbgridcells = 1000;
    R.ampli_cells = R.mass_cells = R.shear_cells = R.poten_cells = R.dpl_cells = 1000;
    R.mass_z = 0.4;
    R.poten_z = R.dpl_z = 0.8;
    R.z_source = 0.8;

    std::vector<Halo> main;
    std::vector<Potfile> pots;
    // block end: a line whose first word starts with "end"
    auto block_end = [&](size_t i) { return starts(tok(L[i], 0), "end"); };
    // the cosmology first, wherever it stands in the file: kpc radii of every block depend on it
    for (size_t i = 0; i < L.size(); ++i) {
        const std::string first = tok(L[i], 0);
        if (starts(first, "fini")) break;
        if (starts(first, "cosmolog")) {
            for (++i; i < L.size() && !block_end(i); ++i) {
                const std::string k = tok(L[i], 0), v = tok(L[i], 1);
                if (k == "model") C.model = std::atoi(v.c_str());
                else if (k == "H0") {
                    C.H0 = std::atof(v.c_str());
                    C.h = C.H0 / 50.;
                } else if (k == "omegaM" || k == "omega") C.omegaM = std::atof(v.c_str());
                else if (k == "omegaX" || k == "lambda") C.omegaX = std::atof(v.c_str());
                else if (k == "wX" || k == "q" || k == "w0") C.wX = std::atof(v.c_str());
                else if (k == "wa" || k == "n" || k == "delta" || k == "w1") C.wa = std::atof(v.c_str());
                else if (k == "omegaK") C.curvature = std::atof(v.c_str());
            }
            if (C.curvature == 0.) C.omegaX = 1 - C.omegaM;
            else C.curvature = C.omegaM + C.omegaX - 1;
        }
    }
    for (size_t i = 0; i < L.size(); ++i) {
        const std::string first = tok(L[i], 0);
        if (starts(first, "fini")) break;
        if (first.empty()) continue;
        if (starts(first, "cosmolog")) {
            while (i + 1 < L.size() && !block_end(i + 1)) ++i;
        } else if (starts(first, "runmode")) {
            for (++i; i < L.size() && !block_end(i); ++i) {
                const std::string& s = L[i];
                const std::string k = tok(s, 0);
                char buf[4096];
                double v = 0;
                if (k == "nbgridcells") std::sscanf(s.c_str(), " %*s %d", &R.nbgridcells);
                else if (k == "inverse") std::sscanf(s.c_str(), " %*s %d", &R.inverse);
                else if (k == "ampli") {
                    std::sscanf(s.c_str(), " %*s %d %d %lf %4095s", &R.ampli, &R.ampli_cells, &v, buf);
                    R.ampli_z = v;
                } else if (k == "mass") {
                    std::sscanf(s.c_str(), " %*s %d %d %lf %4095s", &R.mass, &R.mass_cells, &v, buf);
                    R.mass_z = v;
                } else if (k == "shear") {
                    std::sscanf(s.c_str(), " %*s %d %d %lf %4095s", &R.shear, &R.shear_cells, &v, buf);
                    R.shear_z = v;
                } else if (k == "poten") {
                    std::sscanf(s.c_str(), " %*s %d %d %lf %4095s", &R.poten, &R.poten_cells, &v, buf);
                    R.poten_z = v;
                } else if (k == "dpl") {
                    std::sscanf(s.c_str(), " %*s %d %d %lf", &R.dpl, &R.dpl_cells, &v);
                    R.dpl_z = v;
                } else if (k == "reference") {
                    double a = 0, b = 0;
                    std::sscanf(s.c_str(), " %*s %d %lf %lf", &R.ref_mode, &a, &b);
                    R.ref_ra = a;
                    R.ref_dec = b;
                } else if (k == "image") {
                    int on = 0;
                    if (std::sscanf(s.c_str(), " %*s %d %4095s", &on, buf) == 2 && on) P.imagefile = buf;
                }
            }
        } else if (starts(first, "source")) {
            for (++i; i < L.size() && !block_end(i); ++i)
                if (tok(L[i], 0) == "z_source") R.z_source = std::atof(tok(L[i], 1).c_str());
        } else if (starts(first, "image")) {
            for (++i; i < L.size() && !block_end(i); ++i) {
                if (tok(L[i], 0) == "multfile") {
                    int on = 0;
                    char buf[4096];
                    if (std::sscanf(L[i].c_str(), " %*s %d %4095s", &on, buf) == 2 && on >= 1) P.imagefile = buf;
                }
            }
        } else if (starts(first, "grille") || starts(first, "frame") || starts(first, "champ")) {
            lt_grid_param& F = P.frame;
            for (++i; i < L.size() && !block_end(i); ++i) {
                const std::string k = tok(L[i], 0);
                const double v = std::atof(tok(L[i], 1).c_str());
                if (k == "dmax") {
                    F.xmin = F.ymin = -v;
                    F.xmax = F.ymax = v;
                    F.rmax = v * std::sqrt(2.);
                } else if (k == "xmin") F.xmin = v;
                else if (k == "xmax") F.xmax = v;
                else if (k == "ymin") F.ymin = v;
                else if (k == "ymax") F.ymax = v;
            }
        } else if (starts(first, "potfile")) {
            Potfile pf;
            for (++i; i < L.size() && !block_end(i); ++i) {
                const std::string& s = L[i];
                const std::string k = tok(s, 0);
                double a = 0, b = 0;
                if (k == "filein") {
                    char buf[4096];
                    if (std::sscanf(s.c_str(), " %*s %d %4095s", &pf.ftype, buf) == 2) pf.file = buf;
                } else if (k == "type") pf.type = std::atoi(tok(s, 1).c_str());
                else if (k == "zlens" || k == "z_lens") pf.zlens = std::atof(tok(s, 1).c_str());
                else if (k == "mag0" || k == "r200") pf.mag0 = std::atof(tok(s, 1).c_str());
                else if (k == "core") pf.core = std::atof(tok(s, 1).c_str());
                else if (k == "corekpc") pf.corekpc = std::atof(tok(s, 1).c_str());
                else if (k == "cut") {
                    std::sscanf(s.c_str(), " %*s %d%lf%lf", &pf.ircut, &a, &b);
                    pf.cut1 = a;
                } else if (k == "cutkpc") {
                    std::sscanf(s.c_str(), " %*s %d%lf%lf", &pf.ircut, &a, &b);
                    pf.cutkpc1 = a;
                } else if (k == "slope" || k == "m200slope") {
                    std::sscanf(s.c_str(), " %*s %d%lf%lf", &pf.islope, &a, &b);
                    pf.slope1 = a;
                } else if (k == "sigma") {
                    std::sscanf(s.c_str(), " %*s %d%lf%lf", &pf.isigma, &a, &b);
                    pf.sigma1 = a;
                } else if (k == "vdslope" || k == "c200slope") {
                    std::sscanf(s.c_str(), " %*s %d%lf%lf", &pf.ivdslope, &a, &b);
                    pf.vdslope1 = a;
                }
            }
            pots.push_back(pf);
        } else if (starts(first, "potent")) {
            Halo h;
            if (!main.empty()) h.type = main.back().type;  // a block without `profil` keeps the last one
            double core_kpc = 0, cut_kpc = 0;
            for (++i; i < L.size() && !block_end(i); ++i) {
                const std::string k = tok(L[i], 0);
                const double v = std::atof(tok(L[i], 1).c_str());
                if (k == "profil" || k == "profile") h.type = std::atoi(tok(L[i], 1).c_str());
                else if (k == "x_centre" || k == "x_center") h.x = v;
                else if (k == "y_centre" || k == "y_center") h.y = v;
                else if (k == "ellipticitymass" || k == "ellipticity" || k == "ellipticite") h.ell = v;
                else if (k == "ellipticity_angle" || k == "angle_pos") h.angle = v * dtr;
                else if (k == "rcore" || k == "core_radius") h.rcore = v;
                else if (k == "rcut" || k == "cut_radius") h.rcut = v;
                else if (k == "core_radius_kpc") core_kpc = v;
                else if (k == "cut_radius_kpc") cut_kpc = v;
                else if (k == "z_lens") h.z = v;
                else if (k == "v_disp") h.vdisp = v;
            }
            if (h.z == 0.)
                return set_error(LT_ERR_INVALID, "potential " + std::to_string(main.size() + 1) + ": no z_lens");
            const double d1 = kpc_per_arcsec(h.z, C);
            if (core_kpc != 0.) h.rcore = core_kpc / d1;
            if (cut_kpc != 0.) h.rcut = cut_kpc / d1;
            potential_parameters(h);
            main.push_back(h);
        }
    }
    R.nhalos_main = (int)main.size();

    // slot the potentials by ascending type, file order within a type
    for (int t = 0; t < 100; ++t)
        for (const Halo& h : main)
            if (h.type == t) P.push(h);
    for (const Halo& h : main)
        if (h.type < 0 || h.type >= 100)
            return set_error(LT_ERR_TYPE, "potential profile " + std::to_string(h.type) + " out of range");

    // potfiles
    for (size_t k = 0; k < pots.size(); ++k) {
        Potfile& pf = pots[k];
        if (pf.ftype != 1 && pf.ftype != 3)
            return set_error(LT_ERR_INVALID, "potfile " + pf.file + ": only catalogue types 1 and 3 are read");
        const std::string fn = resolve(pf.file, dir);
        if (fn.empty()) return set_error(LT_ERR_NOT_FOUND, "potfile " + pf.file + " not found");
        std::ifstream cat(fn.c_str());
        const double d1 = kpc_per_arcsec(pf.zlens, C);
        const double cut = pf.cut1 == DBL_MAX && pf.cutkpc1 != DBL_MAX ? pf.cutkpc1 / d1 : pf.cut1;
        const double core = pf.core == -1.0 && pf.corekpc != -1 ? pf.corekpc / d1 : pf.core;
        int ref_mode = 0;
        double ref_ra = 0, ref_dec = 0;
        int n = 0;
        for (std::string s; std::getline(cat, s);) {
            if (starts(s, "#REFERENCE")) {
                std::sscanf(s.c_str(), " %*s %d%lf%lf", &ref_mode, &ref_ra, &ref_dec);
                continue;
            }
            if (s.empty() || s[0] == '#' || tok(s, 0).empty()) continue;
            double cx, cy, aa, bb, th, mg, lm;
            if (std::sscanf(s.c_str(), "%*s%lf%lf%lf%lf%lf%lf%lf", &cx, &cy, &aa, &bb, &th, &mg, &lm) != 7)
                return set_error(LT_ERR_INVALID, "potfile " + fn + ": a line has missing values: " + s);
            Halo h;
            h.type = pf.type;
            h.z = pf.zlens;
            h.ell = aa == bb ? 0. : (aa * aa - bb * bb) / (aa * aa + bb * bb);
            if (h.ell < 0) return set_error(LT_ERR_INVALID, "potfile " + fn + ": negative ellipticity");
            h.x = cx;
            h.y = cy;
            wcs_round_trip(h.x, h.y, ref_mode, ref_ra, ref_dec);
            h.theta = th;
            h.lum = lm;
            h.mag = mg;
            h.angle = th * dtr;
            // scaling relations (src/module_readParameters.cpp:1609-1625)
            if (h.mag != 0) {
                h.rcore = core * std::pow(10., 0.4 * (pf.mag0 - h.mag) / 2.);
                h.vdisp = pf.sigma1 * std::pow(10., 0.4 * (pf.mag0 - h.mag) / pf.vdslope1);
                h.rcut = cut * std::pow(10., 0.4 * (pf.mag0 - h.mag) * 2. / pf.slope1);
            }
            clump_parameters(h);
            P.push(h);
            ++n;
        }
        R.npotfiles += 1;
        (void)n;
    }
    R.nhalos = (int)P.type.size();

    // D_ls / D_s of every halo for the source plane of the run (src/module_readParameters.cpp:230-257)
    P.dlsds.resize(P.type.size());
    for (size_t i = 0; i < P.type.size(); ++i)
        P.dlsds[i] = cosmo_object_object(P.z[i], R.z_source, C) / cosmo_observer_object(R.z_source, C);

    // multiple images (src/module_readParameters.cpp:597-634, 852-935)
    if (!P.imagefile.empty()) {
        const std::string fn = resolve(P.imagefile, dir);
        if (fn.empty()) return set_error(LT_ERR_NOT_FOUND, "image file " + P.imagefile + " not found");
        std::ifstream im(fn.c_str());
        int old = 0;
        for (std::string s; std::getline(im, s);) {
            if (s.empty() || s[0] == '#' || tok(s, 0).empty()) continue;
            double cx, cy, a, b, th, zz, mg = 0;
            if (std::sscanf(s.c_str(), "%*s %lf %lf %lf %lf %lf %lf %lf", &cx, &cy, &a, &b, &th, &zz, &mg) < 6)
                return set_error(LT_ERR_INVALID, "image file " + fn + ": bad line: " + s);
            const int j = std::atoi(s.c_str());
            if (P.nimages.empty() || j != old) {
                P.nimages.push_back(0);
                old = j;
            }
            lt_galaxy g{};
            g.center.x = cx;
            g.center.y = cy;
            g.shape.a = a;
            g.shape.b = b;
            g.shape.theta = th;
            g.redshift = zz;
            g.mag = mg;
            P.images.push_back(g);
            P.nimages.back() += 1;
        }
        R.nsets = (int)P.nimages.size();
        R.nimagestot = (int)P.images.size();
    }
    return LT_OK;
}

}  // namespace

extern "C" {

int lt_par_load(lt_par* out, const char* path)
{
    if (!out || !path) return set_error(LT_ERR_INVALID, "NULL argument");
    *out = nullptr;
    lt_par_s* p = new lt_par_s();
    int st = parse(*p, path);
    if (st) {
        delete p;
        return st;
    }
    *out = p;
    return LT_OK;
}

int lt_par_get(lt_par p, lt_par_info* info, lt_potential_soa* lens, lt_grid_param* frame, lt_cosmo_param* cosmo)
{
    if (!p) return set_error(LT_ERR_NOT_FOUND, "parameter set is NULL");
    if (info) *info = p->info;
    if (frame) *frame = p->frame;
    if (cosmo) *cosmo = p->cosmo;
    if (lens) {
        std::memset(lens, 0, sizeof(*lens));
        const bool any = !p->type.empty();
        lens->type = any ? p->type.data() : nullptr;
        lens->position_x = any ? p->x.data() : nullptr;
        lens->position_y = any ? p->y.data() : nullptr;
        lens->b0 = any ? p->b0.data() : nullptr;
        lens->vdisp = any ? p->vdisp.data() : nullptr;
        lens->ellipticity_angle = any ? p->angle.data() : nullptr;
        lens->ellipticity = any ? p->ell.data() : nullptr;
        lens->ellipticity_potential = any ? p->epot.data() : nullptr;
        lens->rcore = any ? p->rcore.data() : nullptr;
        lens->rcut = any ? p->rcut.data() : nullptr;
        lens->z = any ? p->z.data() : nullptr;
        lens->mag = any ? p->mag.data() : nullptr;
        lens->lum = any ? p->lum.data() : nullptr;
        lens->theta = any ? p->theta.data() : nullptr;
        lens->anglecos = any ? p->cs.data() : nullptr;
        lens->anglesin = any ? p->sn.data() : nullptr;
        lens->dlsds = any ? p->dlsds.data() : nullptr;
    }
    return LT_OK;
}

int lt_par_images(lt_par p, int* nimages_per_set, lt_galaxy* images)
{
    if (!p) return set_error(LT_ERR_NOT_FOUND, "parameter set is NULL");
    if (nimages_per_set)
        for (size_t k = 0; k < p->nimages.size(); ++k) nimages_per_set[k] = p->nimages[k];
    if (images)
        for (size_t k = 0; k < p->images.size(); ++k) images[k] = p->images[k];
    return LT_OK;
}

int lt_par_destroy(lt_par p)
{
    delete p;
    return LT_OK;
}

}  // extern "C"
