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
//   limit blocks         src/module_readParameters.cpp:1734-1937
//   bayes.dat            src/module_readParameters.cpp:53-108 (reading), 132-227 (column mapping)
// Reference behaviour kept on purpose: potentials are slotted by ascending type (file order within
// a type) and potfile clumps follow them in catalogue order; a flat universe (omegaK = 0) takes
// omegaX = 1 - omegaM whatever `lambda` says; a PIEMD with zero ellipticity gets epot = 1e-5; potfile
// positions go through the relative -> absolute -> relative coordinate round trip; `profil` inside a
// potfile block is ignored (its `type` line sets the profile). Unlike the reference, a missing
// catalogue or redshift is an error status instead of exit(-1).
#include <algorithm>
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
    int first = 0, count = 0;  // clump slots in the lens model
};

// the parameters a limit block can free, in the order the bayes.dat columns of one potential
// follow (src/module_readParameters.cpp:139-188)
enum LimitParam { LIM_X, LIM_Y, LIM_EPOT, LIM_ELL, LIM_ANGLE, LIM_RCORE, LIM_VDISP, LIM_RCUT, LIM_Z, LIM_N };
struct Limit {
    int block[LIM_N] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
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
    // optimisation state: limit blocks and lens-model slots of the potentials in file order, the
    // potfiles' scaling state, the number of image-redshift parameters, the bayes.dat rows
    std::vector<Limit> limits;
    std::vector<int> main_slot;
    std::vector<Potfile> pots;
    int n_z_param = 0;
    int bayes_ncols = 0;
    std::vector<std::vector<double>> bayes;
    // runmode map file names: ampli, shear, mass, poten, dpl x, dpl y
    std::string names[6];
    lt_cline_info cline{};

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

// b0 and the ellipticity pair of lens-model slot i after a change of its parameters
// (module_readParameters_calculatePotentialparameter_SOA, src/module_readParameters.cpp:2873-2944;
// unlike the reader of potentiel blocks, the SIS branch keeps the potential ellipticity)
void soa_parameters(lt_par_s& P, size_t i)
{
    double& ell = P.ell[i];
    double& epot = P.epot[i];
    switch (P.type[i]) {
    case 5:
        P.b0[i] = 4. * LT_PI_C2 * P.vdisp[i] * P.vdisp[i];
        if (ell == 0. && epot != 0.) epot = ell / 3.;
        else ell = epot * 3.;
        break;
    case 8:
    case 81:
        P.b0[i] = 6. * LT_PI_C2 * P.vdisp[i] * P.vdisp[i];
        if (ell == 0. && epot != 0.) ell = 2. * epot / (1. + epot * epot);
        else if (ell == 0. && epot == 0.) epot = 0.00001;
        else epot = (1. - std::sqrt(1 - ell * ell)) / ell;
        break;
    default:
        if (ell == 0. && epot != 0.) ell = epot * 3.;
        else epot = ell / 3.;
        break;
    }
}

// the scaling relations of a potfile applied to its clumps (setScalingRelations,
// src/module_readParameters.cpp:1539-1625): core radius, velocity dispersion and cut radius from
// the clump magnitude; `core` and `cut1` are in arcsec by then
void scale_potfile(lt_par_s& P, const Potfile& pf)
{
    for (int i = pf.first; i < pf.first + pf.count; ++i) {
        const double dm = 0.4 * (pf.mag0 - P.mag[i]);
        if (P.mag[i] != 0) {
            P.rcore[i] = pf.core * std::pow(10., dm / 2.);
            P.vdisp[i] = pf.sigma1 * std::pow(10., dm / pf.vdslope1);
            P.rcut[i] = pf.cut1 * std::pow(10., dm * 2. / pf.slope1);
        }
        soa_parameters(P, (size_t)i);
    }
}

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
    P.cline.dmax = 1;
    P.cline.step = 1;
    P.cline.limit_high = 10;
    P.cline.nbgridcells = 0;
    R.nbgridcells = 1000;
    R.ampli_cells = R.mass_cells = R.shear_cells = R.poten_cells = R.dpl_cells = 1000;
    R.mass_z = 0.4;
    R.poten_z = R.dpl_z = 0.8;
    R.z_source = 0.8;

    std::vector<Halo> main;
    std::vector<Potfile>& pots = P.pots;
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
                char buf[4096] = "", buf2[4096] = "";
                double v = 0;
                if (k == "nbgridcells") std::sscanf(s.c_str(), " %*s %d", &R.nbgridcells);
                else if (k == "inverse") std::sscanf(s.c_str(), " %*s %d", &R.inverse);
                else if (k == "ampli") {
                    std::sscanf(s.c_str(), " %*s %d %d %lf %4095s", &R.ampli, &R.ampli_cells, &v, buf);
                    R.ampli_z = v;
                    P.names[0] = buf;
                } else if (k == "mass") {
                    std::sscanf(s.c_str(), " %*s %d %d %lf %4095s", &R.mass, &R.mass_cells, &v, buf);
                    R.mass_z = v;
                    P.names[2] = buf;
                } else if (k == "shear") {
                    std::sscanf(s.c_str(), " %*s %d %d %lf %4095s", &R.shear, &R.shear_cells, &v, buf);
                    R.shear_z = v;
                    P.names[1] = buf;
                } else if (k == "poten") {
                    std::sscanf(s.c_str(), " %*s %d %d %lf %4095s", &R.poten, &R.poten_cells, &v, buf);
                    R.poten_z = v;
                    P.names[3] = buf;
                } else if (k == "dpl") {
                    std::sscanf(s.c_str(), " %*s %d %d %lf %4095s %4095s", &R.dpl, &R.dpl_cells, &v, buf, buf2);
                    R.dpl_z = v;
                    P.names[4] = buf;
                    P.names[5] = buf2;
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
                } else if (tok(L[i], 0) == "z_m_limit") {
                    P.n_z_param += 1;  // one bayes.dat column per image-family redshift
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
        } else if (starts(first, "cline")) {
            // critical-line planes (the reference's compiled-out reader,
            // src/module_readParameters.cpp:1008-1095): nplan n z1 z2 ..., dmax, pas / step /
            // limitLow, limitHigh, nbgridcells
            lt_cline_info& K = P.cline;
            for (++i; i < L.size() && !block_end(i); ++i) {
                const std::string k = tok(L[i], 0);
                if (k == "nplan") {
                    K.nplan = std::min(std::max(std::atoi(tok(L[i], 1).c_str()), 0), LT_CLINE_MAXPLANES);
                    for (int q = 0; q < LT_CLINE_MAXPLANES; ++q) {
                        const std::string w = tok(L[i], 2 + q);
                        if (w.empty()) break;
                        K.z[q] = std::atof(w.c_str());
                    }
                } else if (k == "dmax") K.dmax = std::atof(tok(L[i], 1).c_str());
                else if (k == "pas" || k == "step" || k == "limitLow") K.step = std::atof(tok(L[i], 1).c_str());
                else if (k == "limitHigh") K.limit_high = std::atof(tok(L[i], 1).c_str());
                else if (k == "nbgridcells") K.nbgridcells = std::atoi(tok(L[i], 1).c_str());
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
        } else if (starts(first, "limit")) {
            // the k-th limit block frees parameters of the k-th potentiel block
            Limit lim;
            for (++i; i < L.size() && !block_end(i); ++i) {
                const std::string k = tok(L[i], 0);
                int block = 0;
                std::sscanf(L[i].c_str(), " %*s %d", &block);
                if (k == "x_centre" || k == "x_center") lim.block[LIM_X] = block;
                else if (k == "y_centre" || k == "y_center") lim.block[LIM_Y] = block;
                else if (k == "v_disp") lim.block[LIM_VDISP] = block;
                else if (k == "ellip_pot") lim.block[LIM_EPOT] = block;
                else if (k == "ellipticitymass" || k == "ellipticity" || k == "ellipticite" || k == "gamma")
                    lim.block[LIM_ELL] = block;
                else if (k == "ellipticity_angle" || k == "angle_pos") lim.block[LIM_ANGLE] = block;
                else if (k == "rcut" || k == "cut_radius" || k == "cut_radius_kpc") lim.block[LIM_RCUT] = block;
                else if (k == "rcore" || k == "core_radius" || k == "core_radius_kpc") lim.block[LIM_RCORE] = block;
                else if (k == "z_lens") lim.block[LIM_Z] = block;
            }
            P.limits.push_back(lim);
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
    P.main_slot.assign(main.size(), 0);
    for (int t = 0; t < 100; ++t)
        for (size_t k = 0; k < main.size(); ++k)
            if (main[k].type == t) {
                P.main_slot[k] = (int)P.type.size();
                P.push(main[k]);
            }
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
        if (pf.ftype == 1 && pf.cut1 == DBL_MAX && pf.cutkpc1 == DBL_MAX)
            return set_error(LT_ERR_INVALID, "potfile " + pf.file + ": cut length not defined");
        if (pf.ftype == 1 && pf.core == -1 && pf.corekpc == -1)
            return set_error(LT_ERR_INVALID, "potfile " + pf.file + ": core length not defined");
        std::ifstream cat(fn.c_str());
        const double d1 = kpc_per_arcsec(pf.zlens, C);
        if (pf.cut1 == DBL_MAX && pf.cutkpc1 != DBL_MAX) pf.cut1 = pf.cutkpc1 / d1;
        if (pf.core == -1.0 && pf.corekpc != -1) pf.core = pf.corekpc / d1;
        pf.first = (int)P.type.size();
        int ref_mode = 0;
        double ref_ra = 0, ref_dec = 0;
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
            P.push(h);
            pf.count += 1;
        }
        scale_potfile(P, pf);
        R.npotfiles += 1;
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

// bayes.dat: one column per '#' line, one sample per other line (src/module_readParameters.cpp:53-108).
// Blank lines are skipped; a line with more values than columns is an error (the reference's
// stream would carry the surplus into the next sample), a line with fewer keeps its count so
// that the mapping can refuse a column that is absent.
int bayes_read(lt_par_s& P, const std::string& path)
{
    std::ifstream in(path.c_str());
    if (!in) return set_error(LT_ERR_NOT_FOUND, "bayes file " + path + " not found");
    int ncols = 0;
    std::vector<std::vector<double>> rows;
    int line = 0;
    for (std::string s; std::getline(in, s);) {
        ++line;
        if (starts(s, "#")) {
            ncols += 1;
            continue;
        }
        std::istringstream ss(s);
        std::vector<double> row;
        std::string w;
        while (ss >> w) {
            char* end = nullptr;
            const double v = std::strtod(w.c_str(), &end);
            if (end == w.c_str() || *end != '\0')
                return set_error(LT_ERR_INVALID, path + ":" + std::to_string(line) + ": not a number: " + w);
            row.push_back(v);
        }
        if (row.empty()) continue;
        if ((int)row.size() > ncols)
            return set_error(LT_ERR_INVALID, path + ":" + std::to_string(line) + ": " + std::to_string(row.size()) +
                                                 " values for " + std::to_string(ncols) + " columns");
        rows.push_back(std::move(row));
    }
    P.bayes_ncols = ncols;
    P.bayes.swap(rows);
    return LT_OK;
}

// the lens model of bayes sample k (module_readParameters_setbayesmapmodels,
// src/module_readParameters.cpp:132-227): from column 2 on, the freed parameters of each potential
// in file order (x, y, epot, ellipticity, angle [deg], rcore, vdisp, rcut, z), b0 and the
// ellipticity pair recomputed; with potfiles, the image-redshift columns skipped and then each
// potfile's cut radius [arcsec] and sigma if freed, its scaling relations re-applied. In place, as
// the reference: parameters not in the file keep their values, D_ls/D_s is not recomputed.
int bayes_apply(lt_par_s& P, int k)
{
    if (k < 0 || k >= (int)P.bayes.size())
        return set_error(LT_ERR_INVALID, "bayes sample " + std::to_string(k) + " of " + std::to_string(P.bayes.size()));
    const std::vector<double>& row = P.bayes[(size_t)k];
    const double dtr = std::acos(-1.) / 180.;
    size_t col = 2;
    auto next = [&]() { return row[col++]; };
    // every needed column must be there (checked first: a refused sample leaves the model as it was)
    {
        size_t need = 2;
        for (size_t i = 0; i < P.main_slot.size(); ++i)
            if (i < P.limits.size())
                for (int q = 0; q < LIM_N; ++q) need += P.limits[i].block[q] >= 1;
        if (!P.pots.empty()) {
            need += (size_t)P.n_z_param;
            for (const Potfile& pf : P.pots) need += (pf.ircut > 0) + (pf.isigma > 0);
        }
        if (need > row.size())
            return set_error(LT_ERR_INVALID, "bayes sample " + std::to_string(k) + " has " + std::to_string(row.size()) +
                                                 " values, the model's free parameters need " + std::to_string(need));
    }
    for (size_t i = 0; i < P.main_slot.size(); ++i) {
        const size_t s = (size_t)P.main_slot[i];
        if (i < P.limits.size()) {
            const Limit& lim = P.limits[i];
            if (lim.block[LIM_X] >= 1) P.x[s] = next();
            if (lim.block[LIM_Y] >= 1) P.y[s] = next();
            if (lim.block[LIM_EPOT] >= 1) P.epot[s] = next();
            if (lim.block[LIM_ELL] >= 1) P.ell[s] = next();
            if (lim.block[LIM_ANGLE] >= 1) {
                P.angle[s] = next() * dtr;
                P.cs[s] = std::cos(P.angle[s]);
                P.sn[s] = std::sin(P.angle[s]);
            }
            if (lim.block[LIM_RCORE] >= 1) P.rcore[s] = next();
            if (lim.block[LIM_VDISP] >= 1) P.vdisp[s] = next();
            if (lim.block[LIM_RCUT] >= 1) P.rcut[s] = next();
            if (lim.block[LIM_Z] >= 1) P.z[s] = next();
        }
        soa_parameters(P, s);
    }
    if (!P.pots.empty()) {
        col += (size_t)P.n_z_param;
        for (Potfile& pf : P.pots) {
            if (pf.ircut > 0) pf.cut1 = next();
            if (pf.isigma > 0) pf.sigma1 = next();
            scale_potfile(P, pf);
        }
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

int lt_par_bayes_load(lt_par p, const char* path, int* nsamples, int* ncolumns)
{
    if (!p || !path) return set_error(LT_ERR_INVALID, "NULL argument");
    const int st = bayes_read(*p, path);
    if (st) return st;
    if (nsamples) *nsamples = (int)p->bayes.size();
    if (ncolumns) *ncolumns = p->bayes_ncols;
    return LT_OK;
}

int lt_par_bayes_apply(lt_par p, int sample)
{
    if (!p) return set_error(LT_ERR_NOT_FOUND, "parameter set is NULL");
    return bayes_apply(*p, sample);
}

int lt_par_cline(lt_par p, lt_cline_info* cline)
{
    if (!p || !cline) return set_error(LT_ERR_INVALID, "NULL argument");
    *cline = p->cline;
    return LT_OK;
}

int lt_par_map_name(lt_par p, int kind, int which, char* buf, int cap)
{
    if (!p || !buf || cap <= 0) return set_error(LT_ERR_INVALID, "NULL argument");
    int k = -1;
    switch (kind) {
    case LT_MAP_AMPLI: k = 0; break;
    case LT_MAP_SHEAR: k = 1; break;
    case LT_MAP_MASS: k = 2; break;
    case LT_MAP_POT: k = 3; break;
    case LT_MAP_DPL: k = which ? 5 : 4; break;
    default: return set_error(LT_ERR_INVALID, "map kind " + std::to_string(kind) + " has no runmode file name");
    }
    const std::string& n = p->names[k];
    if ((int)n.size() >= cap) return set_error(LT_ERR_INVALID, "name buffer too small for " + n);
    std::memcpy(buf, n.c_str(), n.size() + 1);
    return LT_OK;
}

int lt_par_destroy(lt_par p)
{
    delete p;
    return LT_OK;
}

}  // extern "C"
