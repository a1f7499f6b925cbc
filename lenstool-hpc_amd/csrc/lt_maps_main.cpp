// lt_maps_main.cpp -- lenstool_maps: the maps of a Lenstool parameter file, for its model or for every
// sample of a bayes.dat, written as FITS files. The GPU part of the reference's utils/maps/main.cpp
// (:413-587: per sample, setbayesmapmodels then map_grid_{mass,ampli,shear,dpl,potential}_GPU and
// module_writeFits) and of utils/exec/main.cpp (:397-545, one model). Here every sample of a map
// kind goes to the GPU in batches: one launch computes the map of all samples of a batch (the
// samples are a grid dimension of the kernel), instead of one launch and one round trip per sample.
//
//   lenstool_maps <file.par> [--bayes FILE] [--out DIR] [--precision fp64|fp32c] [--batch N]
//                 [--samples N] [--quiet]
//
// --batch N: samples per launch (default 64, and no more than 4 GiB of planes per launch);
// --samples N: the first N samples of the bayes file only.
//
// Output names follow the reference: <DIR>/<name>_<k>.fits for sample k, <DIR>/<name>.fits without
// samples; displacement maps as <name1>_x / <name2>_y (with _<k> for samples: the reference writes
// every sample's displacement map to the same file, only its last sample survives there).
// Map planes: the runmode line's cells x cells over the frame, z from the runmode line; the mass
// map's lens plane is the runmode's z (the reference passes runmode.z_mass), the source plane the
// source block's z_source; other maps use the first halo's z as lens plane. With a `cline` block of
// nplan >= 1 source planes, the critical lines and caustics of every sample for every plane go to
// <DIR>/critic_<plane>[_<k>].dat and caustic_<plane>[_<k>].dat, one segment per line
// (x0 y0 x1 y1 kind; kind 0 tangential, 1 radial), on the square frame of half-width dmax.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "lenstool_hip/lt_capi.h"

namespace {

struct Opts {
    std::string par, bayes, out = ".";
    int precision = LT_PREC_FP64, batch = 64, samples = -1;
    bool quiet = false;
};

int usage()
{
    std::fprintf(stderr, "usage: lenstool_maps <file.par> [--bayes FILE] [--out DIR] [--precision fp64|fp32c] "
                         "[--batch N] [--samples N] [--quiet]\n");
    return 2;
}

bool fail(int st, const char* what)
{
    if (st == LT_OK) return false;
    std::fprintf(stderr, "lenstool_maps: %s: %s (status %d)\n", what, lt_last_error(), st);
    return true;
}

// owned copy of a lens model (lt_par_get hands out the parameter set's own arrays, which the next
// sample overwrites)
struct Soa {
    int n = 0;
    std::vector<int> type;
    std::vector<double> d[16];
    lt_potential_soa soa{};

    void take(const lt_potential_soa& s, int nh)
    {
        n = nh;
        type.assign(s.type, s.type + nh);
        const double* src[16] = {s.position_x, s.position_y, s.b0, s.vdisp, s.ellipticity_angle, s.ellipticity,
                                 s.ellipticity_potential, s.rcore, s.rcut, s.z, s.mag, s.lum, s.theta, s.anglecos,
                                 s.anglesin, s.dlsds};
        for (int k = 0; k < 16; ++k) d[k].assign(src[k], src[k] + nh);
        std::memset(&soa, 0, sizeof soa);
        soa.type = type.data();
        double** dst[16] = {&soa.position_x, &soa.position_y, &soa.b0, &soa.vdisp, &soa.ellipticity_angle,
                            &soa.ellipticity, &soa.ellipticity_potential, &soa.rcore, &soa.rcut, &soa.z, &soa.mag,
                            &soa.lum, &soa.theta, &soa.anglecos, &soa.anglesin, &soa.dlsds};
        for (int k = 0; k < 16; ++k) *dst[k] = d[k].data();
    }
};

struct MapJob {
    int kind, mode, cells;
    double z;
    std::string name[2];
};

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

int main(int argc, char** argv)
{
    Opts o;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto val = [&]() -> const char* { return i + 1 < argc ? argv[++i] : nullptr; };
        if (a == "--bayes") {
            const char* v = val();
            if (!v) return usage();
            o.bayes = v;
        } else if (a == "--out") {
            const char* v = val();
            if (!v) return usage();
            o.out = v;
        } else if (a == "--precision") {
            const char* v = val();
            if (!v) return usage();
            if (!std::strcmp(v, "fp64")) o.precision = LT_PREC_FP64;
            else if (!std::strcmp(v, "fp32c")) o.precision = LT_PREC_FP32C;
            else return usage();
        } else if (a == "--batch") {
            const char* v = val();
            if (!v || (o.batch = std::atoi(v)) < 1) return usage();
        } else if (a == "--samples") {
            const char* v = val();
            if (!v) return usage();
            o.samples = std::atoi(v);
        } else if (a == "--quiet") {
            o.quiet = true;
        } else if (!a.empty() && a[0] != '-' && o.par.empty()) {
            o.par = a;
        } else {
            return usage();
        }
    }
    if (o.par.empty()) return usage();

    lt_par par = nullptr;
    if (fail(lt_par_load(&par, o.par.c_str()), o.par.c_str())) return 1;
    lt_par_info info{};
    lt_grid_param frame{};
    lt_cosmo_param cosmo{};
    lt_potential_soa lens{};
    if (fail(lt_par_get(par, &info, &lens, &frame, &cosmo), "lt_par_get")) return 1;
    int nsamples = 1;
    const bool bayes = !o.bayes.empty();
    if (bayes) {
        int ncols = 0;
        if (fail(lt_par_bayes_load(par, o.bayes.c_str(), &nsamples, &ncols), o.bayes.c_str())) return 1;
        if (o.samples >= 0 && o.samples < nsamples) nsamples = o.samples;
    }
    if (nsamples < 1) {
        std::fprintf(stderr, "lenstool_maps: no samples\n");
        return 1;
    }

    // the runmode's maps, in the reference's order: mass, ampli, shear, dpl, potential
    std::vector<MapJob> jobs;
    auto add = [&](int on, int kind, int cells, double z) {
        if (on <= 0) return;
        MapJob j{kind, on, cells, z, {}};
        char buf[4096];
        for (int w = 0; w < (kind == LT_MAP_DPL ? 2 : 1); ++w) {
            if (fail(lt_par_map_name(par, kind, w, buf, sizeof buf), "lt_par_map_name")) std::exit(1);
            j.name[w] = buf;
        }
        jobs.push_back(j);
    };
    add(info.mass, LT_MAP_MASS, info.mass_cells, info.mass_z);
    add(info.ampli, LT_MAP_AMPLI, info.ampli_cells, info.ampli_z);
    add(info.shear, LT_MAP_SHEAR, info.shear_cells, info.shear_z);
    add(info.dpl, LT_MAP_DPL, info.dpl_cells, info.dpl_z);
    add(info.poten, LT_MAP_POT, info.poten_cells, info.poten_z);
    lt_cline_info cline{};
    if (fail(lt_par_cline(par, &cline), "lt_par_cline")) return 1;
    if (jobs.empty() && cline.nplan <= 0) {
        std::fprintf(stderr, "lenstool_maps: the parameter file asks for no map and no critical line\n");
        return 1;
    }

    // the samples' models, in sample order (each applied in place, as the reference does)
    std::vector<Soa> models((size_t)nsamples);
    for (int k = 0; k < nsamples; ++k) {
        if (bayes && fail(lt_par_bayes_apply(par, k), "lt_par_bayes_apply")) return 1;
        if (fail(lt_par_get(par, nullptr, &lens, nullptr, nullptr), "lt_par_get")) return 1;
        models[(size_t)k].take(lens, info.nhalos);
    }
    const int nh = info.nhalos;

    if (fail(lt_init(0), "lt_init")) return 1;
    void* stream = nullptr;
    if (fail(lt_stream_create(&stream), "lt_stream_create")) return 1;
    double t_gpu = 0, t_all = -now_s();
    long long pixels = 0;

    for (const MapJob& j : jobs) {
        const int nc = j.cells;
        if (nc < 2) {
            std::fprintf(stderr, "lenstool_maps: map of %d cells\n", nc);
            return 1;
        }
        const double dx = (frame.xmax - frame.xmin) / (nc - 1), dy = (frame.ymax - frame.ymin) / (nc - 1);
        const int use = j.kind == LT_MAP_DPL ? LT_MAP_GRADIENT : j.kind == LT_MAP_POT ? LT_MAP_POT
                        : j.kind == LT_MAP_MASS ? LT_MAP_MASS : LT_MAP_HESSIAN;
        const int prec = o.precision;
        const int nplanes = j.kind == LT_MAP_DPL ? 2 : 1;
        const long long npix = (long long)nc * nc;
        // samples per launch: --batch, and at most 4 GiB of map planes (device and host copy)
        const long long plane_bytes = npix * (long long)sizeof(double) * nplanes;
        const int bcap = (int)std::max(1LL, (4LL << 30) / plane_bytes);
        const int bmax = std::min(std::min(o.batch, nsamples), bcap);
        std::vector<double*> dev(nplanes, nullptr);
        for (int p = 0; p < nplanes; ++p)
            if (fail(lt_dev_alloc((void**)&dev[p], npix * bmax * (long long)sizeof(double)), "lt_dev_alloc")) return 1;
        std::vector<double> host((size_t)(npix * bmax) * nplanes);
        lt_model model = nullptr;
        int model_n = 0;
        // samples in runs of equal map factors (the lens plane of a sample is its first halo's z,
        // which a bayes.dat can free), at most bmax per launch
        for (int k0 = 0; k0 < nsamples;) {
            const double zl0 = j.kind == LT_MAP_MASS ? j.z : models[(size_t)k0].d[9][0];
            const double zs = j.kind == LT_MAP_MASS ? info.z_source : j.z;
            double r = 0, mult = 1;
            if (fail(lt_map_factors(j.kind, j.mode, &cosmo, zl0, zs, dx, dy, &r, &mult), "lt_map_factors")) return 1;
            int k1 = k0 + 1;
            while (k1 < nsamples && k1 - k0 < bmax &&
                   (j.kind == LT_MAP_MASS || models[(size_t)k1].d[9][0] == zl0))
                ++k1;
            const int nb = k1 - k0;
            std::vector<lt_potential_soa> batch;
            for (int k = k0; k < k1; ++k) batch.push_back(models[(size_t)k].soa);
            const double t0 = now_s();
            if (model && model_n != nb) {
                lt_model_destroy(model);
                model = nullptr;
            }
            if (!model) {
                if (fail(lt_model_create_batch(&model, batch.data(), nb, nh, use), "lt_model_create_batch")) return 1;
                model_n = nb;
            } else if (fail(lt_model_update_batch(model, batch.data(), nb, stream), "lt_model_update_batch")) {
                return 1;
            }
            lt_grid_desc g{frame.xmin, frame.ymin, dx, dy, nc, nc, 0, 0, nc};
            if (fail(lt_map_device(model, j.kind, j.mode, dev[0], nplanes > 1 ? dev[1] : nullptr, nullptr, nullptr,
                                   nullptr, r, mult, &g, prec, 0, stream),
                     "lt_map_device"))
                return 1;
            for (int p = 0; p < nplanes; ++p)
                if (fail(lt_memcpy_d2h(host.data() + (size_t)(npix * bmax) * p, dev[p],
                                       npix * nb * (long long)sizeof(double), stream),
                         "lt_memcpy_d2h"))
                    return 1;
            if (fail(lt_stream_sync(stream), "lt_stream_sync")) return 1;
            t_gpu += now_s() - t0;
            pixels += npix * nb;
            for (int k = k0; k < k1; ++k)
                for (int p = 0; p < nplanes; ++p) {
                    std::string name = j.name[p];
                    if (j.kind == LT_MAP_DPL) name += p == 0 ? "_x" : "_y";
                    if (bayes) name += "_" + std::to_string(k);
                    const std::string path = o.out + "/" + name + ".fits";
                    const double* img = host.data() + (size_t)(npix * bmax) * p + (size_t)(npix * (k - k0));
                    const bool absolute = !(info.ref_ra == 0. && info.ref_dec == 0.);
                    if (fail(lt_fits_write(path.c_str(), img, nc, nc, frame.xmin, frame.xmax, frame.ymin, frame.ymax,
                                           absolute, info.ref_ra, info.ref_dec),
                             path.c_str()))
                        return 1;
                    if (!o.quiet) std::printf("WRITE: %s [%d x %d]\n", path.c_str(), nc, nc);
                }
            k0 = k1;
        }
        if (model) lt_model_destroy(model);
        for (double* p : dev) lt_dev_free(p);
    }
    // critical lines and caustics of every sample for every `cline` source plane, on the square
    // frame of half-width dmax with nbgridcells (or 2 dmax / step + 1) pixels a side
    int nlines = 0;
    if (cline.nplan > 0) {
        const int nc = cline.nbgridcells > 1 ? cline.nbgridcells
                                             : std::min(16385, std::max(2, (int)(2 * cline.dmax / cline.step) + 1));
        lt_grid_param cf = frame;
        cf.xmin = cf.ymin = -cline.dmax;
        cf.xmax = cf.ymax = cline.dmax;
        long long room = 1 << 16;
        std::vector<double> crit, caus;
        std::vector<int> kinds;
        for (int p = 0; p < cline.nplan; ++p)
            for (int k = 0; k < nsamples; ++k) {
                const Soa& m = models[(size_t)k];
                long long nseg = 0;
                int st = LT_ERR_OVERFLOW;
                while (st == LT_ERR_OVERFLOW) {
                    crit.resize((size_t)room * 4);
                    caus.resize((size_t)room * 4);
                    kinds.resize((size_t)room);
                    st = lt_critic_lines(&cosmo, &cf, &m.soa, nh, m.d[9][0], cline.z[p], nc, nc, o.precision,
                                         crit.data(), caus.data(), kinds.data(), room, &nseg);
                    if (st == LT_ERR_OVERFLOW) room = nseg;
                }
                // LT_ERR_TYPE: critical lines done, caustics unavailable (external-shear halos)
                const bool no_caustics = st == LT_ERR_TYPE;
                if (no_caustics) {
                    if (!o.quiet) std::fprintf(stderr, "lenstool_maps: %s\n", lt_last_error());
                } else if (fail(st, "lt_critic_lines")) {
                    return 1;
                }
                const std::string tag = std::to_string(p) + (bayes ? "_" + std::to_string(k) : std::string());
                for (int w = 0; w < (no_caustics ? 1 : 2); ++w) {
                    const std::string path = o.out + (w ? "/caustic_" : "/critic_") + tag + ".dat";
                    std::FILE* f = std::fopen(path.c_str(), "w");
                    if (!f) {
                        std::fprintf(stderr, "lenstool_maps: cannot write %s\n", path.c_str());
                        return 1;
                    }
                    std::fprintf(f, "# %s of source plane z = %.6f: x0 y0 x1 y1 [arcsec] kind (0 tangential, 1 radial)\n",
                                 w ? "caustic segments" : "critical-line segments", cline.z[p]);
                    const double* v = w ? caus.data() : crit.data();
                    for (long long q = 0; q < nseg; ++q)
                        std::fprintf(f, "%.10f %.10f %.10f %.10f %d\n", v[4 * q], v[4 * q + 1], v[4 * q + 2],
                                     v[4 * q + 3], kinds[(size_t)q]);
                    std::fclose(f);
                    if (!o.quiet) std::printf("WRITE: %s [%lld segments]\n", path.c_str(), nseg);
                }
                ++nlines;
            }
    }
    t_all += now_s();
    lt_stream_destroy(stream);
    lt_par_destroy(par);
    std::printf("maps: %d sample(s) x %zu map kind(s), %d halos; GPU %.3f s (models, maps, copies), total %.3f s, "
                "%.3e halo*pixel/s on the GPU; %d critical-line set(s)\n",
                nsamples, jobs.size(), nh, t_gpu, t_all, t_gpu > 0 ? (double)pixels * nh / t_gpu : 0.0, nlines);
    return 0;
}
