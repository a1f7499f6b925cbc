// lt_shim.cpp -- the reference's GPU call signatures (include/lenstool_hip/lenstool_gpu.hpp) on top of
// the C-ABI. Each function cites the reference function it stands in for.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "lenstool_hip/lenstool_gpu.hpp"
#include "lenstool_hip/lt_capi.h"

static_assert(sizeof(Potential_SOA) == sizeof(lt_potential_soa), "Potential_SOA layout");
static_assert(sizeof(grid_param) == sizeof(lt_grid_param), "grid_param layout");
static_assert(sizeof(cosmo_param) == sizeof(lt_cosmo_param), "cosmo_param layout");
static_assert(sizeof(galaxy) == sizeof(lt_galaxy), "galaxy layout");
static_assert(sizeof(point) == sizeof(lt_point), "point layout");

namespace {

thread_local int g_status = LT_OK;

const lt_potential_soa* L(const Potential_SOA* p) { return reinterpret_cast<const lt_potential_soa*>(p); }
const lt_grid_param* G(const grid_param* p) { return reinterpret_cast<const lt_grid_param*>(p); }
const lt_cosmo_param* Cz(const cosmo_param* p) { return reinterpret_cast<const lt_cosmo_param*>(p); }

void report(int st, const char* fn)
{
    g_status = st;
    if (st == LT_OK) return;
    std::fprintf(stderr, "lenstool_hip: %s: %s (status %d)\n", fn, lt_last_error(), st);
    const char* e = std::getenv("LT_EXIT_ON_ERROR");
    if (e && *e && std::strcmp(e, "0") != 0) std::exit(-1);
}

double zl_of(const Potential_SOA* lens) { return (lens && lens->z) ? lens->z[0] : 0.0; }

// The generic path for a map function that is not one of this library's: Hessian planes on the
// device, then the caller's function on them (what the reference's map_grid_*_GPU do).
void generic_map(map_gpu_function_t f, type_t* map, const cosmo_param* cosmo, const grid_param* frame,
                 const Potential_SOA* lens, int nhalos, double zl, double zs, type_t dx, type_t dy, int nx, int ny,
                 int istart, int jstart, int use, const char* fn)
{
    lt_model m = nullptr;
    double* planes[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    const long long n = (long long)nx * ny;
    int st = lt_model_create(&m, L(lens), nhalos, use);
    for (int k = 0; k < 5 && st == LT_OK; ++k) st = lt_dev_alloc((void**)&planes[k], n * (long long)sizeof(double));
    if (st == LT_OK) {
        lt_grid_desc g{frame->xmin, frame->ymin, dx, dy, nx, ny, istart, jstart, nx};
        st = lt_map_device(m, LT_MAP_HESSIAN, 0, planes[0], planes[1], planes[2], planes[3], nullptr, 1.0, 1.0, &g,
                           LT_PREC_FP64, 0, nullptr);
    }
    if (st == LT_OK) st = lt_stream_sync(nullptr);
    if (st == LT_OK) {
        const double dl0s = lt_cosmo_object_object(zl, zs, Cz(cosmo));
        const double dos = lt_cosmo_observer_object(zs, Cz(cosmo));
        const double dol = lt_cosmo_observer_object(zl, Cz(cosmo));
        f(planes[4], planes[0], planes[1], planes[2], planes[3], dl0s, dos, dol, cosmo->h, zs, nx, ny, frame);
        st = lt_stream_sync(nullptr);
    }
    if (st == LT_OK) st = lt_memcpy_d2h(map, planes[4], n * (long long)sizeof(double), nullptr);
    for (int k = 0; k < 5; ++k) lt_dev_free(planes[k]);
    if (m) lt_model_destroy(m);
    report(st, fn);
}

// epilogue of a map function on device planes
void device_epilogue(int kind, int mode, type_t* map, type_t* a, type_t* b, type_t* c, double r, double mult, int nx,
                     int ny, const char* fn)
{
    report(lt_map_epilogue_device(kind, mode, map, nullptr, nullptr, nullptr, nullptr, a, b, c, (long long)nx * ny, r,
                                  mult, nullptr),
           fn);
}

double mass_mult(int mode, double dl0s, double ds, double dl, double h, int nx, int ny, const grid_param* frame)
{
    switch (mode) {
    case 2: return LT_CH4PIG * h / dl / dl0s * ds;
    case 3: {
        const double dx = (frame->xmax - frame->xmin) / (nx - 1);
        const double dy = (frame->ymax - frame->ymin) / (ny - 1);
        return LT_MCRIT12 / h * dx * dy * dl / dl0s * ds;
    }
    case 4: return LT_CH0_4PIG * h / dl / dl0s * ds;
    default: return 1.0;
    }
}

}  // namespace

int lenstool_gpu_last_status() { return g_status; }

// ---- gradient ------------------------------------------------------------------------------------
// gradient_grid_GPU, src/grid_gradient_GPU.cu:77-84 (square) and :87-179
void gradient_grid_GPU(type_t* gx, type_t* gy, const grid_param* frame, const Potential_SOA* lens, int nhalos,
                       int nbgridcells)
{
    const type_t dx = (frame->xmax - frame->xmin) / (nbgridcells - 1);
    const type_t dy = (frame->ymax - frame->ymin) / (nbgridcells - 1);
    gradient_grid_GPU(gx, gy, frame, lens, nhalos, dx, dy, nbgridcells, nbgridcells, 0, 0);
}

void gradient_grid_GPU(type_t* gx, type_t* gy, const grid_param* frame, const Potential_SOA* lens, int nhalos,
                       type_t dx, type_t dy, int nx, int ny, int istart, int jstart)
{
    report(lt_gradient_grid(gx, gy, G(frame), L(lens), nhalos, dx, dy, nx, ny, istart, jstart, LT_PREC_FP64),
           "gradient_grid_GPU");
}

// gradient_grid_GPU_UM, src/grid_gradient_GPU.cu:216-245 (managed memory is ordinary host-visible
// memory for this library: the same path)
void gradient_grid_GPU_UM(type_t* gx, type_t* gy, const grid_param* frame, const Potential_SOA* lens, int nhalos,
                          type_t dx, type_t dy, int nx, int ny, int istart, int jstart)
{
    report(lt_gradient_grid(gx, gy, G(frame), L(lens), nhalos, dx, dy, nx, ny, istart, jstart, LT_PREC_FP64),
           "gradient_grid_GPU_UM");
}

// ---- Hessian -------------------------------------------------------------------------------------
// gradient2_grid_GPU, src/grid_gradient2_GPU.cu:59-65 (square) and :67-183
void gradient2_grid_GPU(type_t* a, type_t* b, type_t* c, type_t* d, const grid_param* frame, const Potential_SOA* lens,
                        int nhalos, int nbgridcells)
{
    const type_t dx = (frame->xmax - frame->xmin) / (nbgridcells - 1);
    const type_t dy = (frame->ymax - frame->ymin) / (nbgridcells - 1);
    gradient2_grid_GPU(a, b, c, d, frame, lens, nhalos, dx, dy, nbgridcells, nbgridcells, 0, 0);
}

void gradient2_grid_GPU(type_t* a, type_t* b, type_t* c, type_t* d, const grid_param* frame, const Potential_SOA* lens,
                        int nhalos, type_t dx, type_t dy, int nx, int ny, int istart, int jstart)
{
    report(lt_gradient2_grid(a, b, c, d, G(frame), L(lens), nhalos, dx, dy, nx, ny, istart, jstart, LT_PREC_FP64),
           "gradient2_grid_GPU");
}

// ---- potential -----------------------------------------------------------------------------------
// potential_grid_GPU, src/grid_potential_GPU.cuh:31-32
void potential_grid_GPU(type_t* phi, const grid_param* frame, const Potential_SOA* lens, int nhalos, type_t dx,
                        type_t dy, int nx, int ny, int istart, int jstart)
{
    report(lt_potential_grid(phi, G(frame), L(lens), nhalos, dx, dy, nx, ny, istart, jstart), "potential_grid_GPU");
}

void potential_grid_GPU(type_t* phi, const grid_param* frame, const Potential_SOA* lens, int nhalos, int nbgridcells)
{
    const type_t dx = (frame->xmax - frame->xmin) / (nbgridcells - 1);
    const type_t dy = (frame->ymax - frame->ymin) / (nbgridcells - 1);
    potential_grid_GPU(phi, frame, lens, nhalos, dx, dy, nbgridcells, nbgridcells, 0, 0);
}

// ---- map functions on device planes (src/grid_map_ampli_GPU.cu:305-547, grid_map_shear_GPU.cu,
// grid_map_mass_GPU.cu:219-305, grid_map_pot_GPU.cu:178-270, grid_map_dpl_GPU.cu:150-208) --------------
#define AMPLI_FN(k)                                                                                               \
    void amplif_##k##_grid_CPU_GPU(type_t* map, type_t* a, type_t* b, type_t* c, type_t* d, type_t dl0s, type_t ds, \
                                   type_t dl, type_t h, type_t z, int nx, int ny, const grid_param* frame)         \
    {                                                                                                             \
        (void)d, (void)dl, (void)h, (void)z, (void)frame;                                                         \
        device_epilogue(LT_MAP_AMPLI, k, map, a, b, c, dl0s / ds, 1.0, nx, ny, "amplif_" #k "_grid_CPU_GPU");      \
    }
AMPLI_FN(1)
AMPLI_FN(2)
AMPLI_FN(3)
AMPLI_FN(4)
AMPLI_FN(5)
AMPLI_FN(6)
#undef AMPLI_FN

#define SHEAR_FN(k)                                                                                              \
    void shear_##k##_grid_CPU_GPU(type_t* map, type_t* a, type_t* b, type_t* c, type_t* d, type_t dl0s, type_t ds, \
                                  type_t dl, type_t h, type_t z, int nx, int ny, const grid_param* frame)         \
    {                                                                                                            \
        (void)d, (void)dl, (void)h, (void)z, (void)frame;                                                        \
        device_epilogue(LT_MAP_SHEAR, k, map, a, b, c, dl0s / ds, 1.0, nx, ny, "shear_" #k "_grid_CPU_GPU");      \
    }
SHEAR_FN(1)
SHEAR_FN(2)
SHEAR_FN(3)
SHEAR_FN(4)
#undef SHEAR_FN

#define MASS_FN(k)                                                                                               \
    void mass_##k##_grid_CPU_GPU(type_t* map, type_t* a, type_t* b, type_t* c, type_t* d, type_t dl0s, type_t ds,  \
                                 type_t dl, type_t h, type_t z, int nx, int ny, const grid_param* frame)          \
    {                                                                                                            \
        (void)d, (void)z;                                                                                        \
        device_epilogue(LT_MAP_MASS, k, map, a, b, c, 1.0, mass_mult(k, dl0s, ds, dl, h, nx, ny, frame), nx, ny,  \
                        "mass_" #k "_grid_CPU_GPU");                                                             \
    }
MASS_FN(1)
MASS_FN(2)
MASS_FN(3)
MASS_FN(4)
#undef MASS_FN

void potential_1_grid_CPU_GPU(type_t* phi, type_t dl0s, type_t ds, type_t dl, type_t h, type_t z, int nx, int ny,
                              const grid_param* frame)
{
    (void)dl, (void)h, (void)z, (void)frame;
    report(lt_scale_device(phi, dl0s / ds, (long long)nx * ny, nullptr), "potential_1_grid_CPU_GPU");
}

void potential_2_grid_CPU_GPU(type_t* phi, type_t dl0s, type_t ds, type_t dl, type_t h, type_t z, int nx, int ny,
                              const grid_param* frame)
{
    (void)dl0s, (void)ds, (void)z, (void)frame;
    report(lt_scale_device(phi, LT_CH2PIG * h / dl, (long long)nx * ny, nullptr), "potential_2_grid_CPU_GPU");
}

// mode 3 is an empty stub in the reference (src/grid_map_pot_GPU.cu:253-270)
void potential_3_grid_CPU_GPU(type_t*, type_t, type_t, type_t, type_t, type_t, int, int, const grid_param*)
{
    report(LT_OK, "potential_3_grid_CPU_GPU");
}

void dpl_grid_CPU_GPU(type_t* map, type_t dl0s, type_t ds, type_t dl, type_t h, type_t z, int nx, int ny,
                      const grid_param* frame)
{
    (void)dl, (void)h, (void)z, (void)frame;
    report(lt_scale_device(map, dl0s / ds, (long long)nx * ny, nullptr), "dpl_grid_CPU_GPU");
}

// ---- selectors (src/grid_map_ampli_GPU.cu:65-90, grid_map_shear_GPU.cu:64-83,
// grid_map_mass_GPU.cu:60-76, grid_map_pot_GPU.cu:64-78); an unknown mode returns NULL and sets an
// error status instead of exit(-1) ------------------------------------------------------------------
static const map_gpu_function_t kAmpli[6] = {amplif_1_grid_CPU_GPU, amplif_2_grid_CPU_GPU, amplif_3_grid_CPU_GPU,
                                             amplif_4_grid_CPU_GPU, amplif_5_grid_CPU_GPU, amplif_6_grid_CPU_GPU};
static const map_gpu_function_t kShear[4] = {shear_1_grid_CPU_GPU, shear_2_grid_CPU_GPU, shear_3_grid_CPU_GPU,
                                             shear_4_grid_CPU_GPU};
static const map_gpu_function_t kMass[4] = {mass_1_grid_CPU_GPU, mass_2_grid_CPU_GPU, mass_3_grid_CPU_GPU,
                                            mass_4_grid_CPU_GPU};
static const map_pot_function_t kPot[3] = {potential_1_grid_CPU_GPU, potential_2_grid_CPU_GPU,
                                           potential_3_grid_CPU_GPU};

template <class F, int N>
static F select_mode(const F (&table)[N], int mode, const char* what)
{
    if (mode >= 1 && mode <= N) {
        g_status = LT_OK;
        return table[mode - 1];
    }
    std::fprintf(stderr, "lenstool_hip: %s mode %d not supported\n", what, mode);
    g_status = LT_ERR_INVALID;
    const char* e = std::getenv("LT_EXIT_ON_ERROR");
    if (e && *e && std::strcmp(e, "0") != 0) std::exit(-1);
    return nullptr;
}

// select_map_dpl_function, src/grid_map_dpl_GPU.cuh:39 (its body is compiled out in the reference,
// src/grid_map_dpl_GPU.cu:58-71, which names dpl_1_grid_CPU_GPU for dpl mode 1): mode 1 gives the
// displacement scaling D_ls/D_s applied in place to `map` (map_grid_dpl_GPU itself ignores its map
// function and applies that scaling to both planes, src/grid_map_dpl_GPU.cu:75-190).
void dpl_1_grid_CPU_GPU(type_t* map, type_t* a, type_t* b, type_t* c, type_t* d, type_t dl0s, type_t ds, type_t dl,
                        type_t h, type_t z, int nx, int ny, const grid_param* frame)
{
    (void)a, (void)b, (void)c, (void)d, (void)dl, (void)h, (void)z, (void)frame;
    report(lt_scale_device(map, dl0s / ds, (long long)nx * ny, nullptr), "dpl_1_grid_CPU_GPU");
}

static const map_gpu_function_t kDpl[1] = {dpl_1_grid_CPU_GPU};
map_gpu_function_t select_map_dpl_function(const runmode_param* rm) { return select_mode(kDpl, rm->dpl, "dpl"); }
map_gpu_function_t select_map_ampli_function(const runmode_param* rm) { return select_mode(kAmpli, rm->amplif, "ampli"); }
map_gpu_function_t select_map_shear_function(const runmode_param* rm) { return select_mode(kShear, rm->shear, "shear"); }
map_gpu_function_t select_map_mass_function(const runmode_param* rm) { return select_mode(kMass, rm->mass, "mass"); }
map_pot_function_t select_map_potential_function(const runmode_param* rm)
{
    return select_mode(kPot, rm->potential, "potential");
}

template <class F, int N>
static int mode_of(const F (&table)[N], F f)
{
    for (int k = 0; k < N; ++k)
        if (table[k] == f) return k + 1;
    return 0;
}

// ---- whole maps ------------------------------------------------------------------------------------
// map_grid_ampli_GPU, src/grid_map_ampli_GPU.cu:93-192: a map function of this library is fused
// into the Hessian kernel; any other is applied to the Hessian planes afterwards.
void map_grid_ampli_GPU(map_gpu_function_t f, type_t* map, const cosmo_param* cosmo, const grid_param* frame,
                        const Potential_SOA* lens, int nhalos, int nbgridcells, int mode, type_t z)
{
    const type_t dx = (frame->xmax - frame->xmin) / (nbgridcells - 1);
    const type_t dy = (frame->ymax - frame->ymin) / (nbgridcells - 1);
    map_grid_ampli_GPU(f, map, cosmo, frame, lens, nhalos, mode, z, dx, dy, nbgridcells, nbgridcells, 0, 0);
}

void map_grid_ampli_GPU(map_gpu_function_t f, type_t* map, const cosmo_param* cosmo, const grid_param* frame,
                        const Potential_SOA* lens, int nhalos, int mode, type_t z, type_t dx, type_t dy, int nx, int ny,
                        int istart, int jstart)
{
    (void)mode;
    const int k = mode_of(kAmpli, f);
    if (k)
        report(lt_map_grid(LT_MAP_AMPLI, k, map, nullptr, Cz(cosmo), G(frame), L(lens), nhalos, zl_of(lens), z, dx, dy,
                           nx, ny, istart, jstart, LT_PREC_FP64),
               "map_grid_ampli_GPU");
    else
        generic_map(f, map, cosmo, frame, lens, nhalos, zl_of(lens), z, dx, dy, nx, ny, istart, jstart, LT_MAP_HESSIAN,
                    "map_grid_ampli_GPU");
}

// map_grid_shear_GPU, src/grid_map_shear_GPU.cu:86-205
void map_grid_shear_GPU(map_gpu_function_t f, type_t* map, const cosmo_param* cosmo, const grid_param* frame,
                        const Potential_SOA* lens, int nhalos, int nbgridcells, int mode, type_t z)
{
    const type_t dx = (frame->xmax - frame->xmin) / (nbgridcells - 1);
    const type_t dy = (frame->ymax - frame->ymin) / (nbgridcells - 1);
    map_grid_shear_GPU(f, map, cosmo, frame, lens, nhalos, mode, z, dx, dy, nbgridcells, nbgridcells, 0, 0);
}

void map_grid_shear_GPU(map_gpu_function_t f, type_t* map, const cosmo_param* cosmo, const grid_param* frame,
                        const Potential_SOA* lens, int nhalos, int mode, type_t z, type_t dx, type_t dy, int nx, int ny,
                        int istart, int jstart)
{
    (void)mode;
    const int k = mode_of(kShear, f);
    if (k)
        report(lt_map_grid(LT_MAP_SHEAR, k, map, nullptr, Cz(cosmo), G(frame), L(lens), nhalos, zl_of(lens), z, dx, dy,
                           nx, ny, istart, jstart, LT_PREC_FP64),
               "map_grid_shear_GPU");
    else
        generic_map(f, map, cosmo, frame, lens, nhalos, zl_of(lens), z, dx, dy, nx, ny, istart, jstart, LT_MAP_HESSIAN,
                    "map_grid_shear_GPU");
}

// map_grid_mass_GPU, src/grid_map_mass_GPU.cu:79-216 (Kmap Hessian; zl = 0 means lens->z[0])
void map_grid_mass_GPU(map_gpu_function_t f, type_t* map, const cosmo_param* cosmo, const grid_param* frame,
                       const Potential_SOA* lens, int nhalos, int nbgridcells, int mode, type_t zl, type_t zs)
{
    const type_t dx = (frame->xmax - frame->xmin) / (nbgridcells - 1);
    const type_t dy = (frame->ymax - frame->ymin) / (nbgridcells - 1);
    map_grid_mass_GPU(f, map, cosmo, frame, lens, nhalos, mode, zl, zs, dx, dy, nbgridcells, nbgridcells, 0, 0);
}

void map_grid_mass_GPU(map_gpu_function_t f, type_t* map, const cosmo_param* cosmo, const grid_param* frame,
                       const Potential_SOA* lens, int nhalos, int mode, type_t zl, type_t zs, type_t dx, type_t dy,
                       int nx, int ny, int istart, int jstart)
{
    (void)mode;
    const int k = mode_of(kMass, f);
    if (zl == 0) zl = zl_of(lens);
    if (k)
        report(lt_map_grid(LT_MAP_MASS, k, map, nullptr, Cz(cosmo), G(frame), L(lens), nhalos, zl, zs, dx, dy, nx, ny,
                           istart, jstart, LT_PREC_FP64),
               "map_grid_mass_GPU");
    else
        generic_map(f, map, cosmo, frame, lens, nhalos, zl, zs, dx, dy, nx, ny, istart, jstart, LT_MAP_MASS,
                    "map_grid_mass_GPU");
}

// map_grid_potential_GPU, src/grid_map_pot_GPU.cu:81-176
void map_grid_potential_GPU(map_pot_function_t f, type_t* map, const cosmo_param* cosmo, const grid_param* frame,
                            const Potential_SOA* lens, int nhalos, int nbgridcells, int mode, type_t z)
{
    const type_t dx = (frame->xmax - frame->xmin) / (nbgridcells - 1);
    const type_t dy = (frame->ymax - frame->ymin) / (nbgridcells - 1);
    map_grid_potential_GPU(f, map, cosmo, frame, lens, nhalos, mode, z, dx, dy, nbgridcells, nbgridcells, 0, 0);
}

void map_grid_potential_GPU(map_pot_function_t f, type_t* map, const cosmo_param* cosmo, const grid_param* frame,
                            const Potential_SOA* lens, int nhalos, int mode, type_t z, type_t dx, type_t dy, int nx,
                            int ny, int istart, int jstart)
{
    (void)mode;
    int k = mode_of(kPot, f);
    if (!k) {
        // a foreign potential function: raw potential on the device, then the function on it
        double* phi = nullptr;
        const long long n = (long long)nx * ny;
        int st = lt_dev_alloc((void**)&phi, n * (long long)sizeof(double));
        lt_model m = nullptr;
        if (st == LT_OK) st = lt_model_create(&m, L(lens), nhalos, LT_MAP_POT);
        if (st == LT_OK) {
            lt_grid_desc g{frame->xmin, frame->ymin, dx, dy, nx, ny, istart, jstart, nx};
            st = lt_map_device(m, LT_MAP_POT, 3, phi, nullptr, nullptr, nullptr, nullptr, 1.0, 1.0, &g, LT_PREC_FP64, 0,
                               nullptr);
        }
        if (st == LT_OK) st = lt_stream_sync(nullptr);
        if (st == LT_OK) {
            const double zl = zl_of(lens);
            f(phi, lt_cosmo_object_object(zl, z, Cz(cosmo)), lt_cosmo_observer_object(z, Cz(cosmo)),
              lt_cosmo_observer_object(zl, Cz(cosmo)), cosmo->h, z, nx, ny, frame);
            st = lt_stream_sync(nullptr);
        }
        if (st == LT_OK) st = lt_memcpy_d2h(map, phi, n * (long long)sizeof(double), nullptr);
        lt_dev_free(phi);
        if (m) lt_model_destroy(m);
        report(st, "map_grid_potential_GPU");
        return;
    }
    report(lt_map_grid(LT_MAP_POT, k, map, nullptr, Cz(cosmo), G(frame), L(lens), nhalos, zl_of(lens), z, dx, dy, nx, ny,
                       istart, jstart, LT_PREC_FP64),
           "map_grid_potential_GPU");
}

// map_grid_dpl_GPU, src/grid_map_dpl_GPU.cu:75-190 (the reference ignores mapfunction and scales
// both gradient planes by D_ls/D_s)
void map_grid_dpl_GPU(map_gpu_function_t f, type_t* dplx, type_t* dply, const cosmo_param* cosmo,
                      const grid_param* frame, const Potential_SOA* lens, int nhalos, int nbgridcells, int mode,
                      type_t z)
{
    const type_t dx = (frame->xmax - frame->xmin) / (nbgridcells - 1);
    const type_t dy = (frame->ymax - frame->ymin) / (nbgridcells - 1);
    map_grid_dpl_GPU(f, dplx, dply, cosmo, frame, lens, nhalos, mode, z, dx, dy, nbgridcells, nbgridcells, 0, 0);
}

void map_grid_dpl_GPU(map_gpu_function_t f, type_t* dplx, type_t* dply, const cosmo_param* cosmo,
                      const grid_param* frame, const Potential_SOA* lens, int nhalos, int mode, type_t z, type_t dx,
                      type_t dy, int nx, int ny, int istart, int jstart)
{
    (void)f, (void)mode;
    report(lt_map_grid(LT_MAP_DPL, 1, dplx, dply, Cz(cosmo), G(frame), L(lens), nhalos, zl_of(lens), z, dx, dy, nx, ny,
                       istart, jstart, LT_PREC_FP64),
           "map_grid_dpl_GPU");
}

// ---- the reference's launchers on device data ------------------------------------------------------
// module_potentialDerivatives_totalGradient_SOA_CPU_GPU (src/grid_gradient_GPU.cu:365-420),
// module_potentialDerivatives_totalGradient2_SOA_CPU_GPU (src/grid_gradient2_GPU.cu:186-209) and
// module_potentialDerivatives_Kmap_SOA_CPU_GPU (src/grid_gradient2_GPU.cu:212-236) take device output
// planes, a grid_param in device memory and a Potential_SOA that itself lives in device memory, whose
// arrays are device arrays; only type, position_x/y, b0, ellipticity_angle, ellipticity_potential,
// rcore, rcut, anglecos/sin (and dlsds for the Kmap) are valid there (src/grid_gradient2_GPU.cu:101-147).
// They read the descriptor and those arrays back, pack the model on the host and run the map kernel
// into the caller's planes (pitch nx), synchronously as the reference does.
namespace {

struct HostLens {
    Potential_SOA soa{};
    std::vector<int> type;
    std::vector<double> v[11];
};

int read_device_array(std::vector<double>& dst, const double* src, int n)
{
    dst.assign((size_t)n, 0.0);
    if (!src || n == 0) return LT_OK;
    return lt_memcpy_d2h(dst.data(), src, (long long)n * (long long)sizeof(double), nullptr);
}

int read_device_lens(const Potential_SOA* lens_gpu, int nhalos, bool kmap, HostLens& h)
{
    if (!lens_gpu || nhalos < 0) return LT_ERR_INVALID;
    Potential_SOA d{};
    int st = lt_memcpy_d2h(&d, lens_gpu, (long long)sizeof(Potential_SOA), nullptr);
    if (st) return st;
    if (!d.type || !d.position_x || !d.position_y || !d.b0 || !d.ellipticity_potential || !d.anglecos || !d.anglesin)
        return LT_ERR_INVALID;
    h.type.assign((size_t)nhalos, 0);
    if (nhalos && (st = lt_memcpy_d2h(h.type.data(), d.type, (long long)nhalos * 4, nullptr))) return st;
    const double* src[11] = {d.position_x, d.position_y,      d.b0,       d.ellipticity_angle,
                             d.ellipticity_potential, d.rcore, d.rcut, d.anglecos,
                             d.anglesin,   kmap ? d.dlsds : nullptr, nullptr};
    for (int k = 0; k < 10; ++k)
        if ((st = read_device_array(h.v[k], src[k], nhalos))) return st;
    h.soa.type = h.type.data();
    h.soa.position_x = h.v[0].data();
    h.soa.position_y = h.v[1].data();
    h.soa.b0 = h.v[2].data();
    h.soa.ellipticity_angle = d.ellipticity_angle ? h.v[3].data() : nullptr;
    h.soa.ellipticity_potential = h.v[4].data();
    h.soa.rcore = d.rcore ? h.v[5].data() : nullptr;
    h.soa.rcut = d.rcut ? h.v[6].data() : nullptr;
    h.soa.anglecos = h.v[7].data();
    h.soa.anglesin = h.v[8].data();
    h.soa.dlsds = (kmap && d.dlsds) ? h.v[9].data() : nullptr;
    return LT_OK;
}

// kind LT_MAP_GRADIENT (p0, p1) or LT_MAP_HESSIAN (p0..p3); use LT_MAP_GRADIENT, LT_MAP_HESSIAN or
// LT_MAP_MASS (Kmap runs)
void device_launch(int kind, int use, type_t* p0, type_t* p1, type_t* p2, type_t* p3, const grid_param* frame_gpu,
                   const Potential_SOA* lens_gpu, int nhalos, type_t dx, type_t dy, int nx, int ny, int istart,
                   int jstart, bool square, const char* fn)
{
    HostLens h;
    grid_param frame{};
    int st = frame_gpu ? lt_memcpy_d2h(&frame, frame_gpu, (long long)sizeof(grid_param), nullptr) : LT_ERR_INVALID;
    if (st == LT_OK && square) {
        if (nx < 2) st = LT_ERR_INVALID;
        dx = (frame.xmax - frame.xmin) / (nx - 1);
        dy = (frame.ymax - frame.ymin) / (ny - 1);
    }
    if (st == LT_OK) st = read_device_lens(lens_gpu, nhalos, use == LT_MAP_MASS, h);
    lt_model m = nullptr;
    if (st == LT_OK) st = lt_model_create(&m, L(&h.soa), nhalos, use);
    if (st == LT_OK) {
        lt_grid_desc g{frame.xmin, frame.ymin, dx, dy, nx, ny, istart, jstart, nx};
        st = lt_map_device(m, kind, 0, p0, p1, p2, p3, nullptr, 1.0, 1.0, &g, LT_PREC_FP64, 0, nullptr);
    }
    if (st == LT_OK) st = lt_stream_sync(nullptr);
    if (m) lt_model_destroy(m);
    report(st, fn);
}

}  // namespace

void module_potentialDerivatives_totalGradient_SOA_CPU_GPU(type_t* gx, type_t* gy, const grid_param* frame,
                                                           const Potential_SOA* lens_gpu, int nhalos, type_t dx,
                                                           type_t dy, int nx, int ny, int istart, int jstart)
{
    device_launch(LT_MAP_GRADIENT, LT_MAP_GRADIENT, gx, gy, nullptr, nullptr, frame, lens_gpu, nhalos, dx, dy, nx, ny,
                  istart, jstart, false, "module_potentialDerivatives_totalGradient_SOA_CPU_GPU");
}

void module_potentialDerivatives_totalGradient_SOA_CPU_GPU(type_t* gx, type_t* gy, const grid_param* frame,
                                                           const Potential_SOA* lens_gpu, int nbgridcells, int nhalos)
{
    device_launch(LT_MAP_GRADIENT, LT_MAP_GRADIENT, gx, gy, nullptr, nullptr, frame, lens_gpu, nhalos, 0, 0,
                  nbgridcells, nbgridcells, 0, 0, true, "module_potentialDerivatives_totalGradient_SOA_CPU_GPU");
}

void module_potentialDerivatives_totalGradient2_SOA_CPU_GPU(type_t* a, type_t* b, type_t* c, type_t* d,
                                                            const grid_param* frame, const Potential_SOA* lens_gpu,
                                                            int nhalos, type_t dx, type_t dy, int nx, int ny,
                                                            int istart, int jstart)
{
    device_launch(LT_MAP_HESSIAN, LT_MAP_HESSIAN, a, b, c, d, frame, lens_gpu, nhalos, dx, dy, nx, ny, istart, jstart,
                  false, "module_potentialDerivatives_totalGradient2_SOA_CPU_GPU");
}

void module_potentialDerivatives_totalGradient2_SOA_CPU_GPU(type_t* a, type_t* b, type_t* c, type_t* d,
                                                            const grid_param* frame, const Potential_SOA* lens_gpu,
                                                            int nbgridcells, int nhalos)
{
    device_launch(LT_MAP_HESSIAN, LT_MAP_HESSIAN, a, b, c, d, frame, lens_gpu, nhalos, 0, 0, nbgridcells, nbgridcells,
                  0, 0, true, "module_potentialDerivatives_totalGradient2_SOA_CPU_GPU");
}

void module_potentialDerivatives_Kmap_SOA_CPU_GPU(type_t* a, type_t* b, type_t* c, type_t* d, const grid_param* frame,
                                                  const Potential_SOA* lens_gpu, int nhalos, type_t dx, type_t dy,
                                                  int nx, int ny, int istart, int jstart)
{
    device_launch(LT_MAP_HESSIAN, LT_MAP_MASS, a, b, c, d, frame, lens_gpu, nhalos, dx, dy, nx, ny, istart, jstart,
                  false, "module_potentialDerivatives_Kmap_SOA_CPU_GPU");
}

// ---- delensing and chi^2 -----------------------------------------------------------------------------
// delense_barycenter_GPU, src/delense_GPU.cu:227-350 (nsets and nbgridcells from the run mode)
void delense_barycenter_GPU(point* image_pos, int* locimagesfound, int* numimg, const runmode_param* runmode,
                            const Potential_SOA* lens, const grid_param* frame, const int* nimages,
                            const galaxy* sources, double* gx, double* gy)
{
    (void)lens, (void)nimages;
    report(lt_delense_barycenter(reinterpret_cast<lt_point*>(image_pos), locimagesfound, numimg, runmode->nsets,
                                 runmode->nbgridcells, G(frame), reinterpret_cast<const lt_galaxy*>(sources), gx, gy),
           "delense_barycenter_GPU");
}

// mychi_bruteforce_SOA_GPU_grid_gradient_barycentersource, src/chi_CPU_barycenter.cpp:58-287
// (nsets, nbgridcells and nhalos from the run mode; one rank)
void mychi_bruteforce_SOA_GPU_grid_gradient_barycentersource(double* chi, int* error, runmode_param* runmode,
                                                             const Potential_SOA* lens, const grid_param* frame,
                                                             const int* nimages, galaxy* images)
{
    int st = lt_chi_barycentersource(chi, error, runmode->nbgridcells, runmode->nsets, runmode->nhalos, L(lens),
                                     G(frame), nimages, reinterpret_cast<const lt_galaxy*>(images));
    report(st, "mychi_bruteforce_SOA_GPU_grid_gradient_barycentersource");
}

// ---- CPU entry points of the reference, served by the GPU -------------------------------------------
// gradient_grid_CPU, src/grid_gradient_CPU.cpp:33-41 (square; istart/jstart are not used there)
void gradient_grid_CPU(type_t* gx, type_t* gy, const grid_param* frame, const Potential_SOA* lens, int nhalos,
                       int nbgridcells, int istart, int jstart)
{
    (void)istart, (void)jstart;
    const type_t dx = (frame->xmax - frame->xmin) / (nbgridcells - 1);
    const type_t dy = (frame->ymax - frame->ymin) / (nbgridcells - 1);
    report(lt_gradient_grid(gx, gy, G(frame), L(lens), nhalos, dx, dy, nbgridcells, nbgridcells, 0, 0, LT_PREC_FP64),
           "gradient_grid_CPU");
}

// src/grid_gradient_CPU.cpp:61-66
void gradient_grid_CPU(type_t* gx, type_t* gy, const grid_param* frame, const Potential_SOA* lens, int nhalos,
                       type_t dx, type_t dy, int nx, int ny, int istart, int jstart)
{
    report(lt_gradient_grid(gx, gy, G(frame), L(lens), nhalos, dx, dy, nx, ny, istart, jstart, LT_PREC_FP64),
           "gradient_grid_CPU");
}

namespace {
// matrix[] (AoS a, b, c, d) from the device planes
void hessian_aos(matrix* out, const grid_param* frame, const Potential_SOA* lens, int nhalos, type_t dx, type_t dy,
                 int nx, int ny, int istart, int jstart, const char* fn)
{
    const size_t n = (size_t)nx * ny;
    std::vector<double> a(n), b(n), c(n), d(n);
    int st = lt_gradient2_grid(a.data(), b.data(), c.data(), d.data(), G(frame), L(lens), nhalos, dx, dy, nx, ny,
                               istart, jstart, LT_PREC_FP64);
    if (st == LT_OK)
        for (size_t i = 0; i < n; ++i) out[i] = matrix{a[i], b[i], c[i], d[i]};
    report(st, fn);
}
}  // namespace

// gradient2_grid_CPU, src/grid_gradient2_CPU.cpp:30-52
void gradient2_grid_CPU(matrix* out, const grid_param* frame, const Potential_SOA* lens, int nhalos, int nbgridcells,
                        int istart, int jstart)
{
    const type_t dx = (frame->xmax - frame->xmin) / (nbgridcells - 1);
    const type_t dy = (frame->ymax - frame->ymin) / (nbgridcells - 1);
    hessian_aos(out, frame, lens, nhalos, dx, dy, nbgridcells, nbgridcells, istart, jstart, "gradient2_grid_CPU");
}

void gradient2_grid_CPU(matrix* out, const grid_param* frame, const Potential_SOA* lens, int nhalos, int nx, int ny,
                        int istart, int jstart)
{
    const type_t dx = (frame->xmax - frame->xmin) / (nx - 1);
    const type_t dy = (frame->ymax - frame->ymin) / (ny - 1);
    hessian_aos(out, frame, lens, nhalos, dx, dy, nx, ny, istart, jstart, "gradient2_grid_CPU");
}

void gradient2_grid_CPU(matrix* out, const grid_param* frame, const Potential_SOA* lens, int nhalos, type_t dx,
                        type_t dy, int nx, int ny, int istart, int jstart)
{
    hessian_aos(out, frame, lens, nhalos, dx, dy, nx, ny, istart, jstart, "gradient2_grid_CPU");
}

// amplif_grid_CPU, src/grid_amplif_CPU.cpp:38-46 (mode 5: kappa without dlsds, :50-83)
void amplif_grid_CPU(type_t* amplif, const cosmo_param* cosmo, const grid_param* frame, const Potential_SOA* lens,
                     int nhalos, int nbgridcells, int mode_amp, type_t z, int istart, int jstart)
{
    const type_t dx = (frame->xmax - frame->xmin) / (nbgridcells - 1);
    const type_t dy = (frame->ymax - frame->ymin) / (nbgridcells - 1);
    report(lt_map_grid(LT_MAP_AMPLI, mode_amp, amplif, nullptr, Cz(cosmo), G(frame), L(lens), nhalos, zl_of(lens), z, dx,
                       dy, nbgridcells, nbgridcells, istart, jstart, LT_PREC_FP64),
           "amplif_grid_CPU");
}

// module_potentialDerivatives_totalGradient_SOA, src/gradient.cpp:899-930 (one point per call: use
// lt_point_gradients for many)
point module_potentialDerivatives_totalGradient_SOA(const point* p, const Potential_SOA* lens, int nhalos)
{
    point g{0., 0.};
    report(lt_point_gradients(&g.x, &g.y, &p->x, &p->y, 1, L(lens), nhalos),
           "module_potentialDerivatives_totalGradient_SOA");
    return g;
}

// module_potentialDerivatives_totalGradient2_SOA, src/gradient2.cpp:226-254 (a 1 x 1 grid at the point)
matrix module_potentialDerivatives_totalGradient2_SOA(const point* p, const Potential_SOA* lens, int nhalos)
{
    matrix m{0., 0., 0., 0.};
    lt_grid_param f{p->x, p->x, p->y, p->y, 0, 0, 0};
    report(lt_gradient2_grid(&m.a, &m.b, &m.c, &m.d, &f, L(lens), nhalos, 0.0, 0.0, 1, 1, 0, 0, LT_PREC_FP64),
           "module_potentialDerivatives_totalGradient2_SOA");
    return m;
}

// delense_barycenter, src/delense_CPU.cpp:6-138
void delense_barycenter(point* image_pos, int* locimagesfound, int* numimg, const runmode_param* runmode,
                        const Potential_SOA* lens, const grid_param* frame, const int* nimages, const galaxy* sources,
                        double* gx, double* gy)
{
    delense_barycenter_GPU(image_pos, locimagesfound, numimg, runmode, lens, frame, nimages, sources, gx, gy);
}

// chi_computation, src/chi_computation.cpp:10-195
void chi_computation(double* chi, int* images_found, int* numimagesfound, point* imagesposition, const int* nimages,
                     const galaxy* images, const int nsets)
{
    report(lt_chi_computation(chi, images_found, numimagesfound, reinterpret_cast<const lt_point*>(imagesposition),
                              nimages, reinterpret_cast<const lt_galaxy*>(images), nsets),
           "chi_computation");
}

// mychi_bruteforce_SOA_CPU_grid_gradient_barycentersource, src/chi_CPU_barycenter.cpp:58-287
void mychi_bruteforce_SOA_CPU_grid_gradient_barycentersource(double* chi, int* error, runmode_param* runmode,
                                                             const Potential_SOA* lens, const grid_param* frame,
                                                             const int* nimages, galaxy* images)
{
    mychi_bruteforce_SOA_GPU_grid_gradient_barycentersource(chi, error, runmode, lens, frame, nimages, images);
}

// ---- FITS ------------------------------------------------------------------------------------------
// module_writeFits_Image, src/module_writeFits.cpp:113-200
int module_writeFits_Image(char* filename, double* ima, int nx, int ny, double xmin, double xmax, double ymin,
                           double ymax)
{
    const int st = lt_fits_write(filename, ima, nx, ny, xmin, xmax, ymin, ymax, 0, 0., 0.);
    report(st, "module_writeFits_Image");
    return st;
}

// module_writeFits_ImageAbsoluteCoordinates, src/module_writeFits.cpp:225-335
int module_writeFits_ImageAbsoluteCoordinates(char* filename, double* ima, int nx, int ny, double xmin, double xmax,
                                              double ymin, double ymax, double ra, double dec)
{
    const int st = lt_fits_write(filename, ima, nx, ny, xmin, xmax, ymin, ymax, 1, ra, dec);
    report(st, "module_writeFits_ImageAbsoluteCoordinates");
    return st;
}

// module_writeFits, src/module_writeFits.cpp:41-61 (<path>/<name>_<ii>.fits) and :63-80 (<path>/<name>.fits):
// the absolute header unless ra == dec == 0
void module_writeFits(std::string path, std::string filename, int ii, type_t* map, const runmode_param* runmode,
                      int size, const grid_param* frame, type_t ra, type_t dec)
{
    (void)runmode;
    const std::string file = path + "/" + filename + "_" + std::to_string(ii) + ".fits";
    report(lt_fits_write(file.c_str(), map, size, size, frame->xmin, frame->xmax, frame->ymin, frame->ymax,
                         !(ra == 0. && dec == 0.), ra, dec),
           "module_writeFits");
}

void module_writeFits(std::string path, std::string filename, type_t* map, const runmode_param* runmode, int size,
                      const grid_param* frame, type_t ra, type_t dec)
{
    (void)runmode;
    const std::string file = path + "/" + filename + ".fits";
    report(lt_fits_write(file.c_str(), map, size, size, frame->xmin, frame->xmax, frame->ymin, frame->ymax,
                         !(ra == 0. && dec == 0.), ra, dec),
           "module_writeFits");
}
