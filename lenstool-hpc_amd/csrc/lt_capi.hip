// lt_capi.hip -- implementation of the C-ABI (include/lenstool_hip/lt_capi.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <tuple>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "lenstool_hip/lt_capi.h"
#include "lt_internal.h"
#include "lt_kernels.h"

using namespace lt;

namespace {

thread_local std::string g_err;
thread_local int g_device = -1;

}  // namespace

namespace lt {

int set_error(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

int hip_error(hipError_t e, const char* where)
{
    std::string m = std::string(where) + ": " + hipGetErrorString(e);
    return set_error(LT_ERR_HIP, m);
}

int current_device(int* dev)
{
    if (g_device < 0) {
        int n = 0;
        hipError_t e = hipGetDeviceCount(&n);
        if (e != hipSuccess || n <= 0) return set_error(LT_ERR_NO_DEVICE, "no GPU visible to the HIP runtime");
        LT_HIP_TRY(hipGetDevice(&g_device));
    }
    if (g_device >= 64) return set_error(LT_ERR_NO_DEVICE, "device index out of range");
    LT_HIP_TRY(hipSetDevice(g_device));
    *dev = g_device;
    return LT_OK;
}

}  // namespace lt

namespace {

int fail(int code, const std::string& msg) { return set_error(code, msg); }
int hip_fail(hipError_t e, const char* where) { return hip_error(e, where); }
#define HIP_TRY(expr) LT_HIP_TRY(expr)

// Per-device state of the synchronous host-array entry points: a stream and grow-only buffers.
struct DeviceState {
    std::mutex mu;
    hipStream_t stream = nullptr;
    void* halos = nullptr;
    size_t halos_bytes = 0;
    void* planes = nullptr;
    size_t planes_bytes = 0;
};

constexpr int kMaxDevices = 64;
DeviceState g_state[kMaxDevices];

int ensure(void** buf, size_t* have, size_t need)
{
    if (need <= *have) return LT_OK;
    if (*buf) HIP_TRY(hipFree(*buf));
    *buf = nullptr;
    *have = 0;
    size_t n = need + need / 4;
    hipError_t e = hipMalloc(buf, n);
    if (e != hipSuccess) {
        *buf = nullptr;
        (void)hipGetLastError();
        return fail(LT_ERR_NOMEM, "device allocation of " + std::to_string(n) + " bytes failed");
    }
    *have = n;
    return LT_OK;
}

// Pixels per thread chosen by lt_autotune, per (kernel family, precision, halo-count and pixel-count
// octave); 2 where nothing was tuned. Every px gives bit-identical results: only speed differs.
struct TuneKey {
    int fam, prec, hoct, noct;
    bool operator<(const TuneKey& o) const
    {
        return std::tie(fam, prec, hoct, noct) < std::tie(o.fam, o.prec, o.hoct, o.noct);
    }
};
std::mutex g_tune_mu;
std::map<TuneKey, int> g_tuned;

int octave(long long v)
{
    int k = 0;
    while (v > 1) {
        v >>= 1;
        ++k;
    }
    return k;
}

TuneKey tune_key(int kind, int precision, int nhalos, long long npix)
{
    const int fam = (kind == LT_MAP_GRADIENT || kind == LT_MAP_DPL) ? 0 : kind == LT_MAP_POT ? 2 : 1;
    // every family has an FP64 and an FP32C kernel, each tuned on its own
    return TuneKey{fam, precision, octave(nhalos), octave(npix)};
}

int default_px(int kind, int precision, int nhalos, long long npix)
{
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tuned.find(tune_key(kind, precision, nhalos, npix));
    return it == g_tuned.end() ? 2 : it->second;
}

struct DevModel {
    HaloRec* recs = nullptr;
    HaloRec32* recs32 = nullptr;
    Run* runs = nullptr;
    int nruns = 0;
};

// Launch the kernel that produces map `kind` from a device model.
int launch_map(const DevModel& m, int kind, int mode, double* const outs[5], double r, double mult,
               const GridDesc& g, int precision, int px, hipStream_t s, int nhalos)
{
    OutDesc o{};
    o.p0 = outs[0];
    o.p1 = outs[1];
    o.p2 = outs[2];
    o.p3 = outs[3];
    o.p4 = outs[4];
    o.kind = kind;
    o.mode = mode;
    o.r = r;
    o.mult = mult;
    if (px <= 0) px = default_px(kind, precision, nhalos, (long long)g.nx * g.ny);
    const bool hessian_family = kind != LT_MAP_GRADIENT && kind != LT_MAP_DPL && kind != LT_MAP_POT;
    if (px == LT_PX_HALO_LANES && !(hessian_family && precision == LT_PREC_FP32C))
        return fail(LT_ERR_INVALID, "px = LT_PX_HALO_LANES applies to FP32C Hessian-derived maps only");
    if (px != 1 && px != 2 && px != 4 && px != LT_PX_HALO_LANES)
        return fail(LT_ERR_INVALID, "px must be 0, 1, 2, 4 or LT_PX_HALO_LANES");
    if (precision != LT_PREC_FP64 && precision != LT_PREC_FP32C)
        return fail(LT_ERR_INVALID, "precision must be LT_PREC_FP64 or LT_PREC_FP32C");
    hipError_t e;
    switch (kind) {
    case LT_MAP_GRADIENT:
    case LT_MAP_DPL:
        if (precision == LT_PREC_FP32C) e = launch_gradient32(m.recs, m.recs32, m.runs, m.nruns, g, o, px, s);
        else e = launch_gradient64(m.recs, m.runs, m.nruns, g, o, px, s);
        break;
    case LT_MAP_POT:
        if (precision == LT_PREC_FP32C) e = launch_potential32(m.recs, m.recs32, m.runs, m.nruns, g, o, px, s);
        else e = launch_potential64(m.recs, m.runs, m.nruns, g, o, px, s);
        break;
    default:
        // FP32C covers the Hessian-derived and the gradient maps
        if (precision == LT_PREC_FP32C) e = launch_hessian32(m.recs, m.recs32, m.runs, m.nruns, g, o, px, s);
        else e = launch_hessian64(m.recs, m.runs, m.nruns, g, o, px, s);
        break;
    }
    if (e != hipSuccess) return hip_fail(e, "kernel launch");
    return LT_OK;
}

int planes_for_kind(int kind)
{
    switch (kind) {
    case LT_MAP_HESSIAN: return 4;
    case LT_MAP_FIELDS: return 5;
    case LT_MAP_GRADIENT:
    case LT_MAP_DPL: return 2;
    default: return 1;
    }
}

int check_kind_mode(int kind, int mode)
{
    switch (kind) {
    case LT_MAP_AMPLI: if (mode >= 1 && mode <= 6) return LT_OK; break;
    case LT_MAP_SHEAR: if (mode >= 1 && mode <= 4) return LT_OK; break;
    case LT_MAP_MASS: if (mode >= 1 && mode <= 4) return LT_OK; break;
    case LT_MAP_POT: if (mode >= 1 && mode <= 3) return LT_OK; break;
    case LT_MAP_DPL:
    case LT_MAP_FIELDS:
    case LT_MAP_HESSIAN:
    case LT_MAP_GRADIENT: return LT_OK;
    default: return fail(LT_ERR_INVALID, "unknown map kind " + std::to_string(kind));
    }
    return fail(LT_ERR_INVALID, "mode " + std::to_string(mode) + " not valid for map kind " + std::to_string(kind));
}

int check_grid(const lt_grid_param* frame, double dx, double dy, int nx, int ny)
{
    if (!frame) return fail(LT_ERR_INVALID, "frame is NULL");
    if (nx <= 0 || ny <= 0) return fail(LT_ERR_INVALID, "grid size must be positive");
    if (!std::isfinite(frame->xmin) || !std::isfinite(frame->ymin) || !std::isfinite(dx) || !std::isfinite(dy))
        return fail(LT_ERR_INVALID, "non-finite frame or pixel size");
    return LT_OK;
}

// Synchronous host-array path: pack, upload, launch, download.
int host_map(int kind, int mode, Use use, double* const host_out[5], const lt_grid_param* frame,
             const lt_potential_soa* lens, int nhalos, double r, double mult, double dx, double dy, int nx, int ny,
             int istart, int jstart, int precision)
{
    int st = check_kind_mode(kind, mode);
    if (st) return st;
    if ((st = check_grid(frame, dx, dy, nx, ny))) return st;
    int np = planes_for_kind(kind);
    bool any = false;
    for (int k = 0; k < np; ++k) any |= host_out[k] != nullptr;
    if (!any) return fail(LT_ERR_INVALID, "no output array given");
    if ((kind == LT_MAP_GRADIENT || kind == LT_MAP_DPL) && (!host_out[0] || !host_out[1]))
        return fail(LT_ERR_INVALID, "gradient output needs both x and y arrays");
    if (kind != LT_MAP_HESSIAN && kind != LT_MAP_FIELDS && kind != LT_MAP_GRADIENT && kind != LT_MAP_DPL && !host_out[0])
        return fail(LT_ERR_INVALID, "map output array is NULL");

    std::vector<HaloRec> recs;
    std::vector<HaloRec32> recs32;
    std::vector<Run> runs;
    std::string err;
    if ((st = pack_model(lens, nhalos, use, recs, recs32, runs, err))) return fail(st, err);

    int dev;
    if ((st = current_device(&dev))) return st;
    DeviceState& S = g_state[dev];
    std::lock_guard<std::mutex> lock(S.mu);
    if (!S.stream) HIP_TRY(hipStreamCreateWithFlags(&S.stream, hipStreamNonBlocking));

    const size_t n = (size_t)nx * (size_t)ny;
    const size_t rec_bytes = recs.size() * sizeof(HaloRec);
    const size_t rec32_bytes = recs32.size() * sizeof(HaloRec32);
    const size_t run_bytes = runs.size() * sizeof(Run);
    const size_t halo_need = rec_bytes + rec32_bytes + run_bytes + 256;
    if ((st = ensure(&S.halos, &S.halos_bytes, halo_need))) return st;
    if ((st = ensure(&S.planes, &S.planes_bytes, n * np * sizeof(double)))) return st;

    char* base = (char*)S.halos;
    DevModel m;
    m.recs = (HaloRec*)base;
    m.recs32 = (HaloRec32*)(base + rec_bytes);
    m.runs = (Run*)(base + ((rec_bytes + rec32_bytes + 63) / 64) * 64);
    m.nruns = (int)runs.size();
    if (rec_bytes) HIP_TRY(hipMemcpyAsync(m.recs, recs.data(), rec_bytes, hipMemcpyHostToDevice, S.stream));
    if (rec32_bytes) HIP_TRY(hipMemcpyAsync(m.recs32, recs32.data(), rec32_bytes, hipMemcpyHostToDevice, S.stream));
    if (run_bytes) HIP_TRY(hipMemcpyAsync(m.runs, runs.data(), run_bytes, hipMemcpyHostToDevice, S.stream));

    double* dev_out[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    for (int k = 0; k < np; ++k)
        if (host_out[k]) dev_out[k] = (double*)S.planes + k * n;
    GridDesc g{};
    g.xmin = frame->xmin;
    g.ymin = frame->ymin;
    g.dx = dx;
    g.dy = dy;
    g.nx = nx;
    g.ny = ny;
    g.istart = istart;
    g.jstart = jstart;
    g.pitch = nx;
    g.nsamples = 1;
    if ((st = launch_map(m, kind, mode, dev_out, r, mult, g, precision, 0, S.stream, nhalos))) return st;
    for (int k = 0; k < np; ++k)
        if (host_out[k])
            HIP_TRY(hipMemcpyAsync(host_out[k], dev_out[k], n * sizeof(double), hipMemcpyDeviceToHost, S.stream));
    HIP_TRY(hipStreamSynchronize(S.stream));
    return LT_OK;
}

}  // namespace

// =================================================================================================
// exported C-ABI
// =================================================================================================
extern "C" {

const char* lt_last_error(void) { return g_err.c_str(); }

int lt_version(void) { return 100; }

int lt_device_count(int* count)
{
    if (!count) return fail(LT_ERR_INVALID, "count is NULL");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) n = 0;
    *count = n;
    return LT_OK;
}

int lt_init(int device)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(LT_ERR_NO_DEVICE, "no GPU visible to the HIP runtime");
    if (device < 0 || device >= n || device >= kMaxDevices) return fail(LT_ERR_INVALID, "device index out of range");
    HIP_TRY(hipSetDevice(device));
    g_device = device;
    return critic_noalloc_setup(device, nullptr);
}

int lt_sync(void)
{
    int dev, st;
    if ((st = current_device(&dev))) return st;
    if (g_state[dev].stream) HIP_TRY(hipStreamSynchronize(g_state[dev].stream));
    return LT_OK;
}

int lt_dev_alloc(void** ptr, long long bytes)
{
    if (!ptr || bytes < 0) return fail(LT_ERR_INVALID, "bad allocation request");
    *ptr = nullptr;
    int dev, st;
    if ((st = current_device(&dev))) return st;
    hipError_t e = hipMalloc(ptr, (size_t)(bytes > 0 ? bytes : 1));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *ptr = nullptr;
        return fail(LT_ERR_NOMEM, "device allocation of " + std::to_string(bytes) + " bytes failed");
    }
    return LT_OK;
}

int lt_dev_free(void* ptr)
{
    if (ptr) HIP_TRY(hipFree(ptr));
    return LT_OK;
}

int lt_memcpy_h2d(void* dst, const void* src, long long bytes, void* stream)
{
    if (bytes < 0 || (bytes && (!dst || !src))) return fail(LT_ERR_INVALID, "bad copy");
    if (!bytes) return LT_OK;
    if (stream) HIP_TRY(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyHostToDevice, (hipStream_t)stream));
    else HIP_TRY(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyHostToDevice));
    return LT_OK;
}

int lt_memcpy_d2h(void* dst, const void* src, long long bytes, void* stream)
{
    if (bytes < 0 || (bytes && (!dst || !src))) return fail(LT_ERR_INVALID, "bad copy");
    if (!bytes) return LT_OK;
    if (stream) HIP_TRY(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDeviceToHost, (hipStream_t)stream));
    else HIP_TRY(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyDeviceToHost));
    return LT_OK;
}

int lt_memset_d(void* dst, int value, long long bytes, void* stream)
{
    if (bytes < 0 || (bytes && !dst)) return fail(LT_ERR_INVALID, "bad memset");
    if (!bytes) return LT_OK;
    // a fill kernel rather than hipMemsetAsync, so that a captured graph holds no memset node
    hipError_t e = launch_fill_bytes(dst, (unsigned char)value, (size_t)bytes, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "fill kernel");
    if (!stream) HIP_TRY(hipStreamSynchronize(nullptr));
    return LT_OK;
}

int lt_memcpy_d2d(void* dst, const void* src, long long bytes, void* stream)
{
    if (bytes < 0 || (bytes && (!dst || !src))) return fail(LT_ERR_INVALID, "bad copy");
    if (!bytes) return LT_OK;
    hipError_t e = launch_copy(dst, src, (size_t)bytes, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "copy kernel");
    if (!stream) HIP_TRY(hipStreamSynchronize(nullptr));
    return LT_OK;
}

int lt_stream_create(void** stream)
{
    if (!stream) return fail(LT_ERR_INVALID, "stream is NULL");
    int dev, st;
    if ((st = current_device(&dev))) return st;
    hipStream_t s;
    HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = s;
    return LT_OK;
}

int lt_stream_destroy(void* stream)
{
    if (stream) HIP_TRY(hipStreamDestroy((hipStream_t)stream));
    return LT_OK;
}

int lt_stream_sync(void* stream)
{
    HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
    return LT_OK;
}

int lt_event_create(void** event)
{
    if (!event) return fail(LT_ERR_INVALID, "event is NULL");
    int dev, st;
    if ((st = current_device(&dev))) return st;
    hipEvent_t e;
    HIP_TRY(hipEventCreate(&e));
    *event = e;
    return LT_OK;
}

int lt_event_destroy(void* event)
{
    if (event) HIP_TRY(hipEventDestroy((hipEvent_t)event));
    return LT_OK;
}

int lt_event_record(void* event, void* stream)
{
    HIP_TRY(hipEventRecord((hipEvent_t)event, (hipStream_t)stream));
    return LT_OK;
}

int lt_event_sync(void* event)
{
    HIP_TRY(hipEventSynchronize((hipEvent_t)event));
    return LT_OK;
}

int lt_event_elapsed(float* ms, void* start, void* stop)
{
    if (!ms) return fail(LT_ERR_INVALID, "ms is NULL");
    HIP_TRY(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
    return LT_OK;
}

double lt_cosmo_observer_object(double z, const lt_cosmo_param* c) { return c ? cosmo_observer_object(z, *c) : NAN; }

double lt_cosmo_object_object(double z1, double z2, const lt_cosmo_param* c)
{
    return c ? cosmo_object_object(z1, z2, *c) : NAN;
}

double lt_cosmo_dlsds(double zl, double zs, const lt_cosmo_param* c)
{
    if (!c) return NAN;
    if (zl >= zs) return 0.0;
    return cosmo_object_object(zl, zs, *c) / cosmo_observer_object(zs, *c);
}

int lt_gradient_grid(double* gx, double* gy, const lt_grid_param* frame, const lt_potential_soa* lens, int nhalos,
                     double dx, double dy, int nx, int ny, int istart, int jstart, int precision)
{
    double* outs[5] = {gx, gy, nullptr, nullptr, nullptr};
    return host_map(LT_MAP_GRADIENT, 0, Use::Gradient, outs, frame, lens, nhalos, 1.0, 1.0, dx, dy, nx, ny, istart,
                    jstart, precision);
}

int lt_gradient2_grid(double* a, double* b, double* c, double* d, const lt_grid_param* frame,
                      const lt_potential_soa* lens, int nhalos, double dx, double dy, int nx, int ny, int istart,
                      int jstart, int precision)
{
    double* outs[5] = {a, b, c, d, nullptr};
    return host_map(LT_MAP_HESSIAN, 0, Use::Hessian, outs, frame, lens, nhalos, 1.0, 1.0, dx, dy, nx, ny, istart,
                    jstart, precision);
}

int lt_kmap_grid(double* a, double* b, double* c, double* d, const lt_grid_param* frame, const lt_potential_soa* lens,
                 int nhalos, double dx, double dy, int nx, int ny, int istart, int jstart, int precision)
{
    double* outs[5] = {a, b, c, d, nullptr};
    return host_map(LT_MAP_HESSIAN, 0, Use::Kmap, outs, frame, lens, nhalos, 1.0, 1.0, dx, dy, nx, ny, istart,
                    jstart, precision);
}

int lt_potential_grid(double* phi, const lt_grid_param* frame, const lt_potential_soa* lens, int nhalos, double dx,
                      double dy, int nx, int ny, int istart, int jstart)
{
    double* outs[5] = {phi, nullptr, nullptr, nullptr, nullptr};
    return host_map(LT_MAP_POT, 3, Use::Potential, outs, frame, lens, nhalos, 1.0, 1.0, dx, dy, nx, ny, istart,
                    jstart, LT_PREC_FP64);
}

int lt_map_factors(int kind, int mode, const lt_cosmo_param* cosmo, double z_lens, double z_source, double dx,
                   double dy, double* r_out, double* mult_out)
{
    int st = check_kind_mode(kind, mode);
    if (st) return st;
    if (!cosmo || !r_out || !mult_out) return fail(LT_ERR_INVALID, "NULL argument");
    // distances as the reference map wrappers form them (src/grid_map_ampli_GPU.cu:110-112,
    // src/grid_map_mass_GPU.cu:101-108)
    const double dl0s = cosmo_object_object(z_lens, z_source, *cosmo);
    const double dos = cosmo_observer_object(z_source, *cosmo);
    const double dol = cosmo_observer_object(z_lens, *cosmo);
    const double r = dl0s / dos;
    double mult = 1.0;
    const double h = cosmo->h;
    if (kind == LT_MAP_MASS) {
        switch (mode) {
        case 2: mult = LT_CH4PIG * h / dol / dl0s * dos; break;           // g/cm^2
        case 3: mult = LT_MCRIT12 / h * dx * dy * dol / dl0s * dos; break;  // 1e12 Msun/pixel
        case 4: mult = LT_CH0_4PIG * h / dol / dl0s * dos; break;           // 1e12 Msun/kpc^2
        default: mult = 1.0; break;
        }
    } else if (kind == LT_MAP_POT) {
        mult = mode == 1 ? r : mode == 2 ? LT_CH2PIG * h / dol : 1.0;
    }
    *r_out = r;
    *mult_out = mult;
    return LT_OK;
}

int lt_map_grid(int kind, int mode, double* map0, double* map1, const lt_cosmo_param* cosmo,
                const lt_grid_param* frame, const lt_potential_soa* lens, int nhalos, double z_lens, double z_source,
                double dx, double dy, int nx, int ny, int istart, int jstart, int precision)
{
    int st = check_kind_mode(kind, mode);
    if (st) return st;
    if (kind == LT_MAP_FIELDS || kind == LT_MAP_HESSIAN || kind == LT_MAP_GRADIENT)
        return fail(LT_ERR_INVALID, "lt_map_grid takes the reference map kinds AMPLI, SHEAR, MASS, POT, DPL");
    if (!cosmo) return fail(LT_ERR_INVALID, "cosmo is NULL");
    if (kind == LT_MAP_MASS && z_lens == 0.0) {
        if (!lens || !lens->z || nhalos <= 0) return fail(LT_ERR_INVALID, "mass map: z_lens = 0 needs lens->z[0]");
        z_lens = lens->z[0];
    }
    double r = 0, mult = 1;
    st = lt_map_factors(kind, mode, cosmo, z_lens, z_source, dx, dy, &r, &mult);
    if (st) return st;
    double* outs[5] = {map0, kind == LT_MAP_DPL ? map1 : nullptr, nullptr, nullptr, nullptr};
    Use use = kind == LT_MAP_DPL ? Use::Gradient : kind == LT_MAP_POT ? Use::Potential
              : kind == LT_MAP_MASS ? Use::Kmap : Use::Hessian;
    return host_map(kind, mode, use, outs, frame, lens, nhalos, r, mult, dx, dy, nx, ny, istart, jstart, precision);
}

int lt_map_fields(double* kappa, double* gamma1, double* gamma2, double* det_a, double* mu, const lt_cosmo_param* cosmo,
                  const lt_grid_param* frame, const lt_potential_soa* lens, int nhalos, double z_lens, double z_source,
                  double dx, double dy, int nx, int ny, int istart, int jstart, int precision)
{
    if (!cosmo) return fail(LT_ERR_INVALID, "cosmo is NULL");
    const double r = z_lens >= z_source ? 0.0
                                         : cosmo_object_object(z_lens, z_source, *cosmo) /
                                               cosmo_observer_object(z_source, *cosmo);
    double* outs[5] = {kappa, gamma1, gamma2, det_a, mu};
    return host_map(LT_MAP_FIELDS, 0, Use::Hessian, outs, frame, lens, nhalos, r, 1.0, dx, dy, nx, ny, istart, jstart,
                    precision);
}

// ---- resident models ------------------------------------------------------------------------------
}  // extern "C"

struct lt_model_s {
    int device = 0;
    Use use = Use::Hessian;
    int nhalos = 0;
    int nsamples = 1;
    std::vector<Run> runs_host;
    std::vector<int> types;
    DevModel dm;
    size_t bytes = 0;
    size_t rec_bytes = 0, rec32_bytes = 0;  // all samples
    void* base = nullptr;                   // device: records [S*H], fp32 records [S*H], runs
    size_t off_runs = 0;                    // offset of the runs in `base` and in `staging`
    void* staging = nullptr;                // pinned host image of base (records and runs)
    hipEvent_t staged = nullptr;            // completes when the last upload has read `staging`
    bool staged_pending = false;
};

static Use use_from_kind(int use)
{
    switch (use) {
    case LT_MAP_GRADIENT:
    case LT_MAP_DPL: return Use::Gradient;
    case LT_MAP_MASS: return Use::Kmap;
    case LT_MAP_POT: return Use::Potential;
    default: return Use::Hessian;
    }
}

// the same run structure (the circular-SIS flag may differ; a batch's run is circular only when it is
// in every sample)
static bool same_runs(const std::vector<Run>& a, const std::vector<Run>& b)
{
    if (a.size() != b.size()) return false;
    for (size_t i = 0; i < a.size(); ++i)
        if (a[i].type != b[i].type || a[i].start != b[i].start || a[i].count != b[i].count || a[i].weight != b[i].weight)
            return false;
    return true;
}

// Pack every sample into the pinned staging image (after the previous upload has read it).
static int stage_samples(lt_model_s* m, const lt_potential_soa* lenses, int nsamples)
{
    if (!lenses) return fail(LT_ERR_INVALID, "lens models are NULL");
    if (nsamples != m->nsamples) return fail(LT_ERR_INVALID, "sample count differs from the model's");
    if (m->staged_pending) {
        HIP_TRY(hipEventSynchronize(m->staged));
        m->staged_pending = false;
    }
    const int H = m->nhalos;
    std::vector<HaloRec> recs;
    std::vector<HaloRec32> recs32;
    std::vector<Run> runs;
    std::string err;
    char* st = (char*)m->staging;
    std::vector<Run> batch_runs = m->runs_host;
    for (int s = 0; s < nsamples; ++s) {
        const lt_potential_soa& L = lenses[s];
        for (int i = 0; i < H; ++i)
            if (!L.type || L.type[i] != m->types[i])
                return fail(LT_ERR_INVALID, "sample " + std::to_string(s) + ": halo types differ from the model's");
        int rc = pack_model(&L, H, m->use, recs, recs32, runs, err);
        if (rc) return fail(rc, "sample " + std::to_string(s) + ": " + err);
        if (!same_runs(runs, m->runs_host))
            return fail(LT_ERR_INVALID, "sample " + std::to_string(s) + ": run structure (types / dlsds) differs");
        for (size_t r = 0; r < runs.size(); ++r) batch_runs[r].circ = s ? (batch_runs[r].circ & runs[r].circ) : runs[r].circ;
        if (H) {
            std::memcpy(st + (size_t)s * H * sizeof(HaloRec), recs.data(), (size_t)H * sizeof(HaloRec));
            std::memcpy(st + m->rec_bytes + (size_t)s * H * sizeof(HaloRec32), recs32.data(),
                        (size_t)H * sizeof(HaloRec32));
        }
    }
    if (!batch_runs.empty()) std::memcpy(st + m->off_runs, batch_runs.data(), batch_runs.size() * sizeof(Run));
    return LT_OK;
}

// Queue the copy staging -> device on `s`.
static int upload_staged(lt_model_s* m, hipStream_t s)
{
    const size_t n = m->off_runs + m->runs_host.size() * sizeof(Run);
    if (!n) return LT_OK;
    HIP_TRY(hipMemcpyAsync(m->base, m->staging, n, hipMemcpyHostToDevice, s));
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    HIP_TRY(hipStreamIsCapturing(s, &cs));
    if (cs == hipStreamCaptureStatusNone) {
        HIP_TRY(hipEventRecord(m->staged, s));
        m->staged_pending = true;
    }
    return LT_OK;
}

static void free_model(lt_model_s* m)
{
    if (m->staged) (void)hipEventDestroy(m->staged);
    if (m->staging) (void)hipHostFree(m->staging);
    if (m->base) (void)hipFree(m->base);
    delete m;
}

extern "C" {

int lt_model_create_batch(lt_model* out, const lt_potential_soa* lenses, int nsamples, int nhalos, int use)
{
    if (!out) return fail(LT_ERR_INVALID, "out is NULL");
    *out = nullptr;
    if (!lenses || nsamples < 1) return fail(LT_ERR_INVALID, "need at least one lens model");
    if (nhalos < 0) return fail(LT_ERR_INVALID, "negative halo count");
    int st = check_kind_mode(use, use == LT_MAP_AMPLI || use == LT_MAP_SHEAR || use == LT_MAP_MASS ||
                                      use == LT_MAP_POT ? 1 : 0);
    if (st) return st;
    std::vector<HaloRec> recs;
    std::vector<HaloRec32> recs32;
    std::vector<Run> runs;
    std::string err;
    Use u = use_from_kind(use);
    if ((st = pack_model(&lenses[0], nhalos, u, recs, recs32, runs, err))) return fail(st, "sample 0: " + err);
    int dev;
    if ((st = current_device(&dev))) return st;
    lt_model_s* m = new lt_model_s();
    m->device = dev;
    m->use = u;
    m->nhalos = nhalos;
    m->nsamples = nsamples;
    m->runs_host = runs;
    for (int i = 0; i < nhalos; ++i) m->types.push_back(lenses[0].type[i]);
    m->rec_bytes = (size_t)nsamples * nhalos * sizeof(HaloRec);
    m->rec32_bytes = (size_t)nsamples * nhalos * sizeof(HaloRec32);
    const size_t off_runs = ((m->rec_bytes + m->rec32_bytes + 63) / 64) * 64;
    m->off_runs = off_runs;
    m->bytes = off_runs + runs.size() * sizeof(Run) + 64;
    hipError_t e = hipMalloc(&m->base, m->bytes);
    if (e == hipSuccess) e = hipHostMalloc(&m->staging, m->bytes);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&m->staged, hipEventDisableTiming);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        free_model(m);
        return fail(LT_ERR_NOMEM, "model allocation failed");
    }
    char* b = (char*)m->base;
    m->dm.recs = (HaloRec*)b;
    m->dm.recs32 = (HaloRec32*)(b + m->rec_bytes);
    m->dm.runs = (Run*)(b + off_runs);
    m->dm.nruns = (int)runs.size();
    if ((st = stage_samples(m, lenses, nsamples))) {
        free_model(m);
        return st;
    }
    e = hipMemcpy(m->base, m->staging, off_runs + runs.size() * sizeof(Run), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        free_model(m);
        return hip_fail(e, "model upload");
    }
    *out = m;
    return LT_OK;
}

int lt_model_create(lt_model* out, const lt_potential_soa* lens, int nhalos, int use)
{
    return lt_model_create_batch(out, lens, 1, nhalos, use);
}

int lt_model_update_batch(lt_model m, const lt_potential_soa* lenses, int nsamples, void* stream)
{
    if (!m) return fail(LT_ERR_NOT_FOUND, "model is NULL");
    HIP_TRY(hipSetDevice(m->device));
    int st = stage_samples(m, lenses, nsamples);
    if (st) return st;
    return upload_staged(m, (hipStream_t)stream);
}

int lt_model_update(lt_model m, const lt_potential_soa* lens, int nhalos, void* stream)
{
    if (!m) return fail(LT_ERR_NOT_FOUND, "model is NULL");
    if (nhalos != m->nhalos) return fail(LT_ERR_INVALID, "halo count differs from the model's");
    if (m->nsamples != 1) return fail(LT_ERR_INVALID, "batched model: use lt_model_update_batch");
    return lt_model_update_batch(m, lens, 1, stream);
}

int lt_model_stage(lt_model m, const lt_potential_soa* lenses, int nsamples)
{
    if (!m) return fail(LT_ERR_NOT_FOUND, "model is NULL");
    HIP_TRY(hipSetDevice(m->device));
    return stage_samples(m, lenses, nsamples);
}

int lt_model_upload(lt_model m, void* stream)
{
    if (!m) return fail(LT_ERR_NOT_FOUND, "model is NULL");
    HIP_TRY(hipSetDevice(m->device));
    return upload_staged(m, (hipStream_t)stream);
}

int lt_model_destroy(lt_model m)
{
    if (!m) return LT_OK;
    (void)hipSetDevice(m->device);
    if (m->staged_pending) (void)hipEventSynchronize(m->staged);
    free_model(m);
    return LT_OK;
}

int lt_model_bytes(lt_model m, long long* bytes)
{
    if (!m || !bytes) return fail(LT_ERR_INVALID, "NULL argument");
    *bytes = (long long)m->bytes;
    return LT_OK;
}

int lt_map_device(lt_model m, int kind, int mode, double* out0, double* out1, double* out2, double* out3, double* out4,
                  double r, double mult, const lt_grid_desc* grid, int precision, int px, void* stream)
{
    if (!m) return fail(LT_ERR_NOT_FOUND, "model is NULL");
    if (!grid) return fail(LT_ERR_INVALID, "grid is NULL");
    int st = check_kind_mode(kind, mode);
    if (st) return st;
    Use need = kind == LT_MAP_DPL || kind == LT_MAP_GRADIENT ? Use::Gradient
               : kind == LT_MAP_POT ? Use::Potential
               : kind == LT_MAP_MASS ? Use::Kmap : Use::Hessian;
    if (need != m->use && !(need == Use::Hessian && m->use == Use::Kmap))
        return fail(LT_ERR_INVALID, "model was created for another kind of map");
    if (grid->nx <= 0 || grid->ny <= 0 || grid->pitch < grid->nx) return fail(LT_ERR_INVALID, "bad grid");
    double* outs[5] = {out0, out1, out2, out3, out4};
    if ((kind == LT_MAP_GRADIENT || kind == LT_MAP_DPL) && (!out0 || !out1))
        return fail(LT_ERR_INVALID, "gradient output needs two planes");
    if (kind != LT_MAP_HESSIAN && kind != LT_MAP_FIELDS && !out0) return fail(LT_ERR_INVALID, "out0 is NULL");
    GridDesc g{};
    g.xmin = grid->xmin;
    g.ymin = grid->ymin;
    g.dx = grid->dx;
    g.dy = grid->dy;
    g.nx = grid->nx;
    g.ny = grid->ny;
    g.istart = grid->istart;
    g.jstart = grid->jstart;
    g.pitch = grid->pitch;
    g.nsamples = m->nsamples;
    g.halo_stride = m->nhalos;
    g.plane_stride = grid->pitch * (long long)grid->ny;
    HIP_TRY(hipSetDevice(m->device));
    return launch_map(m->dm, kind, mode, outs, r, mult, g, precision, px, (hipStream_t)stream, m->nhalos);
}

int lt_scale_device(double* x, double s, long long n, void* stream)
{
    if (n < 0 || (n && !x)) return fail(LT_ERR_INVALID, "bad array");
    if (!n) return LT_OK;
    int dev, st;
    if ((st = current_device(&dev))) return st;
    hipError_t e = launch_scale(x, s, (size_t)n, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "scale launch");
    if (!stream) HIP_TRY(hipStreamSynchronize(nullptr));
    return LT_OK;
}

int lt_map_epilogue_device(int kind, int mode, double* out0, double* out1, double* out2, double* out3, double* out4,
                           const double* a, const double* b, const double* c, long long n, double r, double mult,
                           void* stream)
{
    int st = check_kind_mode(kind, mode);
    if (st) return st;
    if (kind != LT_MAP_AMPLI && kind != LT_MAP_SHEAR && kind != LT_MAP_MASS && kind != LT_MAP_FIELDS)
        return fail(LT_ERR_INVALID, "epilogue kinds are AMPLI, SHEAR, MASS and FIELDS");
    if (n < 0 || (n && (!a || !b || !c))) return fail(LT_ERR_INVALID, "bad Hessian planes");
    if (kind != LT_MAP_FIELDS && !out0) return fail(LT_ERR_INVALID, "out0 is NULL");
    if (!n) return LT_OK;
    int dev;
    if ((st = current_device(&dev))) return st;
    OutDesc o{};
    o.p0 = out0;
    o.p1 = out1;
    o.p2 = out2;
    o.p3 = out3;
    o.p4 = out4;
    o.kind = kind;
    o.mode = mode;
    o.r = r;
    o.mult = mult;
    hipError_t e = launch_epilogue64(a, b, c, (int64_t)n, o, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "epilogue launch");
    if (!stream) HIP_TRY(hipStreamSynchronize(nullptr));
    return LT_OK;
}

// ---- launch-shape autotuner -------------------------------------------------------------------------
int lt_autotune(lt_model m, int kind, int mode, const lt_grid_desc* grid, int precision, void* stream, int* best_px,
                float* ms)
{
    if (!m || !grid || !best_px) return fail(LT_ERR_INVALID, "NULL argument");
    int st = check_kind_mode(kind, mode);
    if (st) return st;
    if (grid->nx <= 0 || grid->ny <= 0) return fail(LT_ERR_INVALID, "bad grid");
    HIP_TRY(hipSetDevice(m->device));
    hipStream_t s = (hipStream_t)stream;
    const int np = planes_for_kind(kind);
    const size_t plane = (size_t)grid->nx * grid->ny * m->nsamples;
    double* buf = nullptr;
    if (hipMalloc(&buf, plane * np * sizeof(double)) != hipSuccess) {
        (void)hipGetLastError();
        return fail(LT_ERR_NOMEM, "autotune scratch planes");
    }
    double* outs[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    for (int k = 0; k < np; ++k) outs[k] = buf + k * plane;
    lt_grid_desc g = *grid;
    g.pitch = g.nx;
    // the shapes whose results are bit-identical to each other (the halo-lane shape sums the FP32 terms
    // in another order: it is used when asked for explicitly)
    const int cand[3] = {1, 2, 4};
    const int ncand = 3;
    float best = 1e30f;
    int bpx = 2;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t he = hipEventCreate(&e0);
    if (he == hipSuccess) he = hipEventCreate(&e1);
    for (int i = 0; i < ncand && he == hipSuccess && st == LT_OK; ++i) {
        const int px = cand[i];
        st = lt_map_device(m, kind, mode, outs[0], outs[1], outs[2], outs[3], outs[4], 0.7, 1.0, &g, precision, px,
                           stream);  // warm
        if (st == LT_OK) he = hipEventRecord(e0, s);
        for (int rep = 0; rep < 2 && st == LT_OK && he == hipSuccess; ++rep)
            st = lt_map_device(m, kind, mode, outs[0], outs[1], outs[2], outs[3], outs[4], 0.7, 1.0, &g, precision,
                               px, stream);
        if (st == LT_OK && he == hipSuccess) he = hipEventRecord(e1, s);
        if (st == LT_OK && he == hipSuccess) he = hipEventSynchronize(e1);
        float t = 0.f;
        if (st == LT_OK && he == hipSuccess) he = hipEventElapsedTime(&t, e0, e1);
        if (ms) ms[i] = t / 2.f;
        if (st == LT_OK && he == hipSuccess && t < best) {
            best = t;
            bpx = px;
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(buf);
    if (st) return st;
    if (he != hipSuccess) return hip_fail(he, "autotune timing");
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tuned[tune_key(kind, precision, m->nhalos, (long long)grid->nx * grid->ny)] = bpx;
    }
    *best_px = bpx;
    return LT_OK;
}

// ---- captured launch sequences -------------------------------------------------------------------
struct lt_graph_s {
    int device = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

int lt_graph_begin(void* stream)
{
    if (!stream) return fail(LT_ERR_INVALID, "capture needs a caller stream (not the NULL stream)");
    HIP_TRY(hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal));
    return LT_OK;
}

int lt_graph_end(void* stream, lt_graph* out)
{
    if (!out) return fail(LT_ERR_INVALID, "out is NULL");
    *out = nullptr;
    if (!stream) return fail(LT_ERR_INVALID, "capture needs a caller stream (not the NULL stream)");
    int dev, st;
    if ((st = current_device(&dev))) return st;
    hipGraph_t g = nullptr;
    HIP_TRY(hipStreamEndCapture((hipStream_t)stream, &g));
    hipGraphExec_t x = nullptr;
    hipError_t e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        return hip_fail(e, "graph instantiate");
    }
    lt_graph_s* r = new lt_graph_s();
    r->device = dev;
    r->graph = g;
    r->exec = x;
    *out = r;
    return LT_OK;
}

int lt_graph_launch(lt_graph g, void* stream)
{
    if (!g) return fail(LT_ERR_NOT_FOUND, "graph is NULL");
    HIP_TRY(hipSetDevice(g->device));
    HIP_TRY(hipGraphLaunch(g->exec, (hipStream_t)stream));
    return LT_OK;
}

int lt_graph_destroy(lt_graph g)
{
    if (!g) return LT_OK;
    (void)hipSetDevice(g->device);
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    delete g;
    return LT_OK;
}

}  // extern "C"
