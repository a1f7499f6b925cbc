// lt_critic.hip -- critical lines (det A = 0) of a det-A map by marching squares, and caustics (their
// images through the lens equation). The reference reads a `cline` block (structure_hpc.hpp:563-578,
// module_readParameters.cpp:1008-1095, compiled out) but computes no critical line; this is the
// consumer of the Hessian maps that the parameter block describes.
//
// Contour rules. Cell (i, j), i < nx - 1, j < ny - 1, has corners c0 = v(i, j), c1 = v(i+1, j),
// c2 = v(i+1, j+1), c3 = v(i, j+1); corner bit k is set when c_k > 0. An edge whose ends differ in
// bit carries a crossing at the linear interpolation t = v_a / (v_a - v_b) from its first end (bottom
// and top edges from their left end, left and right edges from their lower end), and every
// coordinate is formed from the pixel indices, so two cells that share an edge compute the same
// point bit for bit. Segments are oriented with det A > 0 on their left. The saddle cases (corner
// masks 5 and 10) join the positive corners when the sum of the four corners is positive. Each
// segment is labelled tangential (1 - kappa > 0 at its midpoint, kappa interpolated to its ends like
// the crossing) or radial (otherwise): on a critical line 1 - kappa = +gamma or -gamma.
//
// Launches: one work-group per cell row counts the row's segments (256 consecutive cells at a time,
// coalesced), one work-group turns the row counts into row offsets, one work-group per row writes
// the row's segments at its offset in cell order (a work-group prefix sum per 256 cells). The
// segment order is row-major cell order whatever the scheduling. Memory-bound: each cell row reads
// two plane rows (twice: count and write) and writes 32 B per segment.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>
#include <vector>

#include "lenstool_hip/lt_capi.h"
#include "lt_internal.h"

namespace lt {

constexpr int kCritThreads = 256;

struct CritGrid {
    double xmin, ymin, dx, dy;
    int nx, ny, istart, jstart;
    long long pitch;
};

__device__ __forceinline__ double coord(double min, int i, double d)
{
#pragma clang fp contract(off)
    const double t = (double)i * d;
    return min + t;
}

__device__ __forceinline__ double along(double base, double t, double d)
{
#pragma clang fp contract(off)
    const double u = t * d;
    return base + u;
}

__device__ __forceinline__ double mix(double a, double b, double t)
{
#pragma clang fp contract(off)
    const double u = t * (b - a);
    return a + u;
}

// segments of a cell as pairs of edges (0 bottom, 1 right, 2 top, 3 left), oriented with the
// positive side on the left; n = 0, 1 or 2
struct CellSegs {
    int n;
    int e[4];
};

__device__ __forceinline__ CellSegs cell_segments(double c0, double c1, double c2, double c3)
{
    const int cs = (c0 > 0.) | ((c1 > 0.) << 1) | ((c2 > 0.) << 2) | ((c3 > 0.) << 3);
    CellSegs r{0, {0, 0, 0, 0}};
    auto one = [&](int a, int b) {
        r.n = 1;
        r.e[0] = a;
        r.e[1] = b;
    };
    auto two = [&](int a, int b, int c, int d) {
        r.n = 2;
        r.e[0] = a;
        r.e[1] = b;
        r.e[2] = c;
        r.e[3] = d;
    };
    switch (cs) {
    case 1: one(0, 3); break;
    case 2: one(1, 0); break;
    case 3: one(1, 3); break;
    case 4: one(2, 1); break;
    case 6: one(2, 0); break;
    case 7: one(2, 3); break;
    case 8: one(3, 2); break;
    case 9: one(0, 2); break;
    case 11: one(1, 2); break;
    case 12: one(3, 1); break;
    case 13: one(0, 1); break;
    case 14: one(3, 0); break;
    case 5:
        if ((c0 + c1) + (c2 + c3) > 0.) two(0, 1, 2, 3);  // positives joined: around BR, then TL
        else two(0, 3, 2, 1);                              // positives apart: around BL, then TR
        break;
    case 10:
        if ((c0 + c1) + (c2 + c3) > 0.) two(3, 0, 1, 2);  // positives joined: around BL, then TR
        else two(1, 0, 3, 2);                              // positives apart: around BR, then TL
        break;
    default: break;
    }
    return r;
}

struct Corners {
    double c0, c1, c2, c3;
};

__device__ __forceinline__ Corners corners(const double* v, long long pitch, int i, int j)
{
    const double* r0 = v + (long long)j * pitch;
    const double* r1 = r0 + pitch;
    return Corners{r0[i], r0[i + 1], r1[i + 1], r1[i]};
}

// crossing on edge e of cell (i, j): point and interpolation parameter
__device__ __forceinline__ void edge_point(int e, const CritGrid& g, int i, int j, const Corners& c, double& x,
                                           double& y, double& t)
{
    const int gi = g.istart + i, gj = g.jstart + j;
    switch (e) {
    case 0:
        t = c.c0 / (c.c0 - c.c1);
        x = along(coord(g.xmin, gi, g.dx), t, g.dx);
        y = coord(g.ymin, gj, g.dy);
        break;
    case 1:
        t = c.c1 / (c.c1 - c.c2);
        x = coord(g.xmin, gi + 1, g.dx);
        y = along(coord(g.ymin, gj, g.dy), t, g.dy);
        break;
    case 2:
        t = c.c3 / (c.c3 - c.c2);
        x = along(coord(g.xmin, gi, g.dx), t, g.dx);
        y = coord(g.ymin, gj + 1, g.dy);
        break;
    default:
        t = c.c0 / (c.c0 - c.c3);
        x = coord(g.xmin, gi, g.dx);
        y = along(coord(g.ymin, gj, g.dy), t, g.dy);
        break;
    }
}

__device__ __forceinline__ double edge_kappa(int e, double t, const Corners& k)
{
    switch (e) {
    case 0: return mix(k.c0, k.c1, t);
    case 1: return mix(k.c1, k.c2, t);
    case 2: return mix(k.c3, k.c2, t);
    default: return mix(k.c0, k.c3, t);
    }
}

// inclusive prefix sum over the work-group in LDS; *total = the work-group's sum. Every thread of
// the work-group must call it (it synchronises).
__device__ __forceinline__ int block_scan(int v, int* lds, int* total)
{
    const int t = threadIdx.x;
    lds[t] = v;
    __syncthreads();
    for (int off = 1; off < kCritThreads; off <<= 1) {
        const int add = t >= off ? lds[t - off] : 0;
        __syncthreads();
        lds[t] += add;
        __syncthreads();
    }
    const int r = lds[t];
    *total = lds[kCritThreads - 1];
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(kCritThreads) void k_critic_count(const double* __restrict__ det, CritGrid g,
                                                                int* __restrict__ row_count)
{
    __shared__ int lds[kCritThreads];
    const int j = blockIdx.x;  // cell row 0 .. ny-2
    const int ncell = g.nx - 1;
    int n = 0;
    for (int i = threadIdx.x; i < ncell; i += kCritThreads) {
        const Corners c = corners(det, g.pitch, i, j);
        n += cell_segments(c.c0, c.c1, c.c2, c.c3).n;
    }
    int tot = 0;
    block_scan(n, lds, &tot);
    if (threadIdx.x == 0) row_count[j] = tot;
}

// exclusive prefix of the row counts; row_off[nrows] = total. One work-group.
__global__ __launch_bounds__(kCritThreads) void k_critic_offsets(const int* __restrict__ row_count, int nrows,
                                                                  long long* __restrict__ row_off)
{
    __shared__ int lds[kCritThreads];
    long long carry = 0;
    for (int base = 0; base < nrows; base += kCritThreads) {
        const int r = base + (int)threadIdx.x;
        const int v = r < nrows ? row_count[r] : 0;
        int tot = 0;
        const int inc = block_scan(v, lds, &tot);
        if (r < nrows) row_off[r] = carry + inc - v;
        carry += tot;
    }
    if (threadIdx.x == 0) row_off[nrows] = carry;
}

__global__ __launch_bounds__(kCritThreads) void k_critic_write(const double* __restrict__ det,
                                                                const double* __restrict__ kappa, CritGrid g,
                                                                const long long* __restrict__ row_off,
                                                                long long max_seg, double* __restrict__ segs,
                                                                int* __restrict__ kinds)
{
    __shared__ int lds[kCritThreads];
    const int j = blockIdx.x;
    const int ncell = g.nx - 1;
    long long off = row_off[j];
    for (int i0 = 0; i0 < ncell; i0 += kCritThreads) {
        const int i = i0 + (int)threadIdx.x;
        Corners c{0., 0., 0., 0.};
        CellSegs s{0, {0, 0, 0, 0}};
        if (i < ncell) {
            c = corners(det, g.pitch, i, j);
            s = cell_segments(c.c0, c.c1, c.c2, c.c3);
        }
        int tot = 0;
        const int inc = block_scan(s.n, lds, &tot);
        const long long pos = off + inc - s.n;
        if (s.n) {
            Corners k{0., 0., 0., 0.};
            if (kinds) k = corners(kappa, g.pitch, i, j);
            const int nq = pos + s.n <= max_seg ? s.n : (int)(pos < max_seg ? max_seg - pos : 0);
            for (int q = 0; q < nq; ++q) {
                const long long idx = pos + q;
                double xa, ya, ta, xb, yb, tb;
                edge_point(s.e[2 * q], g, i, j, c, xa, ya, ta);
                edge_point(s.e[2 * q + 1], g, i, j, c, xb, yb, tb);
                double* o = segs + 4 * idx;
                o[0] = xa;
                o[1] = ya;
                o[2] = xb;
                o[3] = yb;
                if (kinds) {
                    const double ka = edge_kappa(s.e[2 * q], ta, k), kb = edge_kappa(s.e[2 * q + 1], tb, k);
                    kinds[idx] = 1. - 0.5 * (ka + kb) > 0. ? 0 : 1;
                }
            }
        }
        off += tot;
    }
}

}  // namespace lt

using namespace lt;

namespace {

int fail(int code, const std::string& msg) { return set_error(code, msg); }

// Row counts (cell rows, padded to 256 bytes) followed by nrows + 1 row offsets.
size_t critic_scratch_bytes(int nrows)
{
    return (((size_t)nrows * sizeof(int) + 255) & ~(size_t)255) + (size_t)(nrows + 1) * sizeof(long long);
}

// The launches of both critic entry points on a caller-provided scratch buffer of
// critic_scratch_bytes(ny - 1) bytes; every scratch element the kernels read is written first, so
// the buffer needs no clearing. Waits for the stream.
int critic_run(const double* det, const double* kappa, const lt_grid_desc* grid, double* segs, int* kinds,
               long long max_segments, long long* nsegments, hipStream_t s, void* buf)
{
    const int nrows = grid->ny - 1;
    CritGrid g{grid->xmin, grid->ymin, grid->dx, grid->dy, grid->nx, grid->ny, grid->istart, grid->jstart, grid->pitch};
    int* row_count = (int*)buf;
    long long* row_off = (long long*)((char*)buf + (((size_t)nrows * sizeof(int) + 255) & ~(size_t)255));
    long long total = 0;
    hipError_t e = hipSuccess;
    hipLaunchKernelGGL(k_critic_count, dim3(nrows), dim3(kCritThreads), 0, s, det, g, row_count);
    e = hipGetLastError();
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_critic_offsets, dim3(1), dim3(kCritThreads), 0, s, row_count, nrows, row_off);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(&total, row_off + nrows, sizeof(long long), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && segs && max_segments > 0) {
        hipLaunchKernelGGL(k_critic_write, dim3(nrows), dim3(kCritThreads), 0, s, det, kappa, g, row_off,
                           max_segments, segs, kinds);
        e = hipGetLastError();
    }
    const hipError_t es = hipStreamSynchronize(s);
    if (e == hipSuccess) e = es;
    if (e != hipSuccess) return hip_error(e, "critical lines");
    *nsegments = total;
    if (segs && total > max_segments)
        return fail(LT_ERR_OVERFLOW, std::to_string(total) + " segments, room for " + std::to_string(max_segments));
    return LT_OK;
}

int critic_check(const double* det, const double* kappa, const lt_grid_desc* grid, double* segs, int* kinds,
                 long long max_segments, long long* nsegments)
{
    if (!det || !grid || !nsegments) return fail(LT_ERR_INVALID, "NULL argument");
    if (grid->nx < 2 || grid->ny < 2 || grid->pitch < grid->nx)
        return fail(LT_ERR_INVALID, "a contour needs at least 2 x 2 pixels and pitch >= nx");
    if (kinds && !kappa) return fail(LT_ERR_INVALID, "segment kinds need the kappa plane");
    if (segs && max_segments < 0) return fail(LT_ERR_INVALID, "negative max_segments");
    if (!std::isfinite(grid->dx) || !std::isfinite(grid->dy)) return fail(LT_ERR_INVALID, "non-finite step");
    return LT_OK;
}

}  // namespace

// Scratch of lt_critic_device_noalloc: a fixed array in the code object (each GPU holds its own copy
// for the life of the process). Host code never uses the array's name as an address: its device
// address on each GPU is taken once with hipGetSymbolAddress and kept in g_noalloc_ptr.
constexpr int kNoAllocRows = LT_CRITIC_NOALLOC_MAX_ROWS;
constexpr size_t kNoAllocBytes = (((size_t)kNoAllocRows * sizeof(int) + 255) & ~(size_t)255) +
                                 (size_t)(kNoAllocRows + 1) * sizeof(long long);
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
void* g_noalloc_ptr[kMaxDevices] = {};

}  // namespace

namespace lt {

// With `dev` current: look up (once per device) the device address of the no-alloc scratch array.
int critic_noalloc_setup(int dev, void** ptr)
{
    if (dev < 0 || dev >= kMaxDevices) return set_error(LT_ERR_NO_DEVICE, "device index out of range");
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (!g_noalloc_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            return hip_error(e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress(critic scratch)");
        }
        g_noalloc_ptr[dev] = p;
    }
    if (ptr) *ptr = g_noalloc_ptr[dev];
    return LT_OK;
}

}  // namespace lt

extern "C" {

int lt_critic_device(const double* det, const double* kappa, const lt_grid_desc* grid, double* segs, int* kinds,
                     long long max_segments, long long* nsegments, void* stream)
{
    int st = critic_check(det, kappa, grid, segs, kinds, max_segments, nsegments);
    if (st) return st;
    int dev;
    if ((st = current_device(&dev))) return st;
    void* buf = nullptr;
    LT_HIP_TRY(hipMalloc(&buf, critic_scratch_bytes(grid->ny - 1)));
    st = critic_run(det, kappa, grid, segs, kinds, max_segments, nsegments, (hipStream_t)stream, buf);
    (void)hipFree(buf);
    return st;
}

int lt_critic_device_noalloc(const double* det, const double* kappa, const lt_grid_desc* grid, double* segs,
                             int* kinds, long long max_segments, long long* nsegments, void* stream)
{
    int st = critic_check(det, kappa, grid, segs, kinds, max_segments, nsegments);
    if (st) return st;
    if (grid->ny - 1 > kNoAllocRows)
        return fail(LT_ERR_CAPACITY, std::to_string(grid->ny) + " rows: the no-alloc contour takes at most " +
                                         std::to_string(kNoAllocRows + 1));
    int dev;
    if ((st = current_device(&dev))) return st;
    void* ptr = nullptr;
    if ((st = critic_noalloc_setup(dev, &ptr))) return st;
    return critic_run(det, kappa, grid, segs, kinds, max_segments, nsegments, (hipStream_t)stream, ptr);
}

int lt_critic_lines(const lt_cosmo_param* cosmo, const lt_grid_param* frame, const lt_potential_soa* lens, int nhalos,
                    double z_lens, double z_source, int nx, int ny, int precision, double* critic, double* caustic,
                    int* kinds, long long max_segments, long long* nsegments)
{
    if (!cosmo || !frame || !lens || !nsegments) return fail(LT_ERR_INVALID, "NULL argument");
    if (nx < 2 || ny < 2) return fail(LT_ERR_INVALID, "a contour needs at least 2 x 2 pixels");
    if ((critic || caustic || kinds) && max_segments < 0) return fail(LT_ERR_INVALID, "negative max_segments");
    double r = 0, mult = 1;
    const double dx = (frame->xmax - frame->xmin) / (nx - 1), dy = (frame->ymax - frame->ymin) / (ny - 1);
    // external-shear (type 14) halos have a Hessian but no deflection kernel (the reference has none
    // either): their critical lines are computed, their caustics are not
    bool no_deflection = false;
    for (int i = 0; i < nhalos && lens->type; ++i) no_deflection |= lens->type[i] == 14;
    int st = lt_map_factors(LT_MAP_AMPLI, 1, cosmo, z_lens, z_source, dx, dy, &r, &mult);
    if (st) return st;
    lt_model m = nullptr;
    if ((st = lt_model_create(&m, lens, nhalos, LT_MAP_HESSIAN))) return st;
    void* s = nullptr;
    double *kap = nullptr, *det = nullptr, *dsegs = nullptr;
    int* dkinds = nullptr;
    const long long npix = (long long)nx * ny;
    const long long room = (critic || caustic || kinds) ? max_segments : 0;
    std::vector<double> segs((size_t)std::max(room, 0LL) * 4);
    long long total = 0;
    int st2 = LT_OK;
    do {
        if ((st = lt_stream_create(&s))) break;
        if ((st = lt_dev_alloc((void**)&kap, npix * (long long)sizeof(double)))) break;
        if ((st = lt_dev_alloc((void**)&det, npix * (long long)sizeof(double)))) break;
        if (room > 0) {
            if ((st = lt_dev_alloc((void**)&dsegs, room * 4 * (long long)sizeof(double)))) break;
            if (kinds && (st = lt_dev_alloc((void**)&dkinds, room * (long long)sizeof(int)))) break;
        }
        lt_grid_desc g{frame->xmin, frame->ymin, dx, dy, nx, ny, 0, 0, nx};
        if ((st = lt_map_device(m, LT_MAP_FIELDS, 0, kap, nullptr, nullptr, det, nullptr, r, 1.0, &g, precision, 0, s)))
            break;
        st2 = lt_critic_device(det, kinds ? kap : nullptr, &g, room > 0 ? dsegs : nullptr, dkinds, room, &total, s);
        if (st2 && st2 != LT_ERR_OVERFLOW) {
            st = st2;
            break;
        }
        const long long got = std::min(total, room);
        if (got > 0) {
            if ((st = lt_memcpy_d2h(segs.data(), dsegs, got * 4 * (long long)sizeof(double), nullptr))) break;
            if (kinds && (st = lt_memcpy_d2h(kinds, dkinds, got * (long long)sizeof(int), nullptr))) break;
            if (critic) std::copy(segs.begin(), segs.begin() + got * 4, critic);
            if (caustic && no_deflection) {
                std::fill(caustic, caustic + got * 4, std::nan(""));
            } else if (caustic) {
                // beta = theta - r * alpha(theta) at both ends of every segment
                std::vector<double> px((size_t)got * 2), py((size_t)got * 2), ax((size_t)got * 2), ay((size_t)got * 2);
                for (long long k = 0; k < got; ++k) {
                    px[2 * k] = segs[4 * k];
                    py[2 * k] = segs[4 * k + 1];
                    px[2 * k + 1] = segs[4 * k + 2];
                    py[2 * k + 1] = segs[4 * k + 3];
                }
                if ((st = lt_point_gradients(ax.data(), ay.data(), px.data(), py.data(), (int)(got * 2), lens, nhalos)))
                    break;
                for (long long k = 0; k < 2 * got; ++k) {
                    caustic[2 * k] = px[k] - r * ax[k];
                    caustic[2 * k + 1] = py[k] - r * ay[k];
                }
            }
        }
    } while (false);
    if (dkinds) lt_dev_free(dkinds);
    if (dsegs) lt_dev_free(dsegs);
    if (det) lt_dev_free(det);
    if (kap) lt_dev_free(kap);
    if (s) lt_stream_destroy(s);
    lt_model_destroy(m);
    if (st) return st;
    *nsegments = total;
    if (st2 == LT_OK && caustic && no_deflection)
        return fail(LT_ERR_TYPE, "critical lines written; caustics unavailable (NaN): the model has external-shear "
                                 "(type 14) halos, which have no deflection kernel");
    return st2;
}

}  // extern "C"
