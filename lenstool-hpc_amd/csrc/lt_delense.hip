// lt_delense.hip -- image-plane -> source-plane delensing and the chi^2 of multiple-image families,
// batched over lens-model samples (blockIdx.z / one thread per sample where the work is serial).
//
//   source barycentres  src/chi_CPU_barycenter.cpp:173-210
//   delensing           src/delense_CPU.cpp:6-138, src/delense_CPU_utils.hpp:81-214
//   chi^2               src/chi_computation.cpp:10-195
//
// The geometry and the chi^2 are FP64 with contraction off, operation for operation as the
// reference expressions, so that for the same gradient planes the found images (and their order:
// the reference's single-thread row-major scan order) and the chi^2 are bit-identical to the CPU
// path. Hits are collected with atomics and then put back into scan order by a per-source sort.
#include <hip/hip_runtime.h>

#include "lenstool_hip/lt_status.h"
#include "lt_delense.h"
#include "lt_delense_geom.h"
#include "lt_device.h"
#include "lt_math.h"

namespace lt {

// ------------------------------------------------------------------------------------------------
// source barycentres: one thread per (sample, family); images in order, sum then divide
// ------------------------------------------------------------------------------------------------
__global__ void k_barycentres(const double* __restrict__ img_gx, const double* __restrict__ img_gy,
                              DelenseDesc d)
{
#pragma clang fp contract(off)
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int s = blockIdx.y;
    if (k >= d.nsets) return;
    const int first = d.first_image[k], ni = d.nimages[k];
    double sx = 0., sy = 0.;
    for (int i = first; i < first + ni; ++i) {
        const int64_t j = (int64_t)s * d.nimg_total + i;
        sx += d.img_x[i] - d.img_dr[i] * img_gx[j];
        sy += d.img_y[i] - d.img_dr[i] * img_gy[j];
    }
    const int64_t o = (int64_t)s * d.nsets + k;
    d.src_x[o] = sx / ni;
    d.src_y[o] = sy / ni;
    d.src_dr[o] = d.img_dr[first];
}

// corners of cell (xid, yid) of a sample's gradient planes, image-plane coordinates as the reference
// forms them (src/delense_CPU.cpp:64-85)
__device__ __forceinline__ CellCorners load_corners(const double* __restrict__ gx, const double* __restrict__ gy,
                                                    const DelenseDesc& d, int xid, int yid)
{
#pragma clang fp contract(off)
    const int n = d.n;
    const int64_t g = (int64_t)yid * n + xid;
    CellCorners c;
    c.xp = d.xmin + xid * d.dx;
    c.yp = d.ymin + yid * d.dy;
    c.xq = c.xp + d.dx;
    c.yq = c.yp + d.dy;
    c.g0x = gx[g];
    c.g0y = gy[g];
    c.g1x = gx[g + n];
    c.g1y = gy[g + n];
    c.g2x = gx[g + 1];
    c.g2y = gy[g + 1];
    c.g3x = gx[g + n + 1];
    c.g3y = gy[g + n + 1];
    return c;
}

// ------------------------------------------------------------------------------------------------
// per-sample bounds of the sources for the scan's cull: max |x|, |y| and the range of D_ls/D_s; a
// non-finite source makes the bounds NaN, which switches the cull off for the whole sample
// ------------------------------------------------------------------------------------------------
__global__ void k_source_bounds(DelenseDesc d)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= d.nsamples) return;
    double pmax = 0., lo = 0., hi = 0., sum = 0.;  // NaN and infinity reach the sum (fmin/fmax drop NaN)
    for (int k = 0; k < d.nsets; ++k) {
        const int64_t sk = (int64_t)s * d.nsets + k;
        const double x = fabs(d.src_x[sk]), y = fabs(d.src_y[sk]), r = d.src_dr[sk];
        sum += x + y + fabs(r);
        pmax = fmax(pmax, fmax(x, y));
        lo = k ? fmin(lo, r) : r;
        hi = k ? fmax(hi, r) : r;
    }
    const bool finite = sum <= 1e300 && pmax <= 1e100 && fabs(lo) <= 1e100 && fabs(hi) <= 1e100;
    const double nan = __builtin_nan("");
    d.src_bound[3 * (int64_t)s] = finite ? pmax : nan;
    d.src_bound[3 * (int64_t)s + 1] = finite ? lo : nan;
    d.src_bound[3 * (int64_t)s + 2] = finite ? hi : nan;
}

// append a hit of source k of sample s in cell `cell` (0 = none, 1 upper, 2 lower triangle)
__device__ __forceinline__ void record_hit(const DelenseDesc& d, int s, int k, int cell, int hit)
{
    if (!hit) return;
    const int64_t sk = (int64_t)s * d.nsets + k;
    const int slot = atomicAdd(&d.count[sk], 1);
    if (slot < DELENSE_CAP) d.hits[sk * DELENSE_CAP + slot] = cell * 2 + (hit - 1);
}

// ------------------------------------------------------------------------------------------------
// delensing scan: one thread per (cell, sample), a wavefront per 8 x 8 tile of cells, all sources
// (staged in LDS, SCAN_BLOCK per pass); hits appended per (sample, source)
// ------------------------------------------------------------------------------------------------
constexpr int SCAN_BLOCK = 256;  // threads per block (4 wavefronts: 16 x 16 cells); also the sources staged in LDS per pass
constexpr int SCAN_SIDE = 16;    // cells per block side
constexpr int SCAN_UNROLL = 4;   // sources culled per step of the per-cell scan loop
constexpr int WAVE_SIZE = 64;
static_assert(SCAN_BLOCK == 4 * WAVE_SIZE && SCAN_SIDE == 16, "four wavefronts of 8 x 8 cells per block");

// One pass over cnt staged sources. SCAN_UNROLL sources per step: one margin (the smallest of their
// culls) decides the step in straight-line code; the reference test runs only where a cull fails (a
// few cells around each image, degenerate or non-finite cells), one source at a time. RHO_ONLY: every
// lane's cell has d_ok (or is past the row end), so the margin needs only the distance.
template <bool RHO_ONLY>
__device__ __forceinline__ void scan_pass(const DelenseDesc& d, const CellCorners& c, const CellFilter& f,
                                          const CellFilter32& f32, bool active, int s, int k0, int cell, int cnt,
                                          const double* lsx, const double* lsy, const double* ldr, const float* lsx32,
                                          const float* lsy32, const float* ldr32)
{
    int j0 = 0;
    for (; j0 + SCAN_UNROLL <= cnt; j0 += SCAN_UNROLL) {
        if constexpr (RHO_ONLY) {
            // FP32 pre-test of the step (lt_delense_geom.h cell_filter32): sound on its own, and it decides
            // nearly every step; the FP64 margin below runs only where it does not
            float mg32 = far_margin_rho32(f32, lsx32[j0], lsy32[j0], ldr32[j0]);
#pragma unroll
            for (int u = 1; u < SCAN_UNROLL; ++u)
                mg32 = __builtin_fminf(mg32, far_margin_rho32(f32, lsx32[j0 + u], lsy32[j0 + u], ldr32[j0 + u]));
            if (!active | (mg32 > 0.f)) continue;
        }
        double mg = RHO_ONLY ? far_margin_rho(f, c, lsx[j0], lsy[j0], ldr[j0]) : far_margin(f, c, lsx[j0], lsy[j0], ldr[j0]);
#pragma unroll
        for (int u = 1; u < SCAN_UNROLL; ++u)
            mg = fmin(mg, RHO_ONLY ? far_margin_rho(f, c, lsx[j0 + u], lsy[j0 + u], ldr[j0 + u])
                                   : far_margin(f, c, lsx[j0 + u], lsy[j0 + u], ldr[j0 + u]));
        if (!active | (f.on & (mg > 0.))) continue;
#pragma unroll 1
        for (int j = j0; j < j0 + SCAN_UNROLL; ++j)
            if (!far_from_cell(f, c, lsx[j], lsy[j], ldr[j])) record_hit(d, s, k0 + j, cell, cell_hit(c, lsx[j], lsy[j], ldr[j]));
    }
#pragma unroll 1
    for (; j0 < cnt; ++j0)
        if (active && !far_from_cell(f, c, lsx[j0], lsy[j0], ldr[j0]))
            record_hit(d, s, k0 + j0, cell, cell_hit(c, lsx[j0], lsy[j0], ldr[j0]));
}



// One pass over cnt staged sources with the tile pre-test (lt_delense_geom.h tile_filter32): lane j
// tests source kc + j against the whole tile, and only the sources it keeps (few: those whose image lies
// on or next to the tile) go through the per-cell tests, one source at a time for all cells.
__device__ __forceinline__ void scan_tile(const DelenseDesc& d, const CellCorners& c, const CellFilter& f,
                                          const CellFilter32& f32, const CellFilter32& tf, bool active, int s, int k0,
                                          int cell, int cnt, const double* lsx, const double* lsy, const double* ldr,
                                          const float* lsx32, const float* lsy32, const float* ldr32)
{
    const int lane = (int)threadIdx.x & (WAVE_SIZE - 1);
    for (int kc = 0; kc < cnt; kc += WAVE_SIZE) {
        const int j = kc + lane;  // < SCAN_BLOCK: kc < cnt <= SCAN_BLOCK is a multiple of the wavefront size
        const float mg = j < cnt ? far_margin_rho32(tf, lsx32[j], lsy32[j], ldr32[j]) : 1.f;
        unsigned long long keep = __ballot(!(mg > 0.f));
        while (keep) {
            const int jj = kc + __builtin_ctzll(keep);
            keep &= keep - 1;
            if (active && !(far_margin_rho32(f32, lsx32[jj], lsy32[jj], ldr32[jj]) > 0.f) &&
                !far_from_cell(f, c, lsx[jj], lsy[jj], ldr[jj]))
                record_hit(d, s, k0 + jj, cell, cell_hit(c, lsx[jj], lsy[jj], ldr[jj]));
        }
    }
}

__global__ __launch_bounds__(SCAN_BLOCK) void k_delense(const double* __restrict__ gx, const double* __restrict__ gy,
                                                        DelenseDesc d)
{
    __shared__ double lsx[SCAN_BLOCK], lsy[SCAN_BLOCK], ldr[SCAN_BLOCK];
    __shared__ float lsx32[SCAN_BLOCK], lsy32[SCAN_BLOCK], ldr32[SCAN_BLOCK];
    const int n = d.n;
    const int lane = (int)threadIdx.x & (WAVE_SIZE - 1), wv = (int)threadIdx.x / WAVE_SIZE;
    const int xid = blockIdx.x * SCAN_SIDE + (wv & 1) * 8 + (lane & 7);
    const int yid = blockIdx.y * SCAN_SIDE + (wv >> 1) * 8 + (lane >> 3);
    const int s = blockIdx.z;
    const bool active = xid < n - 1 && yid < n - 1;
    const int64_t base = (int64_t)s * d.grid_stride;
    const int cell = yid * (n - 1) + xid;
    // the first pass's sources, loaded before the cell set-up so that their latency overlaps it
    const int cnt0 = min(SCAN_BLOCK, d.nsets);
    double sx0 = 0., sy0 = 0., sdr0 = 0.;
    if ((int)threadIdx.x < cnt0) {
        const int64_t sk = (int64_t)s * d.nsets + threadIdx.x;
        sx0 = d.src_x[sk];
        sy0 = d.src_y[sk];
        sdr0 = d.src_dr[sk];
    }
    // lanes past the grid read the last cell of their row or column (in bounds) and record nothing
    const CellCorners c = load_corners(gx + base, gy + base, d, min(xid, n - 2), min(yid, n - 2));
    const double* sb = d.src_bound + 3 * (int64_t)s;
    const CellFilter f = cell_filter(c, sb[0], sb[1], sb[2]);
    const CellFilter32 f32 = cell_filter32(f, c);
    const unsigned long long act = __ballot(active);
    const bool rho_only = __all(!active || f.d_ok);  // uniform over the wavefront
    // the tile pre-test when every active cell has d_ok: bounds about the cell at the tile's centre
    // (lane 36, cell (4, 4) of the tile) when it is in the grid, else the first active lane's
    CellFilter32 tf = f32;
    tf.rf2 = __builtin_inff();
    if (rho_only && act) {
        const int r = ((act >> 36) & 1ull) ? 36 : __builtin_ctzll(act);
        const double xr = read_lane64(c.xp, r), yr = read_lane64(c.yp, r), gxr = read_lane64(c.g0x, r),
                     gyr = read_lane64(c.g0y, r);
        const double drmax = __builtin_fmax(__builtin_fabs(sb[1]), __builtin_fabs(sb[2]));
        const double t = active ? tile_reach(f, c, xr, yr, gxr, gyr, drmax) : 0.;
        tf = tile_filter32(xr, yr, gxr, gyr, wave_max64(t), read_lane64(f.m, r), true);
    }
    const bool tile = tf.rf2 < __builtin_inff();  // uniform over the wavefront
    for (int k0 = 0; k0 < d.nsets; k0 += SCAN_BLOCK) {
        const int cnt = min(SCAN_BLOCK, d.nsets - k0);
        __syncthreads();  // the previous pass is done with the staged sources
        if ((int)threadIdx.x < cnt) {
            const int64_t sk = (int64_t)s * d.nsets + k0 + threadIdx.x;
            const double sx = k0 ? d.src_x[sk] : sx0, sy = k0 ? d.src_y[sk] : sy0, sdr = k0 ? d.src_dr[sk] : sdr0;
            lsx[threadIdx.x] = sx;
            lsy[threadIdx.x] = sy;
            ldr[threadIdx.x] = sdr;
            lsx32[threadIdx.x] = (float)sx;
            lsy32[threadIdx.x] = (float)sy;
            ldr32[threadIdx.x] = (float)sdr;
        }
        __syncthreads();
        if (!act) continue;  // a wavefront past the grid: nothing to scan (it still stages and syncs)
        if (tile) scan_tile(d, c, f, f32, tf, active, s, k0, cell, cnt, lsx, lsy, ldr, lsx32, lsy32, ldr32);
        else if (rho_only) scan_pass<true>(d, c, f, f32, active, s, k0, cell, cnt, lsx, lsy, ldr, lsx32, lsy32, ldr32);
        else scan_pass<false>(d, c, f, f32, active, s, k0, cell, cnt, lsx, lsy, ldr, lsx32, lsy32, ldr32);
    }
}

// ------------------------------------------------------------------------------------------------
// back to scan order, barycentres of the first MAXIMPERSOURCE hits: one thread per (sample, source)
// ------------------------------------------------------------------------------------------------
__global__ void k_delense_finish(const double* __restrict__ gx, const double* __restrict__ gy, DelenseDesc d)
{
#pragma clang fp contract(off)
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int s = blockIdx.y;
    if (k >= d.nsets) return;
    const int64_t sk = (int64_t)s * d.nsets + k;
    int cnt = d.count[sk];
    int status = LT_OK;
    if (cnt > LT_MAXIMPERSOURCE) status = LT_ERR_OVERFLOW;
    int* h = d.hits + sk * DELENSE_CAP;
    if (cnt > DELENSE_CAP) {
        // the slots hold an arbitrary subset of the hits: take the first MAXIMPERSOURCE in scan order
        // again, serially (an error case: the status says the source overflowed)
        const int64_t base = (int64_t)s * d.grid_stride;
        int got = 0;
        for (int yid = 0; yid < d.n - 1 && got < LT_MAXIMPERSOURCE; ++yid)
            for (int xid = 0; xid < d.n - 1 && got < LT_MAXIMPERSOURCE; ++xid) {
                const int hit = cell_hit(load_corners(gx + base, gy + base, d, xid, yid), d.src_x[sk], d.src_y[sk],
                                         d.src_dr[sk]);
                if (hit) h[got++] = (yid * (d.n - 1) + xid) * 2 + (hit - 1);
            }
        cnt = got;
    }
    const int kept = cnt;
    for (int i = 1; i < kept; ++i) {  // insertion sort: cell-major, upper before lower
        int v = h[i], j = i - 1;
        while (j >= 0 && h[j] > v) {
            h[j + 1] = h[j];
            --j;
        }
        h[j + 1] = v;
    }
    const int nout = kept < LT_MAXIMPERSOURCE ? kept : LT_MAXIMPERSOURCE;
    for (int i = 0; i < nout; ++i) {
        const int code = h[i];
        const int cell = code >> 1;
        const int yid = cell / (d.n - 1), xid = cell % (d.n - 1);
        const double xp = d.xmin + xid * d.dx;
        const double yp = d.ymin + yid * d.dy;
        double bx, by;
        if ((code & 1) == 0) {  // upper (x,y) (x,y+dy) (x+dx,y)
            bx = (xp + xp + (xp + d.dx)) / 3.;
            by = (yp + (yp + d.dy) + yp) / 3.;
        } else {  // lower (x+dx,y+dy) (x+dx,y) (x,y+dy)
            bx = ((xp + d.dx) + (xp + d.dx) + xp) / 3.;
            by = ((yp + d.dy) + yp + (yp + d.dy)) / 3.;
        }
        d.pos_x[sk * LT_MAXIMPERSOURCE + i] = bx;
        d.pos_y[sk * LT_MAXIMPERSOURCE + i] = by;
    }
    d.loc[sk] = nout;
    if (status != LT_OK) atomicMax(&d.status[s], status);
}

// ------------------------------------------------------------------------------------------------
// chi^2 matching (src/chi_computation.cpp:10-195, with its two early-stopping loop conditions at
// :110 and :118): one thread per (sample, source) writes the per-image chi terms; one thread per
// sample then sums them in the reference's order.
// ------------------------------------------------------------------------------------------------
__device__ __forceinline__ double pdist(double ax, double ay, double bx, double by)
{
#pragma clang fp contract(off)
    double x = ax - bx, y = ay - by;
    return sqrt(x * x + y * y);
}

__global__ void k_chi_match(DelenseDesc d)
{
#pragma clang fp contract(off)
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int s = blockIdx.y;
    if (k >= d.nsets) return;
    const int64_t sk = (int64_t)s * d.nsets + k;
    const int first = d.first_image[k], ni = d.nimages[k];
    int nfound[LT_MAXIMPERSOURCE];
    double tx[LT_MAXIMPERSOURCE], ty[LT_MAXIMPERSOURCE], dist[LT_MAXIMPERSOURCE];
    for (int i = 0; i < LT_MAXIMPERSOURCE; ++i) {
        nfound[i] = 0;
        tx[i] = ty[i] = 0.;
    }
    const int nf = d.loc[sk];
    for (int ii = 0; ii < nf; ++ii) {
        const double px = d.pos_x[sk * LT_MAXIMPERSOURCE + ii], py = d.pos_y[sk * LT_MAXIMPERSOURCE + ii];
        int idx = 0;
        dist[0] = pdist(px, py, d.img_x[first], d.img_y[first]);
        for (int i = 1; i < ni; i++) {
            dist[i] = pdist(px, py, d.img_x[first + i], d.img_y[first + i]);
            if (dist[i] < dist[idx]) idx = i;
        }
        if (nfound[idx] == 0) {
            tx[idx] = px;
            ty[idx] = py;
            nfound[idx]++;
        } else if (nfound[idx] > 0) {
            double tmpx, tmpy;
            if (dist[idx] < pdist(d.img_x[first + idx], d.img_y[first + idx], tx[idx], ty[idx])) {
                tmpx = tx[idx];
                tmpy = ty[idx];
                tx[idx] = px;
                ty[idx] = py;
            } else {
                tmpx = px;
                tmpy = py;
            }
            int second = 0;
            for (int i = 1; i < ni && (i != idx); i++)
                if (dist[i] > dist[second]) second = i;
            for (int i = 0; i < ni && nfound[i] == 0; i++) {
                if (dist[i] < dist[second]) {
                    second = i;
                    idx = i;
                    tx[idx] = tmpx;
                    ty[idx] = tmpy;
                }
            }
            nfound[idx]++;
        }
    }
    for (int i = 0; i < ni; i++) {
        double term;
        if (nfound[i] > 0) {
            double p1 = d.img_x[first + i] - tx[i];
            double p2 = d.img_y[first + i] - ty[i];
            term = p1 * p1 + p2 * p2;
        } else {
            term = 1000. * ni;
        }
        d.terms[(int64_t)s * d.nimg_total + first + i] = term;
    }
}

__global__ void k_chi_reduce(DelenseDesc d)
{
#pragma clang fp contract(off)
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= d.nsamples) return;
    double chi = 0.;
    int found = 0;
    for (int k = 0; k < d.nsets; ++k) found += d.loc[(int64_t)s * d.nsets + k];
    for (int i = 0; i < d.nimg_total; ++i) chi += d.terms[(int64_t)s * d.nimg_total + i];
    d.chi[s] = chi;
    d.found[s] = found;
}

// ------------------------------------------------------------------------------------------------
// launchers
// ------------------------------------------------------------------------------------------------
hipError_t launch_barycentres(const double* img_gx, const double* img_gy, const DelenseDesc& d, hipStream_t s)
{
    hipLaunchKernelGGL(k_barycentres, dim3((d.nsets + 63) / 64, d.nsamples), dim3(64), 0, s, img_gx, img_gy, d);
    return hipGetLastError();
}

hipError_t launch_delense(const double* gx, const double* gy, const DelenseDesc& d, hipStream_t s)
{
    hipLaunchKernelGGL(k_source_bounds, dim3((d.nsamples + 63) / 64), dim3(64), 0, s, d);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    dim3 blk(SCAN_BLOCK);
    dim3 grd((unsigned)((d.n - 1 + SCAN_SIDE - 1) / SCAN_SIDE), (unsigned)((d.n - 1 + SCAN_SIDE - 1) / SCAN_SIDE),
             (unsigned)d.nsamples);
    hipLaunchKernelGGL(k_delense, grd, blk, 0, s, gx, gy, d);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_delense_finish, dim3((d.nsets + 63) / 64, d.nsamples), dim3(64), 0, s, gx, gy, d);
    return hipGetLastError();
}

hipError_t launch_chi(const DelenseDesc& d, hipStream_t s)
{
    hipLaunchKernelGGL(k_chi_match, dim3((d.nsets + 63) / 64, d.nsamples), dim3(64), 0, s, d);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_chi_reduce, dim3((d.nsamples + 63) / 64), dim3(64), 0, s, d);
    return hipGetLastError();
}

}  // namespace lt
