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
#include "lt_device.h"
#include "lt_math.h"

namespace lt {

struct P2 {
    double x, y;
};

// scalar triple product of (A,1), (B,1), (C,1) (src/delense_CPU_utils.hpp:110-117)
__device__ __forceinline__ double tri_det(P2 A, P2 B, P2 C)
{
#pragma clang fp contract(off)
    return (B.x * C.y - B.y * C.x + A.x * B.y - A.y * B.x + A.y * C.x - A.x * C.y);
}

// inside-or-on-border test (src/delense_CPU_utils.hpp:125-142; :180-193 is the same test)
__device__ __forceinline__ bool tri_inside(P2 P, P2 a, P2 b, P2 c)
{
#pragma clang fp contract(off)
    double d = tri_det(a, b, c);
    if (tri_det(a, b, P) * d < 0.) return false;
    if (tri_det(b, c, P) * d < 0.) return false;
    if (tri_det(c, a, P) * d < 0.) return false;
    return true;
}

__device__ __forceinline__ P2 to_source(double x, double y, double dr, double gx, double gy)
{
#pragma clang fp contract(off)
    P2 s;
    s.x = x - dr * gx;
    s.y = y - dr * gy;
    return s;
}

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

// Which triangle of cell (xid, yid) holds source P after delensing: 0 none, 1 upper, 2 lower
// (src/delense_CPU.cpp:60-112: the upper triangle is tested first; the lower only if it misses).
__device__ __forceinline__ int cell_hit(const double* __restrict__ gx, const double* __restrict__ gy,
                                        const DelenseDesc& d, int xid, int yid, double px, double py, double dr)
{
#pragma clang fp contract(off)
    const int n = d.n;
    const int64_t g = (int64_t)yid * n + xid;
    const double xp = d.xmin + xid * d.dx;
    const double yp = d.ymin + yid * d.dy;
    const double xq = xp + d.dx, yq = yp + d.dy;
    const P2 P = {px, py};
    // gradient at the four corners: g (x,y), g+n (x,y+dy), g+1 (x+dx,y), g+n+1 (x+dx,y+dy)
    // upper triangle a=(x,y) b=(x,y+dy) c=(x+dx,y)
    const P2 ua = to_source(xp, yp, dr, gx[g], gy[g]);
    const P2 ub = to_source(xp, yq, dr, gx[g + n], gy[g + n]);
    const P2 uc = to_source(xq, yp, dr, gx[g + 1], gy[g + 1]);
    if (tri_inside(P, ua, ub, uc)) return 1;
    // lower triangle a=(x+dx,y+dy) b=(x+dx,y) c=(x,y+dy)
    const P2 la = to_source(xq, yq, dr, gx[g + n + 1], gy[g + n + 1]);
    const P2 lb = to_source(xq, yp, dr, gx[g + 1], gy[g + 1]);
    const P2 lc = to_source(xp, yq, dr, gx[g + n], gy[g + n]);
    return tri_inside(P, la, lb, lc) ? 2 : 0;
}

// ------------------------------------------------------------------------------------------------
// delensing scan: one thread per (cell, sample), all sources; hits appended per (sample, source)
// ------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_delense(const double* __restrict__ gx, const double* __restrict__ gy,
                                                 DelenseDesc d)
{
#pragma clang fp contract(off)
    const int n = d.n;
    const int xid = blockIdx.x * blockDim.x + threadIdx.x;
    const int yid = blockIdx.y;
    const int s = blockIdx.z;
    if (xid >= n - 1 || yid >= n - 1) return;
    const int64_t base = (int64_t)s * d.grid_stride;
    const int cell = yid * (n - 1) + xid;
    for (int k = 0; k < d.nsets; ++k) {
        const int64_t sk = (int64_t)s * d.nsets + k;
        const int hit = cell_hit(gx + base, gy + base, d, xid, yid, d.src_x[sk], d.src_y[sk], d.src_dr[sk]);
        if (hit) {
            const int slot = atomicAdd(&d.count[sk], 1);
            if (slot < DELENSE_CAP) d.hits[sk * DELENSE_CAP + slot] = cell * 2 + (hit - 1);
        }
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
                const int hit = cell_hit(gx + base, gy + base, d, xid, yid, d.src_x[sk], d.src_y[sk], d.src_dr[sk]);
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
    dim3 blk(256);
    dim3 grd((unsigned)((d.n - 1 + 255) / 256), (unsigned)(d.n - 1), (unsigned)d.nsamples);
    hipLaunchKernelGGL(k_delense, grd, blk, 0, s, gx, gy, d);
    hipError_t e = hipGetLastError();
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
