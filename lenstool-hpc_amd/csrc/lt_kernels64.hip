// lt_kernels64.hip -- FP64 lens-map kernels for gfx950: gradient, Hessian (+ fused map epilogue),
// potential. One thread owns PX pixels of one row (columns lane, lane+64, ...); a 64x4 workgroup
// covers 64*PX columns of 4 rows. Halo records are walked in wavefront-uniform order, run by run.
//
// Formulas (restated from the mathematics; reference equivalents cited):
//   SIS gradient  src/gradient.cpp:260-299      PIEMD gradient  src/gradient.hpp:64-76,
//   dPIE gradient src/gradient.cpp:744-809      PIEMD Hessian   src/gradient2.cpp:54-82, 92-205
//   potential     src/potential_GPU.cu:47-116   epilogues       src/grid_map_*_GPU.cu
#include <hip/hip_runtime.h>

#include "lenstool_hip/lt_status.h"
#include "lt_device.h"
#include "lt_kernels.h"
#include "lt_math.h"
#include "lt_terms64.h"

namespace lt {

template <int PX, int KIND>
__global__ __launch_bounds__(WAVE* ROWS) void k_hessian64(const HaloRec* __restrict__ halos,
                                                          const Run* __restrict__ runs, int nruns,
                                                          GridDesc g, OutDesc o)
{
    const int row = blockIdx.y * ROWS + threadIdx.y;
    if (row >= g.ny) return;
    halos = select_sample(g, halos, o);
    const int col0 = blockIdx.x * (WAVE * PX) + threadIdx.x;
    const double py = pixel_coord(g.ymin, row + g.jstart, g.dy);
    double px[PX];
#pragma unroll
    for (int k = 0; k < PX; ++k) px[k] = pixel_coord(g.xmin, col0 + k * WAVE + g.istart, g.dx);
    double tT[PX], tD[PX], tB[PX];
#pragma unroll
    for (int k = 0; k < PX; ++k) tT[k] = tD[k] = tB[k] = 0.;
    for (int r = 0; r < nruns; ++r) {
        const Run run = runs[r];
        double sT[PX], sD[PX], sB[PX];
#pragma unroll
        for (int k = 0; k < PX; ++k) sT[k] = sD[k] = sB[k] = 0.;
        const int end = run.start + run.count;
        switch (run.type) {
        case TYPE_DPIE:
            for (int i = run.start; i < end; ++i) {
                const HaloRec& h = halos[i];
#pragma unroll
                for (int k = 0; k < PX; ++k) hess_piemd<true>(h, px[k], py, sT[k], sD[k], sB[k]);
            }
            break;
        case TYPE_PIEMD:
            for (int i = run.start; i < end; ++i) {
                const HaloRec& h = halos[i];
#pragma unroll
                for (int k = 0; k < PX; ++k) hess_piemd<false>(h, px[k], py, sT[k], sD[k], sB[k]);
            }
            break;
        case TYPE_SIS:
            if (run.circ) {
                // four halos per step: their records' scalar loads issue together, one wait per step
#pragma unroll 4
                for (int i = run.start; i < end; ++i) {
                    const HaloRec& h = halos[i];
#pragma unroll
                    for (int k = 0; k < PX; ++k) hess_sis_circ(h, px[k], py, sT[k], sD[k], sB[k]);
                }
            } else {
                for (int i = run.start; i < end; ++i) {
                    const HaloRec& h = halos[i];
#pragma unroll
                    for (int k = 0; k < PX; ++k) hess_sis(h, px[k], py, sT[k], sD[k], sB[k]);
                }
            }
            break;
        default:  // TYPE_SHEAR: constant lab-frame Hessian
            for (int i = run.start; i < end; ++i) {
                const HaloRec& h = halos[i];
#pragma unroll
                for (int k = 0; k < PX; ++k) {
                    sT[k] += h.h14a + h.h14c;
                    sD[k] += h.h14a - h.h14c;
                    sB[k] += h.h14b;
                }
            }
            break;
        }
        const double w = run.weight;
#pragma unroll
        for (int k = 0; k < PX; ++k) {
            tT[k] = __builtin_fma(w, sT[k], tT[k]);
            tD[k] = __builtin_fma(w, sD[k], tD[k]);
            tB[k] = __builtin_fma(w, sB[k], tB[k]);
        }
    }
#pragma unroll
    for (int k = 0; k < PX; ++k) {
        const int col = col0 + k * WAVE;
        if (col < g.nx) hess_epilogue<KIND>(o, (int64_t)row * g.pitch + col, tT[k], tD[k], tB[k]);
    }
}

template <int PX>
__global__ __launch_bounds__(WAVE* ROWS) void k_gradient64(const HaloRec* __restrict__ halos,
                                                           const Run* __restrict__ runs, int nruns,
                                                           GridDesc g, OutDesc o)
{
    const int row = blockIdx.y * ROWS + threadIdx.y;
    if (row >= g.ny) return;
    halos = select_sample(g, halos, o);
    const int col0 = blockIdx.x * (WAVE * PX) + threadIdx.x;
    const double py = pixel_coord(g.ymin, row + g.jstart, g.dy);
    double px[PX], tx[PX], ty[PX];
#pragma unroll
    for (int k = 0; k < PX; ++k) {
        px[k] = pixel_coord(g.xmin, col0 + k * WAVE + g.istart, g.dx);
        tx[k] = ty[k] = 0.;
    }
    for (int r = 0; r < nruns; ++r) {
        const Run run = runs[r];
        double sx[PX], sy[PX];
#pragma unroll
        for (int k = 0; k < PX; ++k) sx[k] = sy[k] = 0.;
        const int end = run.start + run.count;
        switch (run.type) {
        case TYPE_DPIE:
            for (int i = run.start; i < end; ++i) {
                const HaloRec& h = halos[i];
#pragma unroll
                for (int k = 0; k < PX; ++k) grad_piemd<true>(h, px[k], py, sx[k], sy[k]);
            }
            break;
        case TYPE_PIEMD:
            for (int i = run.start; i < end; ++i) {
                const HaloRec& h = halos[i];
#pragma unroll
                for (int k = 0; k < PX; ++k) grad_piemd<false>(h, px[k], py, sx[k], sy[k]);
            }
            break;
        default:  // TYPE_SIS
            if (run.circ) {
                // four halos per step: their records' scalar loads issue together, one wait per step
#pragma unroll 4
                for (int i = run.start; i < end; ++i) {
                    const HaloRec& h = halos[i];
#pragma unroll
                    for (int k = 0; k < PX; ++k) grad_sis_circ(h, px[k], py, sx[k], sy[k]);
                }
            } else {
                for (int i = run.start; i < end; ++i) {
                    const HaloRec& h = halos[i];
#pragma unroll
                    for (int k = 0; k < PX; ++k) grad_sis(h, px[k], py, sx[k], sy[k]);
                }
            }
            break;
        }
#pragma unroll
        for (int k = 0; k < PX; ++k) {
            tx[k] += sx[k];
            ty[k] += sy[k];
        }
    }
    const double m = o.kind == LT_MAP_DPL ? o.r : 1.0;
#pragma unroll
    for (int k = 0; k < PX; ++k) {
        const int col = col0 + k * WAVE;
        if (col < g.nx) {
            const int64_t idx = (int64_t)row * g.pitch + col;
            o.p0[idx] = tx[k] * m;
            o.p1[idx] = ty[k] * m;
        }
    }
}

template <int PX>
__global__ __launch_bounds__(WAVE* ROWS) void k_potential64(const HaloRec* __restrict__ halos,
                                                            const Run* __restrict__ runs, int nruns,
                                                            GridDesc g, OutDesc o)
{
    const int row = blockIdx.y * ROWS + threadIdx.y;
    if (row >= g.ny) return;
    halos = select_sample(g, halos, o);
    const int col0 = blockIdx.x * (WAVE * PX) + threadIdx.x;
    const double py = pixel_coord(g.ymin, row + g.jstart, g.dy);
    double px[PX], tot[PX], mn[PX];
#pragma unroll
    for (int k = 0; k < PX; ++k) {
        px[k] = pixel_coord(g.xmin, col0 + k * WAVE + g.istart, g.dx);
        tot[k] = 0.;
        mn[k] = 1.;
    }
    for (int r = 0; r < nruns; ++r) {
        const Run run = runs[r];
        double s[PX];
#pragma unroll
        for (int k = 0; k < PX; ++k) s[k] = 0.;
        for (int i = run.start; i < run.start + run.count; ++i) {
            const HaloRec& h = halos[i];
#pragma unroll
            for (int k = 0; k < PX; ++k) s[k] += pot_halo(h, run.type, px[k], py, mn[k]);
        }
#pragma unroll
        for (int k = 0; k < PX; ++k) tot[k] += s[k];
    }
#pragma unroll
    for (int k = 0; k < PX; ++k) {
        const int col = col0 + k * WAVE;
        // a pixel on a PIEMD centre: the reference's potential is 0/0 there
        const double v = mn[k] == 0. ? __builtin_nan("") : tot[k] * o.mult;
        if (col < g.nx) o.p0[(int64_t)row * g.pitch + col] = v;
    }
}

// Gradient at arbitrary points for a batch of lens models (blockIdx.y = sample), e.g. at the
// observed images for the source barycentres. Same per-run accumulation as the grid kernel.
__global__ __launch_bounds__(256) void k_points_gradient64(const HaloRec* __restrict__ halos,
                                                           const Run* __restrict__ runs, int nruns,
                                                           const double* __restrict__ px,
                                                           const double* __restrict__ py, int npts,
                                                           int64_t halo_stride, double* __restrict__ gx,
                                                           double* __restrict__ gy)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t s = blockIdx.y;
    if (i >= npts) return;
    halos += s * halo_stride;
    const double x = px[i], y = py[i];
    double tx = 0., ty = 0.;
    for (int r = 0; r < nruns; ++r) {
        const Run run = runs[r];
        double sx = 0., sy = 0.;
        const int end = run.start + run.count;
        switch (run.type) {
        case TYPE_DPIE:
            for (int h = run.start; h < end; ++h) grad_piemd<true>(halos[h], x, y, sx, sy);
            break;
        case TYPE_PIEMD:
            for (int h = run.start; h < end; ++h) grad_piemd<false>(halos[h], x, y, sx, sy);
            break;
        default:
            if (run.circ)
                for (int h = run.start; h < end; ++h) grad_sis_circ(halos[h], x, y, sx, sy);
            else
                for (int h = run.start; h < end; ++h) grad_sis(halos[h], x, y, sx, sy);
            break;
        }
        tx += sx;
        ty += sy;
    }
    gx[s * npts + i] = tx;
    gy[s * npts + i] = ty;
}

// Epilogue alone, on Hessian planes already in device memory (the reference's map functions
// amplif_k_grid_CPU_GPU & co. work this way, src/grid_map_ampli_GPU.cu:305-547).
template <int KIND>
__global__ __launch_bounds__(256) void k_epilogue(const double* __restrict__ a, const double* __restrict__ b,
                                                  const double* __restrict__ c, int64_t n, OutDesc o)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const double av = a[i], cv = c[i];
        hess_epilogue<KIND>(o, i, av + cv, av - cv, b[i]);
    }
}

// ------------------------------------------------------------------------------------------------
// launchers
// ------------------------------------------------------------------------------------------------
template <int PX>
static hipError_t launch_hessian64_px(const HaloRec* halos, const Run* runs, int nruns, const GridDesc& g,
                                      const OutDesc& o, hipStream_t s)
{
    dim3 blk(WAVE, ROWS, 1), grd = grid_for(g, PX);
    switch (o.kind) {
    case LT_MAP_FIELDS: hipLaunchKernelGGL((k_hessian64<PX, LT_MAP_FIELDS>), grd, blk, 0, s, halos, runs, nruns, g, o); break;
    case LT_MAP_AMPLI: hipLaunchKernelGGL((k_hessian64<PX, LT_MAP_AMPLI>), grd, blk, 0, s, halos, runs, nruns, g, o); break;
    case LT_MAP_SHEAR: hipLaunchKernelGGL((k_hessian64<PX, LT_MAP_SHEAR>), grd, blk, 0, s, halos, runs, nruns, g, o); break;
    case LT_MAP_MASS: hipLaunchKernelGGL((k_hessian64<PX, LT_MAP_MASS>), grd, blk, 0, s, halos, runs, nruns, g, o); break;
    default: hipLaunchKernelGGL((k_hessian64<PX, LT_MAP_HESSIAN>), grd, blk, 0, s, halos, runs, nruns, g, o); break;
    }
    return hipGetLastError();
}

hipError_t launch_hessian64(const HaloRec* halos, const Run* runs, int nruns, const GridDesc& g,
                            const OutDesc& o, int px, hipStream_t s)
{
    switch (px) {
    case 1: return launch_hessian64_px<1>(halos, runs, nruns, g, o, s);
    case 4: return launch_hessian64_px<4>(halos, runs, nruns, g, o, s);
    default: return launch_hessian64_px<2>(halos, runs, nruns, g, o, s);
    }
}

hipError_t launch_epilogue64(const double* a, const double* b, const double* c, int64_t n, const OutDesc& o,
                             hipStream_t s)
{
    int64_t nb = (n + 255) / 256;
    unsigned blocks = (unsigned)(nb > 16384 ? 16384 : (nb ? nb : 1));
    dim3 blk(256), grd(blocks);
    switch (o.kind) {
    case LT_MAP_FIELDS: hipLaunchKernelGGL((k_epilogue<LT_MAP_FIELDS>), grd, blk, 0, s, a, b, c, n, o); break;
    case LT_MAP_AMPLI: hipLaunchKernelGGL((k_epilogue<LT_MAP_AMPLI>), grd, blk, 0, s, a, b, c, n, o); break;
    case LT_MAP_SHEAR: hipLaunchKernelGGL((k_epilogue<LT_MAP_SHEAR>), grd, blk, 0, s, a, b, c, n, o); break;
    case LT_MAP_MASS: hipLaunchKernelGGL((k_epilogue<LT_MAP_MASS>), grd, blk, 0, s, a, b, c, n, o); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_points_gradient64(const HaloRec* halos, const Run* runs, int nruns, const double* px,
                                   const double* py, int npts, int nsamples, int64_t halo_stride, double* gx,
                                   double* gy, hipStream_t s)
{
    hipLaunchKernelGGL(k_points_gradient64, dim3((npts + 255) / 256, nsamples), dim3(256), 0, s, halos, runs, nruns,
                       px, py, npts, halo_stride, gx, gy);
    return hipGetLastError();
}

hipError_t launch_gradient64(const HaloRec* halos, const Run* runs, int nruns, const GridDesc& g,
                             const OutDesc& o, int px, hipStream_t s)
{
    dim3 blk(WAVE, ROWS, 1);
    switch (px) {
    case 1: hipLaunchKernelGGL(k_gradient64<1>, grid_for(g, 1), blk, 0, s, halos, runs, nruns, g, o); break;
    case 4: hipLaunchKernelGGL(k_gradient64<4>, grid_for(g, 4), blk, 0, s, halos, runs, nruns, g, o); break;
    default: hipLaunchKernelGGL(k_gradient64<2>, grid_for(g, 2), blk, 0, s, halos, runs, nruns, g, o); break;
    }
    return hipGetLastError();
}

hipError_t launch_potential64(const HaloRec* halos, const Run* runs, int nruns, const GridDesc& g,
                              const OutDesc& o, int px, hipStream_t s)
{
    dim3 blk(WAVE, ROWS, 1);
    switch (px) {
    case 1: hipLaunchKernelGGL(k_potential64<1>, grid_for(g, 1), blk, 0, s, halos, runs, nruns, g, o); break;
    case 4: hipLaunchKernelGGL(k_potential64<4>, grid_for(g, 4), blk, 0, s, halos, runs, nruns, g, o); break;
    default: hipLaunchKernelGGL(k_potential64<2>, grid_for(g, 2), blk, 0, s, halos, runs, nruns, g, o); break;
    }
    return hipGetLastError();
}

}  // namespace lt
