// lt_kernels32.hip -- FP32-compensated Hessian (+ fused map epilogue) for gfx950.
//
// Same launch shape and run walk as the FP64 kernel (lt_kernels64.hip). Per (halo, pixel):
//   * the lab-frame offset pixel - centre is formed in FP64 and rounded once to FP32 (the FP32
//     offset is the exact offset to FP32 precision, however far the halo is from the origin);
//   * inside the halo's near radius (HaloRec32::near2: cores, where the term is large) the FP64 term
//     is used, outside it the FP32 term below;
//   * FP32 terms are summed in FP32 over chunks of CHUNK halos and each chunk sum is added to FP64
//     run sums (two-level compensated summation: the FP32 sum error is bounded by CHUNK u sum|t|).
//
// FP32 form of the PIEMD I_{0.5} Hessian. With s = sqrt(e), a = 1/(1+e), b = 1/(1-e), cx1 = a/b,
// w = sqrt(r^2 + a^2 x^2 + b^2 y^2), D = x + i (2 r s - y) and N = cx1 x + i (2 s w - y / cx1), the
// reference's closed form (src/gradient2.cpp:54-82) is the Hessian of Im/Re ln(N / D), whose two
// pieces N'/N and D'/D are both singular at D = 0, i.e. at halo-frame (0, 2 r s); the closed form
// cancels them there and in FP32 loses ~1e-5 relative over the whole upper half plane. Writing
// v = r cx1 + 2 s a b y, one has w^2 - v^2 = a^2 |D|^2 and N = cx1 D + 2 i s (w - v), hence
//   N / D = Q / S,  S = w + v,  Q = r + cx1 w + 2 i s a^2 x
// (Re Q = cx1 S + 2 s a^2 (2 r s - y) = r + cx1 w because cx1 b = a and a (1 + e) = 1). With
// U = b^2 y / w + 2 s a b and t = 2 s w - y / cx1 the three second derivatives are, up to the
// factor 2 s a^2 ci:
//   phi_xx ~ (Re Q - cx1 a^2 x^2 / w) / |Q|^2
//   phi_xy ~ x (2 s a^2 - cx1 U) / |Q|^2
//   phi_yy ~ (Re Q + U t) / |Q|^2
// Re Q >= r > 0, so |Q| never vanishes and nothing cancels near D = 0; one rsq and one rcp per core
// radius. (Derivation: differentiate ln Q - ln S with dw/dx = a^2 x / w, dw/dy = b^2 y / w,
// dv/dy = 2 s a b, and use |D|^2 / S = (w - v) / a^2, 2 s (w - v) + cx1 (2 r s - y) = t.)
#include <hip/hip_runtime.h>

#include "lenstool_hip/lt_capi.h"
#include "lenstool_hip/lt_status.h"
#include "lt_device.h"
#include "lt_kernels.h"
#include "lt_math.h"
#include "lt_terms64.h"

namespace lt {

constexpr int CHUNK = 8;

// The gradient kernel stages the FP32 records of a run through LDS in tiles of TILE records (8 KB),
// which the four waves of the work-group then read with broadcast LDS loads. On gfx950 an FP32 VALU
// instruction with an SGPR operand issues at half the rate of one whose operands are all VGPRs
// (measured on independent chains: 0.57 against 1.0-1.09 wave instructions per ns per SIMD; FP64
// issues at 0.57 either way), and records read from LDS arrive in VGPRs where scalar loads of a
// wavefront-uniform record leave them in SGPRs. TILE is a multiple of CHUNK, so the chunks (and the
// sums) are those of the untiled walk.
constexpr int TILE = 64;
static_assert(TILE % CHUNK == 0, "tiles hold whole chunks");

// n records recs[i0 .. i0+n) into tile[0 .. n), 16 bytes per thread and step; the caller brackets
// it with barriers (every thread of the work-group calls it with the same arguments)
template <typename Rec>
__device__ __forceinline__ void stage_tile(Rec* __restrict__ tile, const Rec* __restrict__ recs, int i0, int n)
{
    static_assert(sizeof(Rec) % 16 == 0, "records are whole 16-byte vectors");
    constexpr int V = sizeof(Rec) / 16;
    const float4* src = reinterpret_cast<const float4*>(recs + i0);
    float4* dst = reinterpret_cast<float4*>(tile);
    for (int j = threadIdx.y * WAVE + threadIdx.x; j < n * V; j += WAVE * ROWS) dst[j] = src[j];
}

// Record j of a tile by value, only its first NV 16-byte vectors read (the fields the FP32 terms
// use; the rest of the copy is left unset and never read): one set of wide LDS loads per halo at
// the top of its iteration, which both pixels' branches then share.
template <int NV, typename Rec>
__device__ __forceinline__ Rec hot_fields(const Rec* __restrict__ tile, int j)
{
    static_assert(NV * 16 <= (int)sizeof(Rec), "within the record");
    Rec r;
    const float4* src = reinterpret_cast<const float4*>(tile + j);
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        const float4 v = src[q];
        __builtin_memcpy(reinterpret_cast<char*>(&r) + 16 * q, &v, 16);
    }
    return r;
}

__device__ __forceinline__ float rsq32(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ float rcp32(float x) { return __builtin_amdgcn_rcpf(x); }

// Halo-frame coordinates of a pixel in FP32 (pixel and centre both as hi + lo floats).
struct Frame32 {
    float x, y;
};

// the row part (the same for every pixel of a thread): ty snx, ty csy, ty^2, ty
struct Row32 {
    float tys, tyc, ty2, ty;
};

__device__ __forceinline__ Row32 row32(const HaloRec32& h, double yh, double py)
{
    const float ty = (float)(py - yh);
    return Row32{ty * h.snx, ty * h.csy, ty * ty, ty};
}

// lab-frame x offset of a pixel from the halo centre: the FP64 difference rounded once to FP32
__device__ __forceinline__ float offset_x(double xh, double px)
{
    return (float)(px - xh);
}

// halo-frame coordinates, scaled for PIEMD records (X = cx1 a x, Y = cx1 b y; lt_terms64.h) and
// plain for SIS records (csx = csy = cos, snx = sny = sin)
__device__ __forceinline__ Frame32 frame32(const HaloRec32& h, const Row32& r, float tx)
{
    Frame32 f;
    f.x = __builtin_fmaf(tx, h.csx, r.tys);
    f.y = __builtin_fmaf(-tx, h.sny, r.tyc);
    return f;
}

// numerators of xx - yy and xy, their common factor 1 / |Q|^2, and the trace xx + yy = 1 / W (lt_terms64.h)
struct Q32 {
    float dd, xy, iq, iW;
};

// pixel-only quantities shared by the core and cut terms (the scaled form of lt_terms64.h)
struct QShared {
    float wq0, Qi2, A, Y, Qi, kx2, ky4;
};

__device__ __forceinline__ QShared q_shared(const HaloRec32& h, float X, float Y)
{
    QShared s;
    s.Y = Y;
    s.A = X * X;
    s.wq0 = __builtin_fmaf(Y, Y, s.A);
    s.Qi = h.k1 * X;
    s.Qi2 = s.Qi * s.Qi;
    s.kx2 = h.k2 * X;
    s.ky4 = h.k4 * Y;
    return s;
}

__device__ __forceinline__ Q32 q_term(const HaloRec32& h, const QShared& s, float r, float r2c)
{
    const float W2 = s.wq0 + r2c;
    const float iW = rsq32(W2);
    const float W = W2 * iW;
    const float Qr = W + r;
    const float U = __builtin_fmaf(s.Y, iW, h.k5);
    Q32 o;
    o.dd = -__builtin_fmaf(s.A, iW, U * __builtin_fmaf(h.k3, W, -s.ky4));
    o.xy = __builtin_fmaf(-s.kx2, U, s.Qi);
    o.iq = rcp32(__builtin_fmaf(Qr, Qr, s.Qi2));
    o.iW = iW;
    return o;
}

// add one halo-frame Hessian (trace T = xx + yy, difference D = xx - yy, off-diagonal B = xy), already
// differenced over the core radii, in the lab frame
__device__ __forceinline__ void acc_lab(const HaloRec32& h, float T, float D, float B, float& fT, float& fD,
                                        float& fB)
{
    fT = __builtin_fmaf(h.hT, T, fT);
    fD = __builtin_fmaf(h.hc2, D, __builtin_fmaf(h.hs2m2, B, fD));
    fB = __builtin_fmaf(h.hs2h, D, __builtin_fmaf(h.hc2, B, fB));
}

template <bool CUT>
__device__ __forceinline__ void hess_piemd32(const HaloRec32& h, const Frame32& f, float& fT, float& fD, float& fB)
{
    const QShared s = q_shared(h, f.x, f.y);
    const Q32 a = q_term(h, s, h.ra, h.ra2c);
    if (CUT) {
        const Q32 b = q_term(h, s, h.rb, h.rb2c);
        acc_lab(h, a.iW - b.iW, __builtin_fmaf(a.dd, a.iq, -(b.dd * b.iq)), __builtin_fmaf(a.xy, a.iq, -(b.xy * b.iq)),
                fT, fD, fB);
    } else {
        acc_lab(h, a.iW, a.dd * a.iq, a.xy * a.iq, fT, fD, fB);
    }
}

// SIS / SIE: halo-frame Hessian amph / R^3 (y^2, -xy, x^2), R^2 = (1-e) x^2 + (1+e) y^2
__device__ __forceinline__ void hess_sis32(const HaloRec32& h, const Frame32& f, float& fT, float& fD, float& fB)
{
    const float x2 = f.x * f.x, y2 = f.y * f.y;
    const float iR = rsq32(__builtin_fmaf(h.sa, x2, h.sb * y2));
    const float g = iR * iR * iR;
    acc_lab(h, g * (x2 + y2), g * (y2 - x2), -(g * f.x * f.y), fT, fD, fB);
}

// circular SIS (a run of them): the lab-frame offsets tx, ty give the Hessian directly (hT = b0)
__device__ __forceinline__ void hess_sis32_circ(const HaloRec32& h, float tx, float ty, float ty2, float& fT, float& fD,
                                                float& fB)
{
    const float x2 = tx * tx;
    const float iR = rsq32(x2 + ty2);
    const float ga = h.hT * (iR * iR * iR);
    fT = __builtin_fmaf(h.hT, iR, fT);
    fD = __builtin_fmaf(ga, ty2 - x2, fD);
    fB = __builtin_fmaf(-ga, tx * ty, fB);
}

// template tag of run32 for a run of circular SIS (not a reference type code)
constexpr int TYPE_SIS_CIRC = -5;

// For runs of SIS halos, the record fields their terms use, held in scalar registers from the top of
// the halo's iteration (one empty asm statement that reads them all), so that their scalar loads issue
// together and are waited for once; left to itself the compiler moves them into the branch after the
// near-core test, once per pixel, each with its own wait, and an SIS term is too short to hide that
// (Bayes maps of 1 dPIE + 10 circular SIS: 10.56 -> 10.41 ms). For PIEMD runs the same pinning
// measured slower (8.95 -> 9.35 ms on the 503-dPIE map): their terms are long enough.
template <int TYPE>
__device__ __forceinline__ void pin_record(const HaloRec32& h, double xh, double yh)
{
    if constexpr (TYPE == TYPE_SIS_CIRC) {
        asm volatile("" ::"s"(xh), "s"(yh), "s"(h.near2), "s"(h.hT));
    } else if constexpr (TYPE == TYPE_SIS) {
        asm volatile("" ::"s"(xh), "s"(yh), "s"(h.near2), "s"(h.hT), "s"(h.csx), "s"(h.snx), "s"(h.csy), "s"(h.sny),
                     "s"(h.hc2), "s"(h.hs2m2), "s"(h.hs2h), "s"(h.sa), "s"(h.sb));
    }
}

// One run of halos of one type: FP32 terms summed per chunk of CHUNK halos, FP64 terms (near
// lanes) straight into the FP64 run sums. The records are wavefront-uniform scalar loads (the
// Hessian keeps them in SGPRs: staged through LDS they cost 20 VGPRs, two waves of occupancy, and
// measured slower, 9.6 against 8.9 ms on the 503-halo 4096^2 map).
template <int TYPE, int PX>
__device__ __forceinline__ void run32(const HaloRec* __restrict__ h64, const HaloRec32* __restrict__ h32, int start,
                                      int end, const double* px, double py, double* sT, double* sD, double* sB)
{
    for (int c0 = start; c0 < end; c0 += CHUNK) {
        const int c1 = min(c0 + CHUNK, end);
        float fT[PX], fD[PX], fB[PX];
#pragma unroll
        for (int k = 0; k < PX; ++k) fT[k] = fD[k] = fB[k] = 0.f;
        for (int i = c0; i < c1; ++i) {
            const HaloRec32& h = h32[i];
            const double xh = h64[i].xh, yh = h64[i].yh;
            pin_record<TYPE>(h, xh, yh);
            const Row32 rw = row32(h, yh, py);
#pragma unroll
            for (int k = 0; k < PX; ++k) {
                const float tx = offset_x(xh, px[k]);
                // the near-core test on the unrotated offset (the distance does not depend on the frame)
                if (__builtin_fmaf(tx, tx, rw.ty2) < h.near2) {
                    const HaloRec& hd = h64[i];
                    if constexpr (TYPE == TYPE_DPIE) hess_piemd<true>(hd, px[k], py, sT[k], sD[k], sB[k]);
                    else if constexpr (TYPE == TYPE_PIEMD) hess_piemd<false>(hd, px[k], py, sT[k], sD[k], sB[k]);
                    else if constexpr (TYPE == TYPE_SIS_CIRC) hess_sis_circ(hd, px[k], py, sT[k], sD[k], sB[k]);
                    else hess_sis(hd, px[k], py, sT[k], sD[k], sB[k]);
                } else if constexpr (TYPE == TYPE_SIS_CIRC) {
                    hess_sis32_circ(h, tx, rw.ty, rw.ty2, fT[k], fD[k], fB[k]);
                } else {
                    const Frame32 f = frame32(h, rw, tx);
                    if constexpr (TYPE == TYPE_DPIE) hess_piemd32<true>(h, f, fT[k], fD[k], fB[k]);
                    else if constexpr (TYPE == TYPE_PIEMD) hess_piemd32<false>(h, f, fT[k], fD[k], fB[k]);
                    else hess_sis32(h, f, fT[k], fD[k], fB[k]);
                }
            }
        }
#pragma unroll
        for (int k = 0; k < PX; ++k) {
            sT[k] += (double)fT[k];
            sD[k] += (double)fD[k];
            sB[k] += (double)fB[k];
        }
    }
}

// The FP64 Hessian sums of one pixel over every run, exactly as the FP64 kernel forms them (same
// terms, same order): the fallback for pixels where det A would amplify the FP32 rounding.
__device__ __forceinline__ void hessian_fp64_pixel(const HaloRec* __restrict__ h64, const Run* __restrict__ runs,
                                                int nruns, double px, double py, double& tT, double& tD, double& tB)
{
    tT = tD = tB = 0.;
    for (int r = 0; r < nruns; ++r) {
        const Run run = runs[r];
        double sT = 0., sD = 0., sB = 0.;
        const int end = run.start + run.count;
        for (int i = run.start; i < end; ++i) {
            const HaloRec& h = h64[i];
            if (run.type == TYPE_DPIE) hess_piemd<true>(h, px, py, sT, sD, sB);
            else if (run.type == TYPE_PIEMD) hess_piemd<false>(h, px, py, sT, sD, sB);
            else if (run.type == TYPE_SIS && run.circ) hess_sis_circ(h, px, py, sT, sD, sB);
            else if (run.type == TYPE_SIS) hess_sis(h, px, py, sT, sD, sB);
            else {
                sT += h.h14a + h.h14c;
                sD += h.h14a - h.h14c;
                sB += h.h14b;
            }
        }
        tT = __builtin_fma(run.weight, sT, tT);
        tD = __builtin_fma(run.weight, sD, tD);
        tB = __builtin_fma(run.weight, sB, tB);
    }
}

// det A = (1 - kappa)^2 - gamma^2 turns an error e in kappa and gamma into up to
// 2 (|1 - kappa| + gamma) e: the FP32 terms leave e below about 4e-8, so where |1 - kappa| + gamma
// exceeds this bound (next to singular or small cores: kappa reaches hundreds at a tenth of an
// arcsecond from an isothermal centre) the pixel's sums are recomputed in FP64, keeping
// |det A (FP32C) - det A (FP64)| under 1e-6 everywhere. A few pixels per halo core pay it.
constexpr double kDetGuard = 8.0;

template <int PX, int KIND>
__global__ __launch_bounds__(WAVE* ROWS) void k_hessian32(const HaloRec* __restrict__ h64,
                                                          const HaloRec32* __restrict__ h32,
                                                          const Run* __restrict__ runs, int nruns,
                                                          GridDesc g, OutDesc o)
{
    const int row = blockIdx.y * ROWS + threadIdx.y;
    if (row >= g.ny) return;
    h64 = select_sample(g, h64, o);
    h32 += (int64_t)blockIdx.z * g.halo_stride;
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
        if (run.type == TYPE_SHEAR) {
            for (int i = run.start; i < end; ++i) {
                const HaloRec& h = h64[i];
#pragma unroll
                for (int k = 0; k < PX; ++k) {
                    sT[k] += h.h14a + h.h14c;
                    sD[k] += h.h14a - h.h14c;
                    sB[k] += h.h14b;
                }
            }
        } else if (run.type == TYPE_DPIE) {
            run32<TYPE_DPIE, PX>(h64, h32, run.start, end, px, py, sT, sD, sB);
        } else if (run.type == TYPE_PIEMD) {
            run32<TYPE_PIEMD, PX>(h64, h32, run.start, end, px, py, sT, sD, sB);
        } else if (run.circ) {
            run32<TYPE_SIS_CIRC, PX>(h64, h32, run.start, end, px, py, sT, sD, sB);
        } else {
            run32<TYPE_SIS, PX>(h64, h32, run.start, end, px, py, sT, sD, sB);
        }
        const double w = run.weight;
#pragma unroll
        for (int k = 0; k < PX; ++k) {
            tT[k] = __builtin_fma(w, sT[k], tT[k]);
            tD[k] = __builtin_fma(w, sD[k], tD[k]);
            tB[k] = __builtin_fma(w, sB[k], tB[k]);
        }
    }
    if constexpr (KIND == LT_MAP_FIELDS || KIND == LT_MAP_AMPLI || KIND == LT_MAP_SHEAR) {
        // one pixel at a time (not unrolled) so that the FP64 loop adds no registers to the kernel; with
        // no runs there is nothing to recompute
        for (int k = 0; k < PX && nruns > 0; ++k) {
            const double kap = 0.5 * tT[k] * o.r;
            const double gb = 0.5 * (__builtin_fabs(tD[k]) + 2. * __builtin_fabs(tB[k])) * __builtin_fabs(o.r);
            if (__builtin_fabs(1. - kap) + gb > kDetGuard) hessian_fp64_pixel(h64, runs, nruns, px[k], py, tT[k], tD[k], tB[k]);
        }
    }
#pragma unroll
    for (int k = 0; k < PX; ++k) {
        const int col = col0 + k * WAVE;
        if (col < g.nx) hess_epilogue<KIND>(o, (int64_t)row * g.pitch + col, tT[k], tD[k], tB[k]);
    }
}

template <int PX>
static hipError_t launch_hessian32_px(const HaloRec* h64, const HaloRec32* h32, const Run* runs, int nruns,
                                      const GridDesc& g, const OutDesc& o, hipStream_t s)
{
    dim3 blk(WAVE, ROWS, 1), grd = grid_for(g, PX);
    switch (o.kind) {
    case LT_MAP_FIELDS: hipLaunchKernelGGL((k_hessian32<PX, LT_MAP_FIELDS>), grd, blk, 0, s, h64, h32, runs, nruns, g, o); break;
    case LT_MAP_AMPLI: hipLaunchKernelGGL((k_hessian32<PX, LT_MAP_AMPLI>), grd, blk, 0, s, h64, h32, runs, nruns, g, o); break;
    case LT_MAP_SHEAR: hipLaunchKernelGGL((k_hessian32<PX, LT_MAP_SHEAR>), grd, blk, 0, s, h64, h32, runs, nruns, g, o); break;
    case LT_MAP_MASS: hipLaunchKernelGGL((k_hessian32<PX, LT_MAP_MASS>), grd, blk, 0, s, h64, h32, runs, nruns, g, o); break;
    default: hipLaunchKernelGGL((k_hessian32<PX, LT_MAP_HESSIAN>), grd, blk, 0, s, h64, h32, runs, nruns, g, o); break;
    }
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------------------
// The same FP32C Hessian with lanes over halos (launch shape LT_PX_HALO_LANES): a wavefront takes a
// row segment of SEG pixels and walks each run 64 halos at a time, lane l holding halo c0 + l, so
// every operand of the FP32 terms is a VGPR (on gfx950 an FP32 instruction with an SGPR operand
// issues at half the rate of an all-VGPR one). Per lane and row, the halo-frame coordinates of the
// segment's first pixel (X0, Y0: the FP64 offset rounded once to FP32, then rotated and scaled) and
// their step per pixel (dX, dY) are formed once; pixel j is at (X0 + j dX, Y0 + j dY). A (halo,
// pixel) pair is near when X^2 + Y^2 < near2s (a region holding the disk of the halo's near radius):
// near pairs are masked out of the FP32 sums by a select, and a second pass, lane j = pixel j, adds
// their FP64 terms for the halos whose near region reaches the segment, evaluating the same FP32
// test with the same operations, so each pair is taken exactly once. Each lane keeps FP32 sums of
// its halos' terms for the SEG pixels (at most FLUSH halos per sum, as the chunks of k_hessian32);
// they are summed over the 64 lanes in FP64 through LDS into the FP64 run sums of lane j = pixel j.
// Same epilogue and det-A guard as k_hessian32.
// ------------------------------------------------------------------------------------------------
constexpr int SEG = 16;
constexpr int FLUSH = 8;  // chunks of 64 halos between two FP64 flushes: <= 8 FP32 terms per lane sum
static_assert(WAVE == 4 * SEG, "the flush adds four quarters of the wavefront");

// one lane's halo over the segment: halo-frame coordinates of pixel 0 and their step per pixel (for
// a circular SIS: the lab-frame x offset and its step, the y offset and its square)
struct Seg32 {
    float X0, Y0, dX, dY, ty2;
};

template <int TYPE>
__device__ __forceinline__ Seg32 seg32(const HaloRec32& h, float tx0, float ty, float dxf)
{
    Seg32 s;
    if constexpr (TYPE == TYPE_SIS_CIRC) {
        s.X0 = tx0;
        s.Y0 = ty;
        s.dX = dxf;
        s.dY = 0.f;
        s.ty2 = ty * ty;
    } else {
        s.X0 = __builtin_fmaf(tx0, h.csx, ty * h.snx);
        s.Y0 = __builtin_fmaf(-tx0, h.sny, ty * h.csy);
        s.dX = dxf * h.csx;
        s.dY = -(dxf * h.sny);
        s.ty2 = 0.f;
    }
    return s;
}

// pixel J of the segment and its near metric: the same operations in the FP32 pass and the near pass
template <int TYPE>
__device__ __forceinline__ float seg_x(const Seg32& s, float J)
{
    return __builtin_fmaf(J, s.dX, s.X0);
}

template <int TYPE>
__device__ __forceinline__ float seg_y(const Seg32& s, float J)
{
    if constexpr (TYPE == TYPE_SIS_CIRC) return s.Y0;
    else return __builtin_fmaf(J, s.dY, s.Y0);
}

template <int TYPE>
__device__ __forceinline__ float near_metric(const Seg32& s, float X, float Y)
{
    if constexpr (TYPE == TYPE_SIS_CIRC) return __builtin_fmaf(X, X, s.ty2);
    else return __builtin_fmaf(Y, Y, X * X);
}

// the FP32 term of a lane's halo at pixel (X, Y) into the lane sums, in the lab frame with the halo's
// amplitudes; a near pair adds exactly zero (the select drops whatever the FP32 form gave there)
template <int TYPE>
__device__ __forceinline__ void add_term32(const HaloRec32& h, const Seg32& s, float X, float Y, bool far, float& aT,
                                           float& aD, float& aB)
{
    if constexpr (TYPE == TYPE_SIS_CIRC) {
        const float x2 = X * X;
        const float iR = far ? rsq32(x2 + s.ty2) : 0.f;
        const float ga = h.hT * (iR * iR * iR);
        aT = __builtin_fmaf(h.hT, iR, aT);
        aD = __builtin_fmaf(ga, s.ty2 - x2, aD);
        aB = __builtin_fmaf(-ga, X * Y, aB);
    } else if constexpr (TYPE == TYPE_SIS) {
        const float x2 = X * X, y2 = Y * Y;
        const float iR = rsq32(__builtin_fmaf(h.sa, x2, h.sb * y2));
        const float g = far ? iR * iR * iR : 0.f;
        acc_lab(h, g * (x2 + y2), g * (y2 - x2), -(g * X * Y), aT, aD, aB);
    } else {
        const QShared q = q_shared(h, X, Y);
        const Q32 a = q_term(h, q, h.ra, h.ra2c);
        float T, D, B;
        if constexpr (TYPE == TYPE_DPIE) {
            const Q32 b = q_term(h, q, h.rb, h.rb2c);
            T = a.iW - b.iW;
            D = __builtin_fmaf(a.dd, a.iq, -(b.dd * b.iq));
            B = __builtin_fmaf(a.xy, a.iq, -(b.xy * b.iq));
        } else {
            T = a.iW;
            D = a.dd * a.iq;
            B = a.xy * a.iq;
        }
        acc_lab(h, far ? T : 0.f, far ? D : 0.f, far ? B : 0.f, aT, aD, aB);
    }
}

template <int TYPE>
__device__ __forceinline__ void term64(const HaloRec& h, double px, double py, double& sT, double& sD, double& sB)
{
    if constexpr (TYPE == TYPE_DPIE) hess_piemd<true>(h, px, py, sT, sD, sB);
    else if constexpr (TYPE == TYPE_PIEMD) hess_piemd<false>(h, px, py, sT, sD, sB);
    else if constexpr (TYPE == TYPE_SIS_CIRC) hess_sis_circ(h, px, py, sT, sD, sB);
    else hess_sis(h, px, py, sT, sD, sB);
}

__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// sum over the 64 lanes of one component's lane sums a[j] (j < SEG) into pixel lane j's FP64 sum s,
// through this wavefront's LDS slice (layout [source lane][pixel]): lane (p = lane % SEG,
// q = lane / SEG) adds the sums of source lanes 16 q .. 16 q + 15 for pixel p in FP64, then the four
// quarters are added (q + 2 into q, then 1 into 0). A fixed order: the same sums on every run.
__device__ __forceinline__ void flush_component(float* __restrict__ lds, const float (&a)[SEG], int lane, double& s)
{
    float4* w = reinterpret_cast<float4*>(lds + lane * SEG);
#pragma unroll
    for (int k = 0; k < SEG / 4; ++k) w[k] = make_float4(a[4 * k], a[4 * k + 1], a[4 * k + 2], a[4 * k + 3]);
    wave_sync();
    const int p = lane & (SEG - 1), q = lane / SEG;
    double part = 0.;
#pragma unroll
    for (int l = 0; l < SEG; ++l) part += (double)lds[(q * SEG + l) * SEG + p];
    wave_sync();
    part += __shfl_down(part, 2 * SEG);
    part += __shfl_down(part, SEG);
    if (lane < SEG) s += part;
}

__device__ __forceinline__ void flush_lanes(float* __restrict__ lds, float (&aT)[SEG], float (&aD)[SEG],
                                            float (&aB)[SEG], int lane, double& sT, double& sD, double& sB)
{
    flush_component(lds, aT, lane, sT);
    flush_component(lds, aD, lane, sD);
    flush_component(lds, aB, lane, sB);
#pragma unroll
    for (int j = 0; j < SEG; ++j) aT[j] = aD[j] = aB[j] = 0.f;
}

// the segment: pixel 0 at px0, pixels 0 .. SEG-1 spanning [xa, xb] (FP64), step dxf (FP32); lane j
// (j < SEG) owns pixel j at pxl
struct SegPix {
    double px0, xa, xb, pxl, py, slack;
    float dxf;
};

template <int TYPE>
__device__ __forceinline__ void run32h(const HaloRec* __restrict__ h64, const HaloRec32* __restrict__ h32, int start,
                                       int end, const SegPix& sp, int lane, float* __restrict__ lds, double& sT,
                                       double& sD, double& sB)
{
    float aT[SEG], aD[SEG], aB[SEG];
#pragma unroll
    for (int j = 0; j < SEG; ++j) aT[j] = aD[j] = aB[j] = 0.f;
    int pending = 0;
    for (int c0 = start; c0 < end; c0 += WAVE) {
        const int i = c0 + lane;
        const bool live = i < end;
        HaloRec32 h{};
        double xh = 0., yh = 0.;
        if (live) {
            h = h32[i];
            xh = h64[i].xh;
            yh = h64[i].yh;
        } else {
            h.near2s = 3.0e38f;  // every pair near: the lane adds nothing and is never a candidate
        }
        const Seg32 sg = seg32<TYPE>(h, (float)(sp.px0 - xh), (float)(sp.py - yh), sp.dxf);
        // does the halo's near region (inside the disk of radius nearr) reach the segment? FP64, with
        // a margin well above the FP32 rounding of the near test
        const double dxs = __builtin_fmax(0., __builtin_fmax(sp.xa - xh, xh - sp.xb));
        const double dys = sp.py - yh;
        const double rr = (double)h.nearr * 1.01 + sp.slack;
        const bool cand = live && __builtin_fma(dxs, dxs, dys * dys) < rr * rr;
#pragma unroll
        for (int j = 0; j < SEG; ++j) {
            const float J = (float)j;
            const float X = seg_x<TYPE>(sg, J), Y = seg_y<TYPE>(sg, J);
            const bool far = !(near_metric<TYPE>(sg, X, Y) < h.near2s);
            add_term32<TYPE>(h, sg, X, Y, far, aT[j], aD[j], aB[j]);
        }
        if (++pending == FLUSH) {
            flush_lanes(lds, aT, aD, aB, lane, sT, sD, sB);
            pending = 0;
        }
        // FP64 terms of the near pairs, candidate halos in lane order, lane j = pixel j
        uint64_t mask = __builtin_amdgcn_ballot_w64(cand);
        while (mask) {
            const int c = __builtin_ctzll(mask);
            mask &= mask - 1;
            Seg32 cs;
            cs.X0 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, sg.X0), c));
            cs.Y0 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, sg.Y0), c));
            cs.dX = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, sg.dX), c));
            cs.dY = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, sg.dY), c));
            cs.ty2 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, sg.ty2), c));
            const float cn2 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, h.near2s), c));
            const float J = (float)lane;
            const float X = seg_x<TYPE>(cs, J), Y = seg_y<TYPE>(cs, J);
            if (lane < SEG && near_metric<TYPE>(cs, X, Y) < cn2) term64<TYPE>(h64[c0 + c], sp.pxl, sp.py, sT, sD, sB);
        }
    }
    if (pending) flush_lanes(lds, aT, aD, aB, lane, sT, sD, sB);
}

template <int KIND>
__global__ __launch_bounds__(WAVE* ROWS) void k_hessian32h(const HaloRec* __restrict__ h64,
                                                           const HaloRec32* __restrict__ h32,
                                                           const Run* __restrict__ runs, int nruns, GridDesc g,
                                                           OutDesc o)
{
    __shared__ __align__(16) float lds_all[ROWS][WAVE * SEG];
    const int row = blockIdx.y * ROWS + threadIdx.y;
    if (row >= g.ny) return;  // the whole wavefront: no work-group barrier is used
    float* lds = lds_all[threadIdx.y];
    h64 = select_sample(g, h64, o);
    h32 += (int64_t)blockIdx.z * g.halo_stride;
    const int lane = threadIdx.x;
    const int col0 = blockIdx.x * SEG;
    SegPix sp;
    sp.py = pixel_coord(g.ymin, row + g.jstart, g.dy);
    sp.px0 = pixel_coord(g.xmin, col0 + g.istart, g.dx);
    const double pxe = pixel_coord(g.xmin, col0 + SEG - 1 + g.istart, g.dx);
    sp.xa = __builtin_fmin(sp.px0, pxe);
    sp.xb = __builtin_fmax(sp.px0, pxe);
    sp.pxl = pixel_coord(g.xmin, col0 + (lane < SEG ? lane : 0) + g.istart, g.dx);
    sp.dxf = (float)g.dx;
    // the FP32 rounding of the near test moves a pixel by at most about 2^-22 (2 d + 2 SEG |dx|), d its
    // distance to the centre: the candidate test's 1 % and this slack cover it
    sp.slack = 1e-4 + 1e-6 * (double)SEG * __builtin_fabs(g.dx);
    double tT = 0., tD = 0., tB = 0.;
    for (int r = 0; r < nruns; ++r) {
        const Run run = runs[r];
        double sT = 0., sD = 0., sB = 0.;
        const int end = run.start + run.count;
        if (run.type == TYPE_SHEAR) {
            for (int i = run.start; i < end; ++i) {
                const HaloRec& h = h64[i];
                sT += h.h14a + h.h14c;
                sD += h.h14a - h.h14c;
                sB += h.h14b;
            }
        } else if (run.type == TYPE_DPIE) {
            run32h<TYPE_DPIE>(h64, h32, run.start, end, sp, lane, lds, sT, sD, sB);
        } else if (run.type == TYPE_PIEMD) {
            run32h<TYPE_PIEMD>(h64, h32, run.start, end, sp, lane, lds, sT, sD, sB);
        } else if (run.circ) {
            run32h<TYPE_SIS_CIRC>(h64, h32, run.start, end, sp, lane, lds, sT, sD, sB);
        } else {
            run32h<TYPE_SIS>(h64, h32, run.start, end, sp, lane, lds, sT, sD, sB);
        }
        tT = __builtin_fma(run.weight, sT, tT);
        tD = __builtin_fma(run.weight, sD, tD);
        tB = __builtin_fma(run.weight, sB, tB);
    }
    const int col = col0 + lane;
    if (lane < SEG && col < g.nx) {
        if constexpr (KIND == LT_MAP_FIELDS || KIND == LT_MAP_AMPLI || KIND == LT_MAP_SHEAR) {
            const double kap = 0.5 * tT * o.r;
            const double gb = 0.5 * (__builtin_fabs(tD) + 2. * __builtin_fabs(tB)) * __builtin_fabs(o.r);
            if (nruns > 0 && __builtin_fabs(1. - kap) + gb > kDetGuard) hessian_fp64_pixel(h64, runs, nruns, sp.pxl, sp.py, tT, tD, tB);
        }
        hess_epilogue<KIND>(o, (int64_t)row * g.pitch + col, tT, tD, tB);
    }
}

static hipError_t launch_hessian32h(const HaloRec* h64, const HaloRec32* h32, const Run* runs, int nruns,
                                    const GridDesc& g, const OutDesc& o, hipStream_t s)
{
    dim3 blk(WAVE, ROWS, 1);
    dim3 grd((unsigned)((g.nx + SEG - 1) / SEG), (unsigned)((g.ny + ROWS - 1) / ROWS),
             (unsigned)(g.nsamples > 0 ? g.nsamples : 1));
    switch (o.kind) {
    case LT_MAP_FIELDS: hipLaunchKernelGGL((k_hessian32h<LT_MAP_FIELDS>), grd, blk, 0, s, h64, h32, runs, nruns, g, o); break;
    case LT_MAP_AMPLI: hipLaunchKernelGGL((k_hessian32h<LT_MAP_AMPLI>), grd, blk, 0, s, h64, h32, runs, nruns, g, o); break;
    case LT_MAP_SHEAR: hipLaunchKernelGGL((k_hessian32h<LT_MAP_SHEAR>), grd, blk, 0, s, h64, h32, runs, nruns, g, o); break;
    case LT_MAP_MASS: hipLaunchKernelGGL((k_hessian32h<LT_MAP_MASS>), grd, blk, 0, s, h64, h32, runs, nruns, g, o); break;
    default: hipLaunchKernelGGL((k_hessian32h<LT_MAP_HESSIAN>), grd, blk, 0, s, h64, h32, runs, nruns, g, o); break;
    }
    return hipGetLastError();
}

hipError_t launch_hessian32(const HaloRec* h64, const HaloRec32* h32, const Run* runs, int nruns, const GridDesc& g,
                            const OutDesc& o, int px, hipStream_t s)
{
    switch (px) {
    case 1: return launch_hessian32_px<1>(h64, h32, runs, nruns, g, o, s);
    case 4: return launch_hessian32_px<4>(h64, h32, runs, nruns, g, o, s);
    case LT_PX_HALO_LANES: return launch_hessian32h(h64, h32, runs, nruns, g, o, s);
    default: return launch_hessian32_px<2>(h64, h32, runs, nruns, g, o, s);
    }
}

// ------------------------------------------------------------------------------------------------
// FP32-compensated gradient (gradient and displacement maps): the same scheme as the Hessian -- the
// offset pixel - centre formed in FP64 and rounded once, FP64 terms inside each halo's near radius,
// FP32 terms summed per chunk of CHUNK halos into FP64 run sums -- with single-precision arctangent
// and logarithm (lt_math.h atan2f_lt, logf_lt).
//
// The FP64 gradient (lt_terms64.h grad_piemd) takes ln of the I_{0.5} ratio N/D with
// N = cx1 x + i (2 s w - y / cx1), D = x + i (2 r s - y) (s = sqrt(e)); both vanish at the halo-frame
// point (0, 2 r s), and next to it N and D are differences of nearly equal numbers, which FP32 cannot
// carry (for the cut radius of a dPIE that point lies tens of arcseconds out). With a = 1/(1+e),
// b = 1/(1-e), w = sqrt(r^2 + a^2 x^2 + b^2 y^2) the same ratio is N/D = Q/S with
//   Q = r + cx1 w + 2 i s a^2 x,   S = w + r cx1 + 2 s a b y > 0
// (w^2 - v^2 = a^2 |D|^2 for v = r cx1 + 2 s a b y, as in the Hessian's division-free form), whose
// parts have no such cancellation. So arg(N/D) = arg Q, ln|N/D| = ln|Q| - ln S, and for the dPIE
//   arg(Za / Zb) = atan2(I (Rb - Ra), Ra Rb + I^2),  ln|Za / Zb| = ln(|Qa|^2 Sb^2 / (|Qb|^2 Sa^2)) / 2
// with R = Re Q and I = Im Q (the same for both radii).
// ------------------------------------------------------------------------------------------------
struct GFrame32 {
    float x, y, x2, y2;
};

__device__ __forceinline__ GFrame32 gframe32(const GradRec32& h, float tx, float ty)
{
    GFrame32 f;
    f.x = __builtin_fmaf(tx, h.cs, ty * h.sn);
    f.y = __builtin_fmaf(ty, h.cs, -(tx * h.sn));
    f.x2 = f.x * f.x;
    f.y2 = f.y * f.y;
    return f;
}

// R = Re Q and |Q|^2, S of one core radius
struct QS32 {
    float R, Q2, S;
};

__device__ __forceinline__ QS32 qs32(const GradRec32& h, float rem2, float I2, float vy, float r, float r2, float rc)
{
    const float q = r2 + rem2;
    const float w = q * rsq32(q);
    QS32 o;
    o.R = __builtin_fmaf(h.cx1, w, r);
    o.Q2 = __builtin_fmaf(o.R, o.R, I2);
    o.S = w + (rc + vy);
    return o;
}

template <bool CUT>
__device__ __forceinline__ void grad_piemd32(const GradRec32& h, float tx, float ty, float& gx, float& gy)
{
    const GFrame32 f = gframe32(h, tx, ty);
    const float rem2 = __builtin_fmaf(f.x2, h.kx, f.y2 * h.ky);
    const float I = h.ki * f.x;
    const float I2 = I * I;
    const float vy = h.kv * f.y;
    const QS32 a = qs32(h, rem2, I2, vy, h.ra, h.ra2, h.rca);
    float arg, ratio;
    if (CUT) {
        const QS32 b = qs32(h, rem2, I2, vy, h.rb, h.rb2, h.rcb);
        arg = atan2f_lt(I * (b.R - a.R), __builtin_fmaf(a.R, b.R, I2));
        ratio = (a.Q2 * (b.S * b.S)) * rcp32(b.Q2 * (a.S * a.S));
    } else {
        arg = atan2f_lt(I, a.R);
        ratio = a.Q2 * rcp32(a.S * a.S);
    }
    const float lnr = 0.5f * logf_lt(ratio);
    const float re = h.ampg * arg, im = -h.ampg * lnr;
    gx += __builtin_fmaf(re, h.cs, -(im * h.sn));
    gy += __builtin_fmaf(im, h.cs, re * h.sn);
}

__device__ __forceinline__ void grad_sis32(const GradRec32& h, float tx, float ty, float& gx, float& gy)
{
    const GFrame32 f = gframe32(h, tx, ty);
    const float s = h.ampg * rsq32(__builtin_fmaf(h.sa, f.x2, h.sb * f.y2));
    const float rx = h.sa * f.x * s, ry = h.sb * f.y * s;
    gx += __builtin_fmaf(rx, h.cs, -(ry * h.sn));
    gy += __builtin_fmaf(ry, h.cs, rx * h.sn);
}

__device__ __forceinline__ void grad_sis32_circ(const GradRec32& h, float tx, float ty, float& gx, float& gy)
{
    const float s = h.ampg * rsq32(__builtin_fmaf(tx, tx, ty * ty));
    gx = __builtin_fmaf(s, tx, gx);
    gy = __builtin_fmaf(s, ty, gy);
}

// Halos [t0, t1) of a run, their FP32 records in tile[0 .. t1 - t0): FP32 terms summed per chunk of
// CHUNK halos, FP64 terms (near lanes) straight into the FP64 run sums
template <int TYPE, int PX>
__device__ __forceinline__ void tileg32(const HaloRec* __restrict__ h64, const GradRec32* __restrict__ tile, int t0,
                                        int t1, const double* px, double py, double* sx, double* sy)
{
    for (int c0 = t0; c0 < t1; c0 += CHUNK) {
        const int c1 = min(c0 + CHUNK, t1);
        float fx[PX], fy[PX];
#pragma unroll
        for (int k = 0; k < PX; ++k) fx[k] = fy[k] = 0.f;
        for (int i = c0; i < c1; ++i) {
            const GradRec32 h = hot_fields<5>(tile, i - t0);  // cs .. circ
            const double xh = h64[i].xh;
            const float ty = (float)(py - h64[i].yh);
            const float ty2 = ty * ty;
#pragma unroll
            for (int k = 0; k < PX; ++k) {
                const float tx = offset_x(xh, px[k]);
                if (__builtin_fmaf(tx, tx, ty2) < h.near2) {
                    const HaloRec& hd = h64[i];
                    if constexpr (TYPE == TYPE_DPIE) grad_piemd<true>(hd, px[k], py, sx[k], sy[k]);
                    else if constexpr (TYPE == TYPE_PIEMD) grad_piemd<false>(hd, px[k], py, sx[k], sy[k]);
                    else if constexpr (TYPE == TYPE_SIS_CIRC) grad_sis_circ(hd, px[k], py, sx[k], sy[k]);
                    else grad_sis(hd, px[k], py, sx[k], sy[k]);
                } else {
                    if constexpr (TYPE == TYPE_DPIE) grad_piemd32<true>(h, tx, ty, fx[k], fy[k]);
                    else if constexpr (TYPE == TYPE_PIEMD) grad_piemd32<false>(h, tx, ty, fx[k], fy[k]);
                    else if constexpr (TYPE == TYPE_SIS_CIRC) grad_sis32_circ(h, tx, ty, fx[k], fy[k]);
                    else grad_sis32(h, tx, ty, fx[k], fy[k]);
                }
            }
        }
#pragma unroll
        for (int k = 0; k < PX; ++k) {
            sx[k] += (double)fx[k];
            sy[k] += (double)fy[k];
        }
    }
}

// One run of halos of one type, its FP32 records staged tile by tile (staged: 12.9 ms on the
// 503-halo 4096^2 map, scalar loads: 13.9 ms). Called by every thread of the work-group with the
// same run (the barriers need all of them).
template <int TYPE, int PX>
__device__ __forceinline__ void rung32(const HaloRec* __restrict__ h64, const GradRec32* __restrict__ g32,
                                       GradRec32* __restrict__ tile, int start, int end, const double* px, double py,
                                       double* sx, double* sy)
{
    for (int t0 = start; t0 < end; t0 += TILE) {
        const int t1 = min(t0 + TILE, end);
        __syncthreads();  // every wave is done with the previous tile
        stage_tile(tile, g32, t0, t1 - t0);
        __syncthreads();
        tileg32<TYPE, PX>(h64, tile, t0, t1, px, py, sx, sy);
    }
}

template <int PX>
__global__ __launch_bounds__(WAVE* ROWS) void k_gradient32(const HaloRec* __restrict__ h64,
                                                           const GradRec32* __restrict__ g32,
                                                           const Run* __restrict__ runs, int nruns, GridDesc g,
                                                           OutDesc o)
{
    // rows past the grid walk the runs too (the tile barriers need every thread) and store nothing
    __shared__ GradRec32 tile[TILE];
    const int row = blockIdx.y * ROWS + threadIdx.y;
    const bool live = row < g.ny;
    h64 = select_sample(g, h64, o);
    g32 += (int64_t)blockIdx.z * g.halo_stride;
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
        if (run.type == TYPE_DPIE) rung32<TYPE_DPIE, PX>(h64, g32, tile, run.start, end, px, py, sx, sy);
        else if (run.type == TYPE_PIEMD) rung32<TYPE_PIEMD, PX>(h64, g32, tile, run.start, end, px, py, sx, sy);
        else if (run.circ) rung32<TYPE_SIS_CIRC, PX>(h64, g32, tile, run.start, end, px, py, sx, sy);
        else rung32<TYPE_SIS, PX>(h64, g32, tile, run.start, end, px, py, sx, sy);
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
        if (live && col < g.nx) {
            const int64_t idx = (int64_t)row * g.pitch + col;
            o.p0[idx] = tx[k] * m;
            o.p1[idx] = ty[k] * m;
        }
    }
}

hipError_t launch_gradient32(const HaloRec* h64, const HaloRec32* recs32, const Run* runs, int nruns,
                             const GridDesc& g, const OutDesc& o, int px, hipStream_t s)
{
    const GradRec32* g32 = reinterpret_cast<const GradRec32*>(recs32);
    dim3 blk(WAVE, ROWS, 1);
    switch (px) {
    case 1: hipLaunchKernelGGL(k_gradient32<1>, grid_for(g, 1), blk, 0, s, h64, g32, runs, nruns, g, o); break;
    case 4: hipLaunchKernelGGL(k_gradient32<4>, grid_for(g, 4), blk, 0, s, h64, g32, runs, nruns, g, o); break;
    default: hipLaunchKernelGGL(k_gradient32<2>, grid_for(g, 2), blk, 0, s, h64, g32, runs, nruns, g, o); break;
    }
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------------------
// FP32-compensated potential: the form of lt_terms64.h pot_piemd (phi_r = x arg Q + y alpha_y -
// 2 sqrt(e) r ln((1 + e) |Q| / (2 r)) with log1p logarithms and, for the dPIE, the deflection parts
// of the two radii merged without cancellation; the reference's is src/potential_GPU.cu:47-116)
// evaluated in FP32 with the single-precision arctangent, a series log1p and the hardware base-2
// logarithm; the offset pixel - centre formed in FP64 and rounded once; FP32 terms summed per chunk
// of CHUNK halos into FP64 run sums. Every term is FP32 (the potential has no singular core: it is
// finite and smooth at the centre of a cored halo), and the map reports its error against the FP64
// kernel. Floating-point contraction is off in these functions: only the written FMAs are fused, so
// that every launch shape (pixels per thread) evaluates the same operations and gives the same map.
// ------------------------------------------------------------------------------------------------

// ln(1 + a) in FP32 for a > -1: 2 atanh(s), s = a / (2 + a), by its series through s^9 for
// -0.3 < a < 0.4 (|s| < 0.18: the next term is below 4e-9 relative), the hardware logarithm of 1 + a
// elsewhere (where 1 + a is not close to 1)
__device__ __forceinline__ float log1pf_lt(float a)
{
#pragma clang fp contract(off)
    const float sv = a * rcp32(2.f + a);
    const float w = sv * sv;
    float p = __builtin_fmaf(w, 1.f / 9.f, 1.f / 7.f);
    p = __builtin_fmaf(w, p, 1.f / 5.f);
    p = __builtin_fmaf(w, p, 1.f / 3.f);
    const float two_s = sv + sv;
    const float series = __builtin_fmaf(two_s * w, p, two_s);
    return (a > -0.3f && a < 0.4f) ? series : logf_lt(1.f + a);
}

// min(a, b) of two non-NaN floats as one instruction
__device__ __forceinline__ float min32(float a, float b)
{
    float r;
    asm("v_min_f32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}

// one core radius: w, d = w - r, Re Q = c r + cx1 d, lq = ln(|Q|^2 / (c r)^2) (lt_terms64.h pot_rad)
struct PotRad32 {
    float w, d, R, lq;
};

__device__ __forceinline__ PotRad32 pot_rad32(float cx1, float rem2, float I2, float r, float r2, float cr,
                                              float icr2)
{
#pragma clang fp contract(off)
    const float q = r2 + rem2;
    PotRad32 o;
    o.w = q * rsq32(q);
    o.d = rem2 * rcp32(o.w + r);
    const float cd = cx1 * o.d;
    o.R = cd + cr;
    o.lq = log1pf_lt(__builtin_fmaf(cd, __builtin_fmaf(2.f, cr, cd), I2) * icr2);
    return o;
}

// one halo's potential at the FP32 offset (tx, ty); mn = min(mn, rem^2) over PIEMD halos (a pixel on a
// centre, where the reference's potential is 0/0, gets NaN)
template <int TYPE>
__device__ __forceinline__ float pot32(const PotRec32& h, float tx, float ty, float& mn)
{
#pragma clang fp contract(off)
    const float x = __builtin_fmaf(tx, h.cs, ty * h.sn);
    const float y = __builtin_fmaf(ty, h.cs, -(tx * h.sn));
    const float x2 = x * x, y2 = y * y;
    if constexpr (TYPE == TYPE_SIS) {
        const float R2 = __builtin_fmaf(h.sa, x2, h.sb * y2);
        return h.amp * (R2 * rsq32(R2));
    } else {
        const float rem2 = __builtin_fmaf(x2, h.kx, y2 * h.ky);
        mn = min32(mn, rem2);
        const float I = h.gki * x;
        const float I2 = I * I;
        const float v = h.gkv * y;
        const PotRad32 a = pot_rad32(h.cx1, rem2, I2, h.ra, h.ra2, h.cra, h.icra2);
        if constexpr (TYPE == TYPE_PIEMD) {
            const float ls = log1pf_lt((a.d + v) * h.icra);
            const float arg = atan2pf_lt(I, a.R);
            const float ay = __builtin_fmaf(-0.5f, a.lq, ls);
            return h.amp * __builtin_fmaf(x, arg, __builtin_fmaf(y, ay, -(h.sra * a.lq)));
        } else {
            const PotRad32 b = pot_rad32(h.cx1, rem2, I2, h.rb, h.rb2, h.crb, h.icrb2);
            const float sig = (a.d + b.d) * rcp32(a.w + b.w);
            const float dR = h.pdr * __builtin_fmaf(-h.cx1, sig, h.pc);
            const float arg = atan2pf_lt(I * dR, __builtin_fmaf(a.R, b.R, I2));
            const float Q2a = __builtin_fmaf(a.R, a.R, I2);
            const float Sa = (h.cra + a.d) + v, Sb = (h.crb + b.d) + v;
            const float ga = __builtin_fmaf(h.pomc, a.d, v);
            const float Pa = __builtin_fmaf(ga, Sa + a.R, -I2);
            const float dP = __builtin_fmaf(Sb + b.R, h.pomcdr * sig, -(ga * (h.pcdr * (2.f - sig))));
            const float N = __builtin_fmaf(Pa, dR * (a.R + b.R), Q2a * dP);
            const float ay = 0.5f * log1pf_lt(N * rcp32((Sb * Sb) * Q2a));
            const float L = __builtin_fmaf(h.sra, a.lq, -(h.srb * b.lq));
            return h.amp * __builtin_fmaf(x, arg, __builtin_fmaf(y, ay, -L));
        }
    }
}

template <int TYPE, int PX>
__device__ __forceinline__ void runp32(const HaloRec* __restrict__ h64, const PotRec32* __restrict__ p32, int start,
                                       int end, const double* px, double py, double* s, float* mn)
{
#pragma clang fp contract(off)
    for (int c0 = start; c0 < end; c0 += CHUNK) {
        const int c1 = min(c0 + CHUNK, end);
        float f[PX];
#pragma unroll
        for (int k = 0; k < PX; ++k) f[k] = 0.f;
        for (int i = c0; i < c1; ++i) {
            const PotRec32& h = p32[i];
            const double xh = h64[i].xh;
            const float ty = (float)(py - h64[i].yh);
#pragma unroll
            for (int k = 0; k < PX; ++k) {
                const float v = pot32<TYPE>(h, offset_x(xh, px[k]), ty, mn[k]);
                f[k] = f[k] + v;
            }
        }
#pragma unroll
        for (int k = 0; k < PX; ++k) s[k] += (double)f[k];
    }
}

template <int PX>
__global__ __launch_bounds__(WAVE* ROWS) void k_potential32(const HaloRec* __restrict__ h64,
                                                            const PotRec32* __restrict__ p32,
                                                            const Run* __restrict__ runs, int nruns, GridDesc g,
                                                            OutDesc o)
{
    const int row = blockIdx.y * ROWS + threadIdx.y;
    if (row >= g.ny) return;
    h64 = select_sample(g, h64, o);
    p32 += (int64_t)blockIdx.z * g.halo_stride;
    const int col0 = blockIdx.x * (WAVE * PX) + threadIdx.x;
    const double py = pixel_coord(g.ymin, row + g.jstart, g.dy);
    double px[PX], tot[PX];
    float mn[PX];
#pragma unroll
    for (int k = 0; k < PX; ++k) {
        px[k] = pixel_coord(g.xmin, col0 + k * WAVE + g.istart, g.dx);
        tot[k] = 0.;
        mn[k] = 1.f;
    }
    for (int r = 0; r < nruns; ++r) {
        const Run run = runs[r];
        double s[PX];
#pragma unroll
        for (int k = 0; k < PX; ++k) s[k] = 0.;
        const int end = run.start + run.count;
        if (run.type == TYPE_DPIE) runp32<TYPE_DPIE, PX>(h64, p32, run.start, end, px, py, s, mn);
        else if (run.type == TYPE_PIEMD) runp32<TYPE_PIEMD, PX>(h64, p32, run.start, end, px, py, s, mn);
        else runp32<TYPE_SIS, PX>(h64, p32, run.start, end, px, py, s, mn);
#pragma unroll
        for (int k = 0; k < PX; ++k) tot[k] += s[k];
    }
#pragma unroll
    for (int k = 0; k < PX; ++k) {
        const int col = col0 + k * WAVE;
        // a pixel on a PIEMD centre: the reference's potential is 0/0 there
        const double v = mn[k] == 0.f ? __builtin_nan("") : tot[k] * o.mult;
        if (col < g.nx) o.p0[(int64_t)row * g.pitch + col] = v;
    }
}

hipError_t launch_potential32(const HaloRec* h64, const HaloRec32* recs32, const Run* runs, int nruns,
                              const GridDesc& g, const OutDesc& o, int px, hipStream_t s)
{
    const PotRec32* p32 = reinterpret_cast<const PotRec32*>(recs32);
    dim3 blk(WAVE, ROWS, 1);
    switch (px) {
    case 1: hipLaunchKernelGGL(k_potential32<1>, grid_for(g, 1), blk, 0, s, h64, p32, runs, nruns, g, o); break;
    case 4: hipLaunchKernelGGL(k_potential32<4>, grid_for(g, 4), blk, 0, s, h64, p32, runs, nruns, g, o); break;
    default: hipLaunchKernelGGL(k_potential32<2>, grid_for(g, 2), blk, 0, s, h64, p32, runs, nruns, g, o); break;
    }
    return hipGetLastError();
}

}  // namespace lt
