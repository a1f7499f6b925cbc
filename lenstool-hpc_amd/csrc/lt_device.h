// lt_device.h -- device-side data layout shared by the kernels and the host packer.
//
// A lens model lives in HBM as an array of 256-byte halo records (one per halo, in the SOA order of
// the caller, i.e. grouped by type) plus a short array of runs (type, first halo, count, weight).
// Every per-halo constant of the closed forms is precomputed on the host once per model, so the
// per-(halo, pixel) work is only what depends on the pixel. All lanes of a wavefront walk the same
// halo at the same time: the record address is wavefront-uniform and is read with scalar loads.
#pragma once

#include <stdint.h>

namespace lt {

// One halo, 40 doubles. Field meaning by type:
//   all     : xh, yh (centre), cs, sn (cos/sin of the angle), hc2, hs2 (Hessian amplitude times
//             cos/sin of twice the angle), b0, eps, rc, rcut (raw parameters, for the potential)
//   8, 81   : PIEMD constants (see below); ampg / amph gradient / Hessian amplitudes; e2ab .. qs2h
//             the constants of the division-free Hessian form
//   5       : sa = 1 - eps, sb = 1 + eps, ampg = b0, amph = b0 (1 - eps)(1 + eps)
//   14      : h14a, h14b, h14c = constant Hessian (lab frame)
struct alignas(64) HaloRec {
    double xh, yh, cs, sn;          // 0..3
    double hc2, hs2;                // 4, 5
    double kx, ky;                  // 6, 7   1/(1+e)^2, 1/(1-e)^2
    double e2;                      // 8      2 sqrt(e)
    double cx1, icx1, cx1sq;        // 9..11  (1-e)/(1+e), its inverse, its square
    double ra2, rb2;                // 12, 13 rcore^2, rcut^2
    double e2ra, e2rb;              // 14, 15 2 rcore sqrt(e), 2 rcut sqrt(e)
    double ampg;                    // 16     gradient amplitude
    double amph;                    // 17     Hessian amplitude
    double hk1, hk2, hk3, hk4, hk5, hk6;  // 18..23 Hessian bracket coefficients
    double sa, sb;                  // 24, 25 SIS (1 - e), (1 + e)
    double h14a, h14b, h14c;        // 26..28 type-14 constant Hessian
    double b0, eps, rc, rcut;       // 29..32 raw parameters (potential)
    double e2ab, e2a2, cxa2;        // 33..35 2 sqrt(e) a b, 2 sqrt(e) a^2, cx1 a^2 (a, b = 1/(1 +- e))
    double qT, qc2, qs2m2, qs2h;    // 36..39 Hessian amplitudes of the division-free form (lt_terms64.h)
};
static_assert(sizeof(HaloRec) == 320, "halo record is 320 bytes");

// A run of halos of one type (and, for the Kmap, one dlsds weight).
struct alignas(32) Run {
    int32_t type;
    int32_t start;
    int32_t count;
    int32_t pad;
    double weight;      // 1 except for the Kmap (dlsds of the run)
    double pad2;
};
static_assert(sizeof(Run) == 32, "run is 32 bytes");

// Single-precision record of a halo for the FP32-compensated Hessian. The PIEMD terms use a form
// of the I_{0.5} Hessian without the removable 0/0 of the closed form (see lt_kernels32.hip), whose
// constants are these. The centre is split into hi + lo floats so that pixel - centre keeps FP64
// accuracy in FP32 arithmetic. near2: squared halo-frame radius inside which the FP32 rounding of
// this halo's term would exceed the error budget; there the FP64 term of the HaloRec is used.
struct alignas(64) HaloRec32 {
    float xh_hi, xh_lo, yh_hi, yh_lo;   // 0..3
    float cs, sn;                       // 4, 5
    float a2, b2;                       // 6, 7    a^2 = 1/(1+e)^2, b^2 = 1/(1-e)^2
    float e2, cx1, e2ab, e2a2;          // 8..11   2 sqrt(e), (1-e)/(1+e), 2 sqrt(e) a b, 2 sqrt(e) a^2
    float cxa2;                         // 12      cx1 a^2
    float ra, ra2;                      // 13, 14  rcore, rcore^2
    float rb, rb2;                      // 15, 16  rcut, rcut^2
    float icx1, pad0;                   // 17, 18  1 / cx1
    float hT, hc2, hs2m2, hs2h;         // 19..22  lab-frame amplitudes: trace, cos 2t, -2 sin 2t, sin 2t / 2
    float near2;                        // 23
    float sa, sb;                       // 24, 25  SIS: 1 - e, 1 + e
    float pad[6];                       // 26..31
};
static_assert(sizeof(HaloRec32) == 128, "fp32 halo record is 128 bytes");

// Grid description (pixel (col, row) of the output sits at (xmin + (col+istart) dx,
// ymin + (row+jstart) dy); output index row*pitch + col).
struct GridDesc {
    double xmin, ymin, dx, dy;
    int32_t nx, ny, istart, jstart;
    int64_t pitch;
    // batches of lens models on one grid (blockIdx.z = sample): sample s reads halo records
    // [s*halo_stride, s*halo_stride + nhalos) and writes its planes at offset s*plane_stride
    int64_t halo_stride;
    int64_t plane_stride;
    int32_t nsamples;
    int32_t pad;
};

// What the epilogue writes. kind / mode follow lt_status.h (LT_MAP_*).
struct OutDesc {
    double* p0;
    double* p1;
    double* p2;
    double* p3;
    double* p4;
    int32_t kind;
    int32_t mode;
    double r;        // dlsds = D_ls / D_s applied by the map formulas
    double mult;     // mass-map multiplier (mode 2..4), potential multiplier
};

enum : int32_t { TYPE_SIS = 5, TYPE_PIEMD = 8, TYPE_DPIE = 81, TYPE_SHEAR = 14 };

}  // namespace lt
