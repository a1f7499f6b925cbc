// lt_device.h -- device-side data layout shared by the kernels and the host packer.
//
// A lens model lives in HBM as an array of 384-byte halo records (one per halo, in the SOA order of
// the caller, i.e. grouped by type) plus a short array of runs (type, first halo, count, weight).
// Every per-halo constant of the closed forms is precomputed on the host once per model, so the
// per-(halo, pixel) work is only what depends on the pixel. All lanes of a wavefront walk the same
// halo at the same time: the record address is wavefront-uniform and is read with scalar loads.
#pragma once

#include <stdint.h>

namespace lt {

// One halo, 48 doubles. Field meaning by type:
//   all     : xh, yh (centre), cs, sn (cos/sin of the angle), hc2, hs2 (Hessian amplitude times
//             cos/sin of twice the angle)
//   8, 81   : PIEMD constants (see below); ampg / amph gradient / Hessian amplitudes; csx .. k5,
//             ra2c, rb2c, qT .. qs2h the constants of the Hessian's scaled division-free form;
//             psra .. picra those of the potential (lt_terms64.h)
//   5       : sa = 1 - eps, sb = 1 + eps, ampg = b0, amph = b0 (1 - eps)(1 + eps); circ = 1 when
//             eps = 0 (runs of circular SIS form their terms from the lab-frame offsets)
//   14      : h14a, h14b, h14c = constant Hessian (lab frame)
struct alignas(64) HaloRec {
    double xh, yh, cs, sn;          // 0..3
    double hc2, hs2;                // 4, 5
    double kx, ky;                  // 6, 7   1/(1+e)^2, 1/(1-e)^2
    double e2;                      // 8      2 sqrt(e)
    double cx1, gki, gkv;           // 9..11  (1-e)/(1+e); gradient: 2 sqrt(e) a^2, 2 sqrt(e) a b (a, b = 1/(1 +- e))
    double ra2, rb2;                // 12, 13 rcore^2, rcut^2
    double rca, rcb;                // 14, 15 cx1 rcore, cx1 rcut
    double ampg;                    // 16     gradient amplitude
    double amph;                    // 17     Hessian amplitude
    double csx, snx, csy, sny;      // 18..21 cx1 a (cs, sn), cx1 b (cs, sn): the scaled Hessian frame
    double ra2c, rb2c;              // 22, 23 cx1^2 rcore^2, cx1^2 rcut^2
    union {                         // 24..28 by type:
        struct {
            double sa, sb;          //   SIS (1 - e), (1 + e)
            double h14a, h14b, h14c;  // type-14 constant Hessian
        };
        struct {                    //   PIEMD potential, dPIE merging of the two radii (lt_terms64.h):
            double pc, pomc;        //   c = 1 + cx1, 1 - cx1
            double pdr, pcdr, pomcdr;  // D = rcut - rcore, c D, (1 - cx1) D
        };
    };
    double psra, psrb;              // 29, 30 potential: sqrt(e) rcore, sqrt(e) rcut
    double rc, rcut;                // 31, 32 rcore, rcut
    double k1, k2, k3;              // 33..35 2 sqrt(e) a / cx1, b / a, 2 sqrt(e) b / cx1
    double qT, qc2, qs2m2, qs2h;    // 36..39 Hessian amplitudes of the division-free form
    double k4, k5;                  // 40, 41 1 / cx1^2, 2 sqrt(e) a   (a, b = 1/(1 +- e))
    double pcra, pcrb;              // 42, 43 potential: c rcore, c rcut (c = 1 + cx1 = 2 / (1 + e))
    double picra2, picrb2;          // 44, 45 1 / (c rcore)^2, 1 / (c rcut)^2
    double circ;                    // 46     1 for a circular SIS (e = 0)
    double picra;                   // 47     1 / (c rcore)
};
static_assert(sizeof(HaloRec) == 384, "halo record is 384 bytes");

// A run of halos of one type (and, for the Kmap, one dlsds weight).
struct alignas(32) Run {
    int32_t type;
    int32_t start;
    int32_t count;
    int32_t circ;       // 1: every halo of the run is a circular SIS (in every sample of a batch)
    double weight;      // 1 except for the Kmap (dlsds of the run)
    double pad2;
};
static_assert(sizeof(Run) == 32, "run is 32 bytes");

// Single-precision record of a halo for the FP32-compensated Hessian. The PIEMD terms use a form
// of the I_{0.5} Hessian without the removable 0/0 of the closed form (see lt_kernels32.hip), whose
// constants are these. The offset pixel - centre is formed from the FP64 centre of the HaloRec and
// rounded once to FP32. near2: squared halo-frame radius inside which the FP32 rounding of
// this halo's term would exceed the error budget; there the FP64 term of the HaloRec is used.
struct alignas(64) HaloRec32 {
    float csx, snx, csy, sny;           // 0..3    scaled frame (PIEMD: cx1 a (cs, sn), cx1 b (cs, sn); SIS: cs, sn)
    float k1, k2, k3, k4;               // 4..7    as HaloRec
    float k5, ra, ra2c, rb;             // 8..11   2 sqrt(e) a, rcore, cx1^2 rcore^2, rcut
    float rb2c;                         // 12      cx1^2 rcut^2
    float hT, hc2, hs2m2, hs2h;         // 13..16  lab-frame amplitudes: trace, cos 2t, -2 sin 2t, sin 2t / 2
    float near2;                        // 17
    float sa, sb;                       // 18, 19  SIS: 1 - e, 1 + e
    float circ;                         // 20      1 for a circular SIS
    float near2s;                       // 21      near bound on X^2 + Y^2 (scaled frame; lab frame for circular SIS)
    float nearr;                        // 22      radius of a disk about the centre holding that near region
    float pad[9];                       // 23..31
};
static_assert(sizeof(HaloRec32) == 128, "fp32 halo record is 128 bytes");

// Single-precision record of a halo for the FP32-compensated gradient (models packed for gradient
// use hold these in place of HaloRec32, same size). PIEMD terms use the ratio N/D of the I_{0.5}
// gradient in the form Q/S (lt_kernels32.hip): Q = r + cx1 w + i ki x, S = w + r cx1 + kv y, with
// w = sqrt(r^2 + kx x^2 + ky y^2); the near radius follows the HaloRec32 rule.
struct alignas(64) GradRec32 {
    float cs, sn;                       // 0, 1    cos, sin of the angle
    float kx, ky, cx1;                  // 2..4    1/(1+e)^2, 1/(1-e)^2, (1-e)/(1+e)
    float ki, kv;                       // 5, 6    2 sqrt(e)/(1+e)^2, 2 sqrt(e)/((1+e)(1-e))
    float ra, rb, ra2, rb2;             // 7..10   rcore, rcut, their squares
    float rca, rcb;                     // 11, 12  cx1 rcore, cx1 rcut
    float ampg;                         // 13      gradient amplitude
    float near2;                        // 14
    float sa, sb, circ;                 // 15..17  SIS: 1 - e, 1 + e, 1 when circular
    float pad[14];                      // 18..31
};
static_assert(sizeof(GradRec32) == sizeof(HaloRec32), "gradient fp32 record is 128 bytes");

// Single-precision record of a halo for the FP32-compensated potential (models packed for potential
// use hold these in place of HaloRec32, same size): the constants of lt_terms64.h pot_piemd in FP32.
struct alignas(64) PotRec32 {
    float cs, sn;                       // 0, 1    cos, sin of the angle
    float kx, ky, cx1, gki, gkv;        // 2..6    as HaloRec
    float ra, ra2, cra, icra2, icra, sra;  // 7..12  rcore, rcore^2, c rcore, 1/(c rcore)^2, 1/(c rcore), sqrt(e) rcore
    float rb, rb2, crb, icrb2, srb;     // 13..17  the same for rcut (c = 1 + cx1)
    float pc, pomc, pdr, pcdr, pomcdr;  // 18..22  c, 1 - cx1, D = rcut - rcore, c D, (1 - cx1) D
    float amp;                          // 23      PIEMD / dPIE amplitude (ampg); SIS b0
    float sa, sb;                       // 24, 25  SIS: 1 - e, 1 + e
    float pad[6];                       // 26..31
};
static_assert(sizeof(PotRec32) == sizeof(HaloRec32), "potential fp32 record is 128 bytes");

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
