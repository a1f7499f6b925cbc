// lt_delense.h -- device layout of the batched delensing + chi^2 pipeline (lt_delense.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lenstool_hip/lt_types.h"

namespace lt {

// hits kept per (sample, source) before they are put back into scan order; the first
// LT_MAXIMPERSOURCE of them are reported, more than that is flagged LT_ERR_OVERFLOW
constexpr int DELENSE_CAP = 64;

struct DelenseDesc {
    int32_t n;            // delensing grid is n x n pixels, (n-1)^2 cells
    int32_t nsets;        // image families (sources)
    int32_t nsamples;     // lens-model samples
    int32_t nimg_total;   // observed images over all families
    double xmin, ymin, dx, dy;
    int64_t grid_stride;  // elements between the gradient planes of two samples (n*n)
    const int32_t* nimages;      // [nsets]
    const int32_t* first_image;  // [nsets]
    const double* img_x;         // [nimg_total] observed image positions and D_ls/D_s
    const double* img_y;
    const double* img_dr;
    double* src_x;               // [nsamples*nsets] source barycentres
    double* src_y;
    double* src_dr;
    double* src_bound;           // [nsamples*3] max |x|, |y|, min dr, max dr of the sample's sources
    int32_t* count;              // [nsamples*nsets] hits found
    int32_t* hits;               // [nsamples*nsets*DELENSE_CAP] cell*2 + (lower triangle)
    double* pos_x;               // [nsamples*nsets*LT_MAXIMPERSOURCE] image barycentres
    double* pos_y;
    int32_t* loc;                // [nsamples*nsets] images kept per source
    int32_t* status;             // [nsamples]
    double* terms;               // [nsamples*nimg_total] chi^2 term per observed image
    double* chi;                 // [nsamples]
    int32_t* found;              // [nsamples]
};

hipError_t launch_barycentres(const double* img_gx, const double* img_gy, const DelenseDesc& d, hipStream_t s);
hipError_t launch_delense(const double* gx, const double* gy, const DelenseDesc& d, hipStream_t s);
hipError_t launch_chi(const DelenseDesc& d, hipStream_t s);

}  // namespace lt
