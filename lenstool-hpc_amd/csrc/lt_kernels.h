// lt_kernels.h -- host-side declarations of the kernel launchers and of the model packer.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "lenstool_hip/lt_types.h"
#include "lt_device.h"

namespace lt {

// FP64 kernels (lt_kernels64.hip). px = pixels per thread (1, 2 or 4).
hipError_t launch_hessian64(const HaloRec* halos, const Run* runs, int nruns, const GridDesc& g,
                            const OutDesc& o, int px, hipStream_t s);
hipError_t launch_gradient64(const HaloRec* halos, const Run* runs, int nruns, const GridDesc& g,
                             const OutDesc& o, int px, hipStream_t s);
hipError_t launch_potential64(const HaloRec* halos, const Run* runs, int nruns, const GridDesc& g,
                              const OutDesc& o, int px, hipStream_t s);
// gradient at npts points for nsamples lens models (records of sample s start at s*halo_stride);
// outputs [nsamples*npts]
hipError_t launch_points_gradient64(const HaloRec* halos, const Run* runs, int nruns, const double* px,
                                   const double* py, int npts, int nsamples, int64_t halo_stride, double* gx,
                                   double* gy, hipStream_t s);
// map epilogue alone on device Hessian planes (kinds FIELDS, AMPLI, SHEAR, MASS)
hipError_t launch_epilogue64(const double* a, const double* b, const double* c, int64_t n, const OutDesc& o,
                             hipStream_t s);

// FP32-compensated kernels (lt_kernels32.hip).
hipError_t launch_hessian32(const HaloRec* halos64, const HaloRec32* halos32, const Run* runs, int nruns,
                            const GridDesc& g, const OutDesc& o, int px, hipStream_t s);
// gradient / displacement maps; the model must be packed for gradient use (GradRec32 records)
hipError_t launch_gradient32(const HaloRec* halos64, const HaloRec32* recs32, const Run* runs, int nruns,
                             const GridDesc& g, const OutDesc& o, int px, hipStream_t s);

// potential maps; the model must be packed for potential use (PotRec32 records)
hipError_t launch_potential32(const HaloRec* halos64, const HaloRec32* recs32, const Run* runs, int nruns,
                              const GridDesc& g, const OutDesc& o, int px, hipStream_t s);

// Streaming device copy (16-byte vectors plus a byte tail).
hipError_t launch_copy(void* dst, const void* src, size_t bytes, hipStream_t s);
// x *= s on n doubles
hipError_t launch_scale(double* x, double s, size_t n, hipStream_t st);

// What a packed model is used for: the run split differs for the Kmap (runs also break on dlsds).
enum class Use { Gradient, Hessian, Kmap, Potential };

// Pack an SOA lens model into device records and runs. Returns an LT_* status; on error `err`
// names the offending halo.
int pack_model(const lt_potential_soa* lens, int nhalos, Use use, std::vector<HaloRec>& recs,
               std::vector<HaloRec32>& recs32, std::vector<Run>& runs, std::string& err);

// Cosmology (lt_cosmo.cpp): angular distances in units of c/H0 and D_ls/D_s.
double cosmo_observer_object(double z, const lt_cosmo_param& c);
double cosmo_object_object(double z1, double z2, const lt_cosmo_param& c);

}  // namespace lt
