// lt_math.h -- small device math helpers for the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace lt {

// 1/x in FP64 for the strictly positive, normal-range denominators of the closed forms: hardware
// reciprocal estimate r (relative error d below 2^-24 on gfx950), then one third-order correction
// 1/x = r / (1 - e) ~ r (1 + e + e^2) with e = 1 - x r = d, whose truncation error e^3 < 2^-72 is far
// below the rounding of the result (two FMAs and one multiply-free FMA).
__device__ __forceinline__ double rcp64(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    double e = __builtin_fma(-x, r, 1.0);
    return __builtin_fma(r, __builtin_fma(e, e, e), r);
}

// 1/sqrt(q) in FP64 for q > 0: hardware estimate y (relative error below 2^-24), then one
// third-order correction q^(-1/2) = y (1 - e)^(-1/2) ~ y (1 + e/2 + 3 e^2/8) with e = 1 - q y^2;
// truncation error about 5 e^3 / 16 < 2^-70.
__device__ __forceinline__ double rsq64(double q)
{
    double y = __builtin_amdgcn_rsq(q);
    double e = __builtin_fma(-(q * y), y, 1.0);
    double p = __builtin_fma(0.375, e, 0.5);
    return __builtin_fma(y, e * p, y);
}

// The same estimates with one second-order correction, for the Hessian terms: 1/x = r (1 + e) with
// relative error e^2 < 2^-48, 1/sqrt(q) = y (1 + e/2) with relative error 1.5 d^2 < 2^-47.8 (d the
// estimate's relative error, e = 1 - q y^2 = -2d - d^2): a few units in the last place of a double,
// three orders of magnitude inside the 1e-12 the Hessian maps are held to, one instruction less each.
__device__ __forceinline__ double rcp64_n2(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    double e = __builtin_fma(-x, r, 1.0);
    return __builtin_fma(r, e, r);
}

__device__ __forceinline__ double rsq64_n2(double q)
{
    double y = __builtin_amdgcn_rsq(q);
    double e = __builtin_fma(-(q * y), y, 1.0);
    return __builtin_fma(y * 0.5, e, y);
}

// Pixel coordinate exactly as the reference CPU forms it: xmin + (i * d), two roundings, no fused
// multiply-add (src/grid_gradient_CPU.cpp:89-90).
__device__ __forceinline__ double pixel_coord(double xmin, int i, double d)
{
#pragma clang fp contract(off)
    double t = (double)i * d;
    return xmin + t;
}

// Error-free transformations for compensated FP32 sums (Knuth two-sum, Dekker fast-two-sum).
__device__ __forceinline__ void two_sum(float a, float b, float& s, float& e)
{
#pragma clang fp contract(off)
    s = a + b;
    float bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

// Compensated accumulator: value = hi + lo, with the rounding error of every addition kept.
struct CompF32 {
    float hi, lo;
    __device__ __forceinline__ void init() { hi = 0.f; lo = 0.f; }
    __device__ __forceinline__ void add(float v)
    {
        float s, e;
        two_sum(hi, v, s, e);
        hi = s;
        lo += e;
    }
    __device__ __forceinline__ double value() const { return (double)hi + (double)lo; }
};

}  // namespace lt
