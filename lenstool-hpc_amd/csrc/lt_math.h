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

// Polynomial tables of atan_lt / log_lt (lenstool-hpc_amd/csrc/gen_poly.py derives them with
// mpmath; the interpolation error is below 1e-17 relative in both functions).
// kAtanP: degree 10 on [0, 0.171573], interpolation error 3.16e-17
constexpr double kAtanP[11] = {
    -0x1.5555555555555p-2,  // -0.3333333333333333016
    0x1.999999999934cp-3,  // 0.19999999999995520678
    -0x1.2492492436201p-3,  // -0.14285714284666542922
    0x1.c71c71853d7fap-4,  // 0.11111111015256361727
    -0x1.745d0b28a7e37p-4,  // -0.090909045781239018905
    0x1.3b1263064f6b9p-4,  // 0.076921831908260865637
    -0x1.10fa77b1a6d57p-4,  // -0.066645114473819480001
    0x1.dfe6497e96323p-5,  // 0.058581489128022098816
    -0x1.a0999c632b6edp-5,  // -0.050854497379402598613
    0x1.4162c02b1dda3p-5,  // 0.039231658295587191303
    -0x1.3a31b1c0fd3b7p-6,  // -0.019176887119062258989
};
// kLogR: degree 6 on [0, 0.0294373], interpolation error 1.55e-16
constexpr double kLogR[7] = {
    0x1.5555555555558p-2,  // 0.33333333333333348431
    0x1.99999999952e2p-3,  // 0.19999999999949752245
    0x1.2492492df148dp-3,  // 0.14285714312987742461
    0x1.c71c62e5800a1p-4,  // 0.11111105567397539928
    0x1.7462b4ab2ef6bp-4,  // 0.090914445626308610401
    0x1.39fe606542ddep-4,  // 0.076658608002780206343
    0x1.2b584aae78a57p-4,  // 0.073082248425217029302
};

// atan2(y, x) in FP64 for finite arguments, within 3 units in the last place (2.7 measured over 4M
// random inputs against long double, tests/cpp/math_accuracy.hip): the angle of
// (max(|x|,|y|), min(|x|,|y|)) in [0, pi/4] is folded to [-pi/8, pi/8] by the rotation (u, v) ->
// (u + v, v - u) (a rotation by -pi/4 times sqrt 2, which leaves the ratio v/u right), then
// atan(t) = t + t^3 P(t^2) with |t| <= tan(pi/8) and one corrected reciprocal; the octant is put
// back with pi/4, pi/2 - r, pi - r and the sign of y. atan2(0, 0) = 0 (pi for x = -0), NaN in ->
// NaN out. About 40 instructions against the hundred of the library atan2.
// one Horner step p z + c as a three-operand FMA: written out so that the coefficient c, which the
// compiler keeps in a register across the halo loop, is read in place instead of being copied into
// the accumulator of the two-operand form (one instruction per step instead of two)
__device__ __forceinline__ double horner_step(double p, double z, double c)
{
    double r;
    asm("v_fma_f64 %0, %1, %2, %3" : "=v"(r) : "v"(p), "v"(z), "v"(c));
    return r;
}

// min(a, b) of two non-NaN doubles as one instruction (fmin adds a quieting v_max per operand)
__device__ __forceinline__ double min64(double a, double b)
{
    double r;
    asm("v_min_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}

// Largest value of v over the 64 lanes of a wavefront, returned to every lane; every lane of the
// wavefront must be executing. Within each row of 16 lanes by data-parallel moves (partner lane ^ 1,
// ^ 2, then the mirrored lane of the half row and of the row: each step pairs lanes holding the maxima
// of disjoint groups), then the four rows' maxima read from lanes 0, 16, 32 and 48.
template <int CTRL>
__device__ __forceinline__ double dpp_move64(double v)
{
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double read_lane64(double v, int r)
{
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), r);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), r);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double wave_max64(double v)
{
    v = __builtin_fmax(v, dpp_move64<0xB1>(v));   // quad_perm [1, 0, 3, 2]
    v = __builtin_fmax(v, dpp_move64<0x4E>(v));   // quad_perm [2, 3, 0, 1]
    v = __builtin_fmax(v, dpp_move64<0x141>(v));  // row_half_mirror
    v = __builtin_fmax(v, dpp_move64<0x140>(v));  // row_mirror
    return __builtin_fmax(__builtin_fmax(read_lane64(v, 0), read_lane64(v, 16)),
                          __builtin_fmax(read_lane64(v, 32), read_lane64(v, 48)));
}

__device__ __forceinline__ double atan_lt_poly(double t)
{
    const double z = t * t;
    double p = kAtanP[10];
#pragma unroll
    for (int k = 9; k >= 0; --k) p = horner_step(p, z, kAtanP[k]);
    return __builtin_fma(t * z, p, t);
}

__device__ __forceinline__ double atan2_lt(double y, double x)
{
    constexpr double kPi = 3.141592653589793116, kPi2 = 1.5707963267948965580, kPi4 = 0.78539816339744827900;
    constexpr double kTanPi8 = 0.41421356237309503;
    const double ax = __builtin_fabs(x), ay = __builtin_fabs(y);
    const double mx = ax > ay ? ax : ay, mn = ax > ay ? ay : ax;
    const bool rot = mn > kTanPi8 * mx;
    const double u = rot ? mx + mn : mx;
    const double v = rot ? mn - mx : mn;
    double r = atan_lt_poly(v * rcp64(u));
    r = rot ? r + kPi4 : r;
    r = ay > ax ? kPi2 - r : r;
    r = __builtin_signbit(x) ? kPi - r : r;
    r = mx == 0. ? (__builtin_signbit(x) ? kPi : 0.) : r;
    r = __builtin_copysign(r, y);
    return __builtin_isunordered(x, y) ? x + y : r;
}

// atan2(y, x) for x > 0, the case of the I_{0.5} terms (their x is Re Q > 0 or Ra Rb + I^2 > 0): the
// angle lies in (-pi/2, pi/2), so atan2_lt's quadrant, zero and NaN cases drop out; the same folding
// and polynomial otherwise.
__device__ __forceinline__ double atan2p_lt(double y, double x)
{
    constexpr double kPi2 = 1.5707963267948965580, kPi4 = 0.78539816339744827900;
    constexpr double kTanPi8 = 0.41421356237309503;
    const double ay = __builtin_fabs(y);
    const double mx = __builtin_fmax(x, ay), mn = __builtin_fmin(x, ay);
    const bool rot = mn > kTanPi8 * mx;
    const double u = rot ? mx + mn : mx;
    const double v = rot ? mn - mx : mn;
    double r = atan_lt_poly(v * rcp64(u));
    r = rot ? r + kPi4 : r;
    r = ay > x ? kPi2 - r : r;
    return __builtin_copysign(r, y);
}

// FP32 arctangent for the FP32-compensated gradient: the folding of atan2_lt, then
// atan(t) = t + t^3 P(t^2) with P of degree 4 (kAtanP32, gen_poly.py; interpolation error 1.1e-9
// in atan, below FP32 rounding); finite arguments.
constexpr float kAtanP32[5] = {
    -0x1.555554p-2f,  // -0.333333317612
    0x1.99973p-3f,    // 0.199995404836
    -0x1.242036p-3f,  // -0.14263955598
    0x1.b8103p-4f,    // 0.107437314908
    -0x1.08455ep-4f,  // -0.0645192820812
};

__device__ __forceinline__ float atan2f_lt(float y, float x)
{
    constexpr float kPi = 3.14159265f, kPi2 = 1.57079633f, kPi4 = 0.785398163f, kTanPi8 = 0.414213562f;
    const float ax = __builtin_fabsf(x), ay = __builtin_fabsf(y);
    const float mx = __builtin_fmaxf(ax, ay), mn = __builtin_fminf(ax, ay);
    const bool rot = mn > kTanPi8 * mx;
    const float u = rot ? mx + mn : mx;
    const float v = rot ? mn - mx : mn;
    const float t = v * __builtin_amdgcn_rcpf(u);
    const float z = t * t;
    float p = kAtanP32[4];
#pragma unroll
    for (int k = 3; k >= 0; --k) p = __builtin_fmaf(p, z, kAtanP32[k]);
    float r = __builtin_fmaf(t * z, p, t);
    r = rot ? r + kPi4 : r;
    r = ay > ax ? kPi2 - r : r;
    r = __builtin_signbitf(x) ? kPi - r : r;
    r = mx == 0.f ? (__builtin_signbitf(x) ? kPi : 0.f) : r;
    return __builtin_copysignf(r, y);
}

// atan2f_lt for x > 0 (as atan2p_lt)
__device__ __forceinline__ float atan2pf_lt(float y, float x)
{
    constexpr float kPi2 = 1.57079633f, kPi4 = 0.785398163f, kTanPi8 = 0.414213562f;
    const float ay = __builtin_fabsf(y);
    const float mx = __builtin_fmaxf(x, ay), mn = __builtin_fminf(x, ay);
    const bool rot = mn > kTanPi8 * mx;
    const float u = rot ? mx + mn : mx;
    const float v = rot ? mn - mx : mn;
    const float t = v * __builtin_amdgcn_rcpf(u);
    const float z = t * t;
    float p = kAtanP32[4];
#pragma unroll
    for (int k = 3; k >= 0; --k) p = __builtin_fmaf(p, z, kAtanP32[k]);
    float r = __builtin_fmaf(t * z, p, t);
    r = rot ? r + kPi4 : r;
    r = ay > x ? kPi2 - r : r;
    return __builtin_copysignf(r, y);
}

// natural logarithm in FP32 of a normal positive x: the hardware base-2 logarithm times ln 2
__device__ __forceinline__ float logf_lt(float x) { return __builtin_amdgcn_logf(x) * 0.693147181f; }

// natural logarithm in FP64 for finite x > 0, within 3 units in the last place (2.4 measured):
// x = m 2^e with m in [sqrt(1/2), sqrt 2), log m = 2 atanh(s) = 2 s + 2 s^3 R(s^2) with
// s = (m - 1)/(m + 1) (|s| <= 0.1716, one corrected reciprocal), plus e ln 2 with ln 2 split so that
// e ln2_hi is exact. About 30 instructions against the hundred of the library log.
__device__ __forceinline__ double log_lt(double x)
{
    constexpr double kLn2Hi = 0x1.62e42fee00000p-1, kLn2Lo = 0x1.a39ef35793c76p-33;
    constexpr double kSqrtHalf = 0.70710678118654757274;
    int e;
    double m = __builtin_frexp(x, &e);
    const bool lo = m < kSqrtHalf;
    m = lo ? m + m : m;
    e = lo ? e - 1 : e;
    const double f = m - 1.;
    const double sv = f * rcp64(m + 1.);
    const double w = sv * sv;
    double p = kLogR[6];
#pragma unroll
    for (int k = 5; k >= 0; --k) p = horner_step(p, w, kLogR[k]);
    const double s2 = sv + sv;
    const double lm = __builtin_fma(s2 * w, p, s2);
    const double de = (double)e;
    return __builtin_fma(de, kLn2Hi, __builtin_fma(de, kLn2Lo, lm));
}

// ln(1 + a) in FP64 for finite a > -1, within a few units in the last place also for small |a|:
// x = 1 + a rounded, then ln(1 + a) = ln x + ln(1 + d / x) ~ ln x + d / x with d = a - (x - 1) the
// rounding error of x (exact for -1/2 <= a <= 1: x - 1 and a - (x - 1) are then differences of
// numbers within a factor 2 of each other, Sterbenz; outside that range ln x is not small and d / x
// is below its last place); the hardware reciprocal is enough for the correction term.
// log_lt itself is accurate near 1 (its m - 1 is exact there).
__device__ __forceinline__ double log1p_lt(double a)
{
    const double x = 1. + a;
    const double d = a - (x - 1.);
    return __builtin_fma(d, __builtin_amdgcn_rcp(x), log_lt(x));
}

// Pixel coordinate exactly as the reference CPU forms it: xmin + (i * d), two roundings, no fused
// multiply-add (src/grid_gradient_CPU.cpp:89-90).
__device__ __forceinline__ double pixel_coord(double xmin, int i, double d)
{
#pragma clang fp contract(off)
    double t = (double)i * d;
    return xmin + t;
}

}  // namespace lt
