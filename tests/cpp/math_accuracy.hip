// Accuracy of the library's FP64 atan2_lt / atan2p_lt / log_lt (lenstool-hpc_amd/csrc/lt_math.h) on the GPU
// against the host's long double atan2l / logl, in units in the last place, over 4M seeded random inputs
// (all sign combinations, magnitudes 1e-13..1e13, arguments of log near 1; atan2p_lt, for x > 0, at
// (y, |x|)). Prints the three maxima; exit 0.
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>
#include "lt_math.h"
using namespace lt;
__global__ void k(const double* y, const double* x, double* at, double* ap, double* lg, int n)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    at[t] = atan2_lt(y[t], x[t]);
    ap[t] = atan2p_lt(y[t], fabs(x[t]));
    lg[t] = log_lt(fabs(x[t]));
}
// wave_max64 over wavefronts of random values (one in eight lanes NaN-free negative, the rest random
// magnitudes), one result per lane
__global__ void kmax(const double* v, double* out, int n)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    out[t] = wave_max64(v[t]);
}
static double ulp(double v) { return std::nextafter(std::fabs(v), INFINITY) - std::fabs(v); }
int main()
{
    const int n = 1 << 22;
    std::mt19937_64 g(7);
    std::uniform_real_distribution<double> u(-1, 1), le(-30, 30);
    std::vector<double> y(n), x(n), at(n), ap(n), lg(n);
    for (int i = 0; i < n; ++i) {
        x[i] = std::copysign(std::exp(le(g)), u(g));
        y[i] = (i % 4 == 0) ? x[i] * u(g) : std::copysign(std::exp(le(g)), u(g));
        if (i % 97 == 0) x[i] = 1.0 + 1e-6 * u(g);  // log near 1
    }
    double *dy, *dx, *da, *dp, *dl;
    if (hipMalloc(&dy, n * 8) || hipMalloc(&dx, n * 8) || hipMalloc(&da, n * 8) || hipMalloc(&dp, n * 8) ||
        hipMalloc(&dl, n * 8))
        return 1;
    if (hipMemcpy(dy, y.data(), n * 8, hipMemcpyHostToDevice) || hipMemcpy(dx, x.data(), n * 8, hipMemcpyHostToDevice))
        return 1;
    hipLaunchKernelGGL(k, dim3((n + 255) / 256), dim3(256), 0, 0, dy, dx, da, dp, dl, n);
    if (hipMemcpy(at.data(), da, n * 8, hipMemcpyDeviceToHost) || hipMemcpy(ap.data(), dp, n * 8, hipMemcpyDeviceToHost) ||
        hipMemcpy(lg.data(), dl, n * 8, hipMemcpyDeviceToHost))
        return 1;
    double wa = 0, wp = 0, wl = 0, wl_abs = 0;
    for (int i = 0; i < n; ++i) {
        long double ra = atan2l((long double)y[i], (long double)x[i]);
        long double rl = logl(fabsl((long double)x[i]));
        wa = std::fmax(wa, (double)(fabsl(at[i] - ra) / ulp((double)ra)));
        long double rp = atan2l((long double)y[i], fabsl((long double)x[i]));
        wp = std::fmax(wp, (double)(fabsl(ap[i] - rp) / ulp((double)rp)));
        if (rl != 0) wl = std::fmax(wl, (double)(fabsl(lg[i] - rl) / ulp((double)rl)));
        wl_abs = std::fmax(wl_abs, (double)fabsl(lg[i] - rl));
    }
    printf("atan2_lt max error %.2f ulp, log_lt max error %.2f ulp (max abs %.2e), atan2p_lt max error %.2f ulp over %d "
           "inputs\n",
           wa, wl, wl_abs, wp, n);
    // wave_max64: every lane of every wavefront gets that wavefront's maximum (n is a multiple of 256)
    for (int i = 0; i < n; ++i) x[i] = (i % 8 == 3) ? -std::exp(le(g)) : std::copysign(std::exp(le(g)), u(g));
    if (hipMemcpy(dx, x.data(), n * 8, hipMemcpyHostToDevice)) return 1;
    hipLaunchKernelGGL(kmax, dim3(n / 256), dim3(256), 0, 0, dx, da, n);
    if (hipMemcpy(at.data(), da, n * 8, hipMemcpyDeviceToHost)) return 1;
    long bad = 0;
    for (int w = 0; w < n / 64; ++w) {
        double m = x[64 * w];
        for (int l = 1; l < 64; ++l) m = std::fmax(m, x[64 * w + l]);
        for (int l = 0; l < 64; ++l) bad += at[64 * w + l] != m;
    }
    printf("wave_max64 lanes not equal to their wavefront's maximum: %ld of %d\n", bad, n);
    return 0;
}
