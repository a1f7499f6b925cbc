// issue_rate.hip -- VALU issue rates of the instruction forms the map kernels are built from, on the
// GPU it runs on: wavefront instructions per nanosecond per SIMD for FP32 FMAs with all-VGPR operands
// (1, 2, 4 and 8 independent dependency chains per thread), FP32 FMAs with one SGPR operand, FP64
// FMAs, and the FP32 / FP64 reciprocal square root. Every launch runs 16 wavefronts' worth of
// threads per SIMD (so that independent waves hide latency) and writes each thread's result, so
// nothing is folded away. Built with the library's device flags (packed FP32 disabled).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <vector>

namespace {

constexpr int kIters = 4096;   // instructions per thread: kIters * 16
constexpr int kThreads = 256;

template <int CH, bool SGPR>
__global__ __launch_bounds__(kThreads) void k_fma32(const float* __restrict__ in, float* __restrict__ out, float sa,
                                                    float sb)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const float a = SGPR ? sa : in[t & 1023];
    const float b = SGPR ? sb : in[(t + 7) & 1023];
    float x[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) x[c] = in[(t + 13 * c) & 1023];
    for (int i = 0; i < kIters / CH; ++i) {
#pragma unroll
        for (int u = 0; u < 16; ++u) {
#pragma unroll
            for (int c = 0; c < CH; ++c) x[c] = __builtin_fmaf(x[c], a, b);
        }
    }
    float s = 0.f;
#pragma unroll
    for (int c = 0; c < CH; ++c) s += x[c];
    out[t] = s;
}

template <int CH>
__global__ __launch_bounds__(kThreads) void k_fma64(const double* __restrict__ in, double* __restrict__ out)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const double a = in[t & 1023], b = in[(t + 7) & 1023];
    double x[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) x[c] = in[(t + 13 * c) & 1023];
    for (int i = 0; i < kIters / CH; ++i) {
#pragma unroll
        for (int u = 0; u < 16; ++u) {
#pragma unroll
            for (int c = 0; c < CH; ++c) x[c] = __builtin_fma(x[c], a, b);
        }
    }
    double s = 0.;
#pragma unroll
    for (int c = 0; c < CH; ++c) s += x[c];
    out[t] = s;
}

// 4 chains of x <- rsq(x) (x stays near 1)
__global__ __launch_bounds__(kThreads) void k_rsq32(const float* __restrict__ in, float* __restrict__ out)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    float x[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) x[c] = in[(t + 13 * c) & 1023];
    for (int i = 0; i < kIters / 4; ++i) {
#pragma unroll
        for (int u = 0; u < 16; ++u) {
#pragma unroll
            for (int c = 0; c < 4; ++c) x[c] = __builtin_amdgcn_rsqf(x[c]);
        }
    }
    out[t] = x[0] + x[1] + x[2] + x[3];
}

__global__ __launch_bounds__(kThreads) void k_rsq64(const double* __restrict__ in, double* __restrict__ out)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    double x[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) x[c] = in[(t + 13 * c) & 1023];
    for (int i = 0; i < kIters / 4; ++i) {
#pragma unroll
        for (int u = 0; u < 16; ++u) {
#pragma unroll
            for (int c = 0; c < 4; ++c) x[c] = __builtin_amdgcn_rsq(x[c]);
        }
    }
    out[t] = x[0] + x[1] + x[2] + x[3];
}

}  // namespace

#define CHECK(x)                                                                                  \
    do {                                                                                          \
        hipError_t e_ = (x);                                                                      \
        if (e_ != hipSuccess) {                                                                   \
            std::printf("HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__);            \
            return 1;                                                                             \
        }                                                                                         \
    } while (0)

int main()
{
    int dev = 0, cus = 0;
    CHECK(hipGetDevice(&dev));
    CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const int simds = 4 * cus;
    const int blocks = simds * 16 * 64 / kThreads;  // 16 wavefronts per SIMD
    const size_t n = (size_t)blocks * kThreads;
    float *in32 = nullptr, *out32 = nullptr;
    double *in64 = nullptr, *out64 = nullptr;
    CHECK(hipMalloc(&in32, 1024 * sizeof(float)));
    CHECK(hipMalloc(&in64, 1024 * sizeof(double)));
    CHECK(hipMalloc(&out32, n * sizeof(float)));
    CHECK(hipMalloc(&out64, n * sizeof(double)));
    std::vector<float> h32(1024);
    std::vector<double> h64(1024);
    for (int i = 0; i < 1024; ++i) {
        h32[i] = 0.999f + 1e-6f * (float)i;
        h64[i] = 0.999 + 1e-6 * i;
    }
    CHECK(hipMemcpy(in32, h32.data(), 1024 * sizeof(float), hipMemcpyHostToDevice));
    CHECK(hipMemcpy(in64, h64.data(), 1024 * sizeof(double), hipMemcpyHostToDevice));
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    const double winstr = (double)n / 64.0 * kIters * 16.0;  // wavefront instructions per launch
    const dim3 g(blocks), b(kThreads);
    auto run = [&](const char* name, auto launch) -> int {
        for (int w = 0; w < 2; ++w) launch();
        CHECK(hipGetLastError());
        CHECK(hipDeviceSynchronize());
        CHECK(hipEventRecord(e0));
        for (int r = 0; r < 5; ++r) launch();
        CHECK(hipEventRecord(e1));
        CHECK(hipEventSynchronize(e1));
        float ms = 0.f;
        CHECK(hipEventElapsedTime(&ms, e0, e1));
        std::printf("rate %-22s %.3f wave-instr/ns/SIMD\n", name, winstr / (ms / 5.0 * 1e6) / simds);
        return 0;
    };
    int rc = 0;
    rc |= run("fp32_fma_vgpr_1", [&] { hipLaunchKernelGGL((k_fma32<1, false>), g, b, 0, 0, in32, out32, 0.f, 0.f); });
    rc |= run("fp32_fma_vgpr_2", [&] { hipLaunchKernelGGL((k_fma32<2, false>), g, b, 0, 0, in32, out32, 0.f, 0.f); });
    rc |= run("fp32_fma_vgpr_4", [&] { hipLaunchKernelGGL((k_fma32<4, false>), g, b, 0, 0, in32, out32, 0.f, 0.f); });
    rc |= run("fp32_fma_vgpr_8", [&] { hipLaunchKernelGGL((k_fma32<8, false>), g, b, 0, 0, in32, out32, 0.f, 0.f); });
    rc |= run("fp32_fma_sgpr_4", [&] {
        hipLaunchKernelGGL((k_fma32<4, true>), g, b, 0, 0, in32, out32, 0.9991f, 1e-4f);
    });
    rc |= run("fp64_fma_vgpr_4", [&] { hipLaunchKernelGGL((k_fma64<4>), g, b, 0, 0, in64, out64); });
    rc |= run("fp32_rsq_4", [&] { hipLaunchKernelGGL(k_rsq32, g, b, 0, 0, in32, out32); });
    rc |= run("fp64_rsq_4", [&] { hipLaunchKernelGGL(k_rsq64, g, b, 0, 0, in64, out64); });
    CHECK(hipFree(in32));
    CHECK(hipFree(in64));
    CHECK(hipFree(out32));
    CHECK(hipFree(out64));
    return rc;
}
