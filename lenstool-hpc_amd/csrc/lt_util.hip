// lt_util.hip -- small utility kernels.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "lt_internal.h"
#include "lt_kernels.h"

namespace lt {

// Grid-stride streaming copy with 16-byte vectors.
__global__ __launch_bounds__(256) void k_copy16(uint4* __restrict__ dst, const uint4* __restrict__ src, size_t nvec)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nvec; i += stride) dst[i] = src[i];
}

// Grid-stride byte copy (unaligned buffers and the tail of aligned ones).
__global__ __launch_bounds__(256) void k_copy1(unsigned char* __restrict__ dst, const unsigned char* __restrict__ src,
                                               size_t n)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

static unsigned blocks_for(size_t n)
{
    size_t b = (n + 255) / 256;
    if (b > 8192) b = 8192;
    return (unsigned)(b ? b : 1);
}

hipError_t launch_copy(void* dst, const void* src, size_t bytes, hipStream_t s)
{
    const bool aligned = ((uintptr_t)dst % 16 == 0) && ((uintptr_t)src % 16 == 0);
    const size_t nvec = aligned ? bytes / 16 : 0;
    const size_t body = nvec * 16;
    if (nvec)
        hipLaunchKernelGGL(k_copy16, dim3(blocks_for(nvec)), dim3(256), 0, s, (uint4*)dst, (const uint4*)src, nvec);
    if (bytes > body)
        hipLaunchKernelGGL(k_copy1, dim3(blocks_for(bytes - body)), dim3(256), 0, s, (unsigned char*)dst + body,
                           (const unsigned char*)src + body, bytes - body);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_scale(double* __restrict__ x, double s, size_t n)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) x[i] *= s;
}

hipError_t launch_scale(double* x, double s, size_t n, hipStream_t st)
{
    hipLaunchKernelGGL(k_scale, dim3(blocks_for(n)), dim3(256), 0, st, x, s, n);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_fill_i32(int32_t* __restrict__ p, int32_t v, size_t n)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

hipError_t launch_fill_i32(int32_t* p, int32_t v, size_t n, hipStream_t s)
{
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_fill_i32, dim3(blocks_for(n)), dim3(256), 0, s, p, v, n);
    return hipGetLastError();
}

// Byte fill: 16-byte vectors over the aligned body, single bytes at both ends.
__global__ __launch_bounds__(256) void k_fill16(uint4* __restrict__ p, uint4 v, size_t nvec)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nvec; i += stride) p[i] = v;
}

__global__ __launch_bounds__(256) void k_fill1(unsigned char* __restrict__ p, unsigned char v, size_t n)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

hipError_t launch_fill_bytes(void* dst, unsigned char v, size_t bytes, hipStream_t s)
{
    unsigned char* p = (unsigned char*)dst;
    const size_t head = std::min(bytes, (size_t)((16 - (uintptr_t)p % 16) % 16));
    const size_t nvec = (bytes - head) / 16;
    const size_t tail = bytes - head - nvec * 16;
    const uint32_t w = 0x01010101u * v;
    if (head) hipLaunchKernelGGL(k_fill1, dim3(1), dim3(256), 0, s, p, v, head);
    if (nvec)
        hipLaunchKernelGGL(k_fill16, dim3(blocks_for(nvec)), dim3(256), 0, s, (uint4*)(p + head), make_uint4(w, w, w, w),
                           nvec);
    if (tail) hipLaunchKernelGGL(k_fill1, dim3(1), dim3(256), 0, s, p + head + nvec * 16, v, tail);
    return hipGetLastError();
}

}  // namespace lt
