// lt_internal.h -- error reporting and device selection shared by the C-ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

namespace lt {

// record `msg` as the calling thread's last error and return `code`
int set_error(int code, const std::string& msg);
// LT_ERR_HIP with the runtime's message
int hip_error(hipError_t e, const char* where);
// the calling thread's device (made current); LT_ERR_NO_DEVICE without a GPU
int current_device(int* dev);
// fill n ints of device memory with v (a kernel, so it can sit in a captured graph)
hipError_t launch_fill_i32(int32_t* p, int32_t v, size_t n, hipStream_t s);
// fill `bytes` bytes of device memory with v (kernels only: no memset node when captured)
hipError_t launch_fill_bytes(void* dst, unsigned char v, size_t bytes, hipStream_t s);
// with `dev` current: the device address of lt_critic_device_noalloc's scratch array on `dev`, looked
// up once per device (ptr may be NULL: set-up only)
int critic_noalloc_setup(int dev, void** ptr);

}  // namespace lt

#define LT_HIP_TRY(expr)                                                \
    do {                                                                \
        hipError_t e_ = (expr);                                         \
        if (e_ != hipSuccess) return ::lt::hip_error(e_, #expr);        \
    } while (0)
