// Shared host/device helpers for libvxmix (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/vxmix.h"

namespace vx {

// Per-thread message of the most recent failure (read by vx_last_error()).
std::string& last_error();

inline int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    last_error() = buf;
    return code;
}

// Check the launch that was just enqueued.
inline int check_launch(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        return fail(VX_ERR_LAUNCH, "%s: %s", what, hipGetErrorString(e));
    }
    return VX_OK;
}

inline bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

constexpr int kMaxNz = 128;

// Blocks are dispatched round-robin over the 8 XCDs, each with its own L2.
// With gridDim.x a multiple of 8 this gives block b the index of the
// (b / 8)-th block of the x-th eighth of the grid (x = b % 8), so that each
// XCD walks a contiguous range and neighbouring rows share its L2.
__device__ __forceinline__ int64_t xcd_contiguous_block() {
    const int64_t b = blockIdx.x;
    return (b & 7) * (int64_t)(gridDim.x >> 3) + (b >> 3);
}

inline unsigned blocks_mult8(int64_t threads, int block) {
    return (unsigned)(((threads + block - 1) / block + 7) / 8 * 8);
}

}  // namespace vx
