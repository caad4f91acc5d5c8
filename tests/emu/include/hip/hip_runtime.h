/* Test-only stand-in for <hip/hip_runtime.h>: lets g++ compile generated
 * stage kernels as plain C++ so tests can run them lane by lane on the CPU
 * (tests/emu/kun_emu.cpp). Only the qualifiers and built-in index variables
 * the generated kernels use are provided. Never part of the product build. */
#pragma once
#include <stddef.h>
#include <stdint.h>

#define __device__
#define __host__
#define __global__
#define __forceinline__ inline __attribute__((always_inline))
#define __launch_bounds__(n)
/* every emulated lane runs to completion before the next starts and the
 * kernels only touch their own lane's LDS column, so one static array per
 * kernel stands in for a workgroup's LDS */
#define __shared__ static

struct kun_emu_dim3 {
    unsigned x, y, z;
};
extern thread_local kun_emu_dim3 blockIdx, threadIdx, gridDim, blockDim;
