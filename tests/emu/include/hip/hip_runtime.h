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

/* the compiler built-ins the kernels use that g++ lacks, with the meaning
 * the AMDGPU back end gives them */
#include <string.h>
#include <math.h>
static inline bool __builtin_isfpclass(float x, int m) {
    const int c = fpclassify(x);
    const bool neg = signbit(x);
    if (c == FP_SUBNORMAL) return m & (neg ? 0x10 : 0x80);
    if (c == FP_NAN) return m & 0x3;
    if (c == FP_INFINITE) return m & (neg ? 0x4 : 0x200);
    if (c == FP_ZERO) return m & (neg ? 0x20 : 0x40);
    return m & (neg ? 0x8 : 0x100);
}
/* a raw buffer: base address; loads at base + lane offset + scalar offset */
struct __amdgpu_buffer_rsrc_t {
    const char *base;
};
static inline __amdgpu_buffer_rsrc_t __builtin_amdgcn_make_buffer_rsrc(void *p, short, int, int) {
    return __amdgpu_buffer_rsrc_t{(const char *)p};
}
static inline uint32_t __builtin_amdgcn_raw_buffer_load_b32(__amdgpu_buffer_rsrc_t r, int vo, int so, int) {
    uint32_t v;
    memcpy(&v, r.base + (uint32_t)vo + (uint32_t)so, 4);
    return v;
}
static inline uint64_t __builtin_amdgcn_raw_buffer_load_b64(__amdgpu_buffer_rsrc_t r, int vo, int so, int) {
    uint64_t v;
    memcpy(&v, r.base + (uint32_t)vo + (uint32_t)so, 8);
    return v;
}
static inline void __builtin_amdgcn_sched_barrier(int) {}
