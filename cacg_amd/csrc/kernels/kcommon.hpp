// Shared device-side definitions for the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cacg_kernels.h"

#define CACG_WAVE 64

// Fragment type of the FP64 MFMA accumulator (v_mfma_f64_16x16x4f64):
// lane l holds D[l/16 + 4*r][l%16], r = 0..3 (measured on gfx950: the four
// lane groups interleave rows).  Operands: lane l supplies A[l%16][l/16] and
// B[l/16][l%16].
typedef double dbl4 __attribute__((ext_vector_type(4)));

// Products and sums that must round separately (the reference is compiled for
// x86-64 without FMA).  Kernels whose results are meant to equal the
// reference bit for bit use these two helpers only; the clang pragma keeps
// the compiler from contracting them into fused multiply-adds.
__device__ __forceinline__ double mul_rn(double a, double b)
{
#pragma clang fp contract(off)
    return a * b;
}
__device__ __forceinline__ double add_rn(double a, double b)
{
#pragma clang fp contract(off)
    return a + b;
}
__device__ __forceinline__ double sub_rn(double a, double b)
{
#pragma clang fp contract(off)
    return a - b;
}

// FP32 versions for the FP32-basis kernels (no parity target; the same
// separate rounding keeps their results independent of compiler contraction)
__device__ __forceinline__ float mul_rn(float a, float b)
{
#pragma clang fp contract(off)
    return a * b;
}
__device__ __forceinline__ float add_rn(float a, float b)
{
#pragma clang fp contract(off)
    return a + b;
}
__device__ __forceinline__ float sub_rn(float a, float b)
{
#pragma clang fp contract(off)
    return a - b;
}

__device__ __forceinline__ dbl4 mfma16x16x4(double a, double b, dbl4 c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

static inline int cacg_grid(long work, int per_block, int cap)
{
    long g = (work + per_block - 1) / per_block;
    if (g < 1)
        g = 1;
    if (cap > 0 && g > cap)
        g = cap;
    return (int)g;
}

// Kernels that take more than 64 KiB of dynamic LDS (up to the 160 KiB of a
// gfx950 CU) declare it before launch.
template <typename K>
static inline void cacg_allow_lds(K kernel, size_t bytes)
{
    if (bytes > 65536)
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

#define CACG_LAUNCH_CHECK()                                                                       \
    do {                                                                                          \
        hipError_t e_ = hipGetLastError();                                                        \
        if (e_ != hipSuccess)                                                                     \
            return (int)e_;                                                                       \
    } while (0)
