// Split of a float64 value into two bfloat16 planes, x ~ hi + lo: the value is rounded
// to float32, hi = bf16(f), lo = bf16(f - hi), both round to nearest even (finite
// inputs). Shared by every kernel that writes split-BF16 GEMM operands, so that all of
// them produce the same planes for the same value.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

__device__ __forceinline__ uint16_t ccsd_f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float ccsd_bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

__device__ __forceinline__ void ccsd_split_bf16(double x, uint16_t& hi, uint16_t& lo)
{
    const float f = (float)x;
    hi = ccsd_f2bf(f);
    lo = ccsd_f2bf(f - ccsd_bf2f(hi));
}
