// Shared device helpers for the gfx950 batch-indicator kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define VTA_QNAN_F32 __builtin_bit_cast(float, 0x7fc00000u)

__device__ __forceinline__ bool vta_isfinite(float x) {
    return (__builtin_bit_cast(uint32_t, x) & 0x7f800000u) != 0x7f800000u;
}
__device__ __forceinline__ bool vta_isfinite_d(double x) {
    return (__builtin_bit_cast(uint64_t, x) & 0x7ff0000000000000ull) != 0x7ff0000000000000ull;
}

// Round-to-nearest-even f32 -> bf16 (bit pattern), NaN stays NaN.
__device__ __forceinline__ uint16_t vta_bf16_rne(float x) {
    __bf16 h = (__bf16)x;
    return __builtin_bit_cast(uint16_t, h);
}
__device__ __forceinline__ float vta_bf16_to_f32(uint16_t h) {
    return __builtin_bit_cast(float, (uint32_t)h << 16);
}
