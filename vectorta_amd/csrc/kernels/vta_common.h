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

// Unconditional load from a clamped (always in-bounds) index.  Guarding a
// load with a branch makes the compiler wait for it inside the branch, which
// serialises prefetch batches; callers load unconditionally and apply their
// predicate to the value.  Requires hi > lo.
__device__ __forceinline__ float vta_ld_clamped(const float *__restrict__ p, long idx, long lo,
                                                long hi) {
    const long j = idx < lo ? lo : (idx >= hi ? hi - 1 : idx);
    return p[j];
}
