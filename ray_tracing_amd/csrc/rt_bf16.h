// rt_bf16.h -- bfloat16 conversion and MFMA operand types shared by the
// network kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rtk {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// round to nearest even bfloat16 (inputs are finite)
__device__ __forceinline__ uint16_t to_bf16(float x) {
    uint32_t u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float from_bf16(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

}  // namespace rtk
