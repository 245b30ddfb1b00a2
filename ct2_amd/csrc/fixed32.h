// fixed32.h -- float -> ct2 fixed32 increment, shared by the drift kernel and
// the host code that must agree with it bit for bit.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace ct2 {

// fixed32(s) for drift.cpp:12-14 (simd_fixed32(float)): sfmm's rounding is
// not part of the reference tree, so it is defined here as
// RNE(s * 2^32) mod 2^32 for finite s, and 0 for NaN or infinity.
// Every operation is exact except the single rintf, so the result does not
// depend on the compiler's choice of instructions.
__host__ __device__ inline uint32_t fixed32_from_float(float s) {
#pragma clang fp contract(off)
    if (!std::isfinite(s)) return 0u;
    if (std::fabs(s) >= 4294967296.0f) return 0u;  // s * 2^32 is a multiple of 2^41
    const float t = std::rint(s * 4294967296.0f);
    if (std::fabs(t) < 2147483648.0f) return (uint32_t)(int32_t)t;
    // |t| >= 2^31: t is a multiple of 2^8, so this reduction is exact
    const float q = std::floor(t * 0x1p-32f);
    const float r = t - q * 4294967296.0f;
    return (uint32_t)r;
}

}  // namespace ct2
