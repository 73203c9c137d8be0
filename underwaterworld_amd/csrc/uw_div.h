// Division by small runtime divisors (chunk resolutions, sample counts, ring sizes) with a 32-bit
// reciprocal multiply instead of the long integer-division sequence.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

// Division by a small runtime divisor d via a 32-bit reciprocal: exact while g * d < 2^32.
struct uw_divider {
    uint32_t d, m;
};
__host__ __device__ __forceinline__ uw_divider uw_make_divider(uint32_t d) {
    uw_divider r;
    r.d = d;
    r.m = 0xFFFFFFFFu / d + 1u;  // ceil(2^32 / d) for d >= 2 (d == 1 is special-cased by uw_div)
    return r;
}
__device__ __forceinline__ uint32_t uw_div(uint32_t g, uw_divider v) { return v.d == 1u ? g : __umulhi(g, v.m); }
__device__ __forceinline__ uint32_t uw_mod(uint32_t g, uw_divider v) { return g - uw_div(g, v) * v.d; }

