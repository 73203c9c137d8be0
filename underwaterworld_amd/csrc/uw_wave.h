// Wavefront-level prefix sums with DPP (data-parallel primitives) for 64-lane gfx9 waves.
//
// Inclusive scan in six steps, the standard CDNA pattern: Hillis-Steele within each 16-lane row
// with row_shr:1, 2, 4, 8 (a lane whose source lies outside its row adds 0), then the row totals
// are carried across rows with row_bcast:15 (lane 15 of rows 0 and 2 into rows 1 and 3) and
// row_bcast:31 (lane 31 into rows 2 and 3). Each step is a VALU move with a modifier, so the scan
// stays in registers instead of going through LDS permutes.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define UW_DPP_ROW_SHR(n) (0x110 + (n))
#define UW_DPP_ROW_BCAST15 0x142
#define UW_DPP_ROW_BCAST31 0x143

template <int CTRL, int ROW_MASK>
__device__ __forceinline__ uint32_t uw_dpp_u32(uint32_t v) {
    return uint32_t(__builtin_amdgcn_update_dpp(0, int(v), CTRL, ROW_MASK, 0xF, false));
}

__device__ __forceinline__ uint32_t uw_wave_incl_scan(uint32_t v) {
    v += uw_dpp_u32<UW_DPP_ROW_SHR(1), 0xF>(v);
    v += uw_dpp_u32<UW_DPP_ROW_SHR(2), 0xF>(v);
    v += uw_dpp_u32<UW_DPP_ROW_SHR(4), 0xF>(v);
    v += uw_dpp_u32<UW_DPP_ROW_SHR(8), 0xF>(v);
    v += uw_dpp_u32<UW_DPP_ROW_BCAST15, 0xA>(v);
    v += uw_dpp_u32<UW_DPP_ROW_BCAST31, 0xC>(v);
    return v;
}

template <int CTRL, int ROW_MASK>
__device__ __forceinline__ unsigned long long uw_dpp_u64(unsigned long long v) {
    const uint32_t lo = uw_dpp_u32<CTRL, ROW_MASK>(uint32_t(v));
    const uint32_t hi = uw_dpp_u32<CTRL, ROW_MASK>(uint32_t(v >> 32));
    return (static_cast<unsigned long long>(hi) << 32) | lo;
}

__device__ __forceinline__ unsigned long long uw_wave_incl_scan(unsigned long long v) {
    v += uw_dpp_u64<UW_DPP_ROW_SHR(1), 0xF>(v);
    v += uw_dpp_u64<UW_DPP_ROW_SHR(2), 0xF>(v);
    v += uw_dpp_u64<UW_DPP_ROW_SHR(4), 0xF>(v);
    v += uw_dpp_u64<UW_DPP_ROW_SHR(8), 0xF>(v);
    v += uw_dpp_u64<UW_DPP_ROW_BCAST15, 0xA>(v);
    v += uw_dpp_u64<UW_DPP_ROW_BCAST31, 0xC>(v);
    return v;
}

// Value of lane 63 (the wave total after an inclusive scan), as a wave-uniform value.
__device__ __forceinline__ uint32_t uw_wave_last(uint32_t v) { return uint32_t(__builtin_amdgcn_readlane(int(v), 63)); }
__device__ __forceinline__ unsigned long long uw_wave_last(unsigned long long v) {
    const uint32_t lo = uw_wave_last(uint32_t(v)), hi = uw_wave_last(uint32_t(v >> 32));
    return (static_cast<unsigned long long>(hi) << 32) | lo;
}

// Sum over the wave's 64 lanes of an f64 value (the inclusive scan's last lane), wave-uniform; the
// addition order is fixed, so the result does not vary between runs.
__device__ __forceinline__ double uw_wave_sum_f64(double v) {
    unsigned long long u = static_cast<unsigned long long>(__double_as_longlong(v));
#define UW_F64_STEP(CTRL, MASK)                                                                   \
    v += __longlong_as_double(static_cast<long long>(uw_dpp_u64<CTRL, MASK>(u)));                \
    u = static_cast<unsigned long long>(__double_as_longlong(v));
    UW_F64_STEP(UW_DPP_ROW_SHR(1), 0xF)
    UW_F64_STEP(UW_DPP_ROW_SHR(2), 0xF)
    UW_F64_STEP(UW_DPP_ROW_SHR(4), 0xF)
    UW_F64_STEP(UW_DPP_ROW_SHR(8), 0xF)
    UW_F64_STEP(UW_DPP_ROW_BCAST15, 0xA)
    UW_F64_STEP(UW_DPP_ROW_BCAST31, 0xC)
#undef UW_F64_STEP
    return __longlong_as_double(static_cast<long long>(uw_wave_last(u)));
}
