// Wavefront sums of doubles on DPP lane moves and readlane, with no LDS
// round trip (__shfl_xor lowers to ds_bpermute).  Every lane of a group gets
// the same, bitwise identical sum, added in a fixed order.  Call with all
// lanes of the group active.
#pragma once

#include <hip/hip_runtime.h>

namespace bx {

// DPP controls (GFX9): quad_perm [1,0,3,2], quad_perm [2,3,0,1],
// row_half_mirror (lane i <- 7-i of its 8), row_mirror (lane i <- 15-i of its 16)
constexpr int kDppXor1 = 0xB1, kDppXor2 = 0x4E, kDppHalfMirror = 0x141, kDppMirror = 0x140;

template <int CTRL>
__device__ __forceinline__ double dpp_mov(double v) {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)b, CTRL, 0xF, 0xF, false);
    const unsigned hi =
        (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)(b >> 32), CTRL, 0xF, 0xF, false);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

__device__ __forceinline__ double readlane_d(double v, int lane) {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, lane);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), lane);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// sum over aligned groups of 8 lanes
__device__ __forceinline__ double group_sum8(double v) {
    v += dpp_mov<kDppXor1>(v);
    v += dpp_mov<kDppXor2>(v);
    v += dpp_mov<kDppHalfMirror>(v);
    return v;
}
// ... of 16 lanes
__device__ __forceinline__ double group_sum16(double v) {
    v = group_sum8(v);
    v += dpp_mov<kDppMirror>(v);
    return v;
}
// ... of 32 lanes
__device__ __forceinline__ double group_sum32(double v) {
    v = group_sum16(v);
    const double s0 = readlane_d(v, 0) + readlane_d(v, 16);
    const double s1 = readlane_d(v, 32) + readlane_d(v, 48);
    return (threadIdx.x & 32) ? s1 : s0;
}
// ... of the whole wave
__device__ __forceinline__ double wave_sum(double v) {
    v = group_sum16(v);
    return (readlane_d(v, 0) + readlane_d(v, 16)) + (readlane_d(v, 32) + readlane_d(v, 48));
}

}  // namespace bx
