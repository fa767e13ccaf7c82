// wave_ops.h -- exact lane exchanges within a 64-lane gfx950 wave, without LDS.
//
// xor_lane(v, m) gives every lane i the value lane i ^ m holds (what __shfl_xor(v, m, 64)
// returns), and quad_lane(v, q) the value of lane 4 (i / 4) + q (__shfl(v, 4 (i / 4) + q, 64)).
// __shfl_xor compiles to ds_bpermute_b32, a round trip through the LDS crossbar of about a
// hundred cycles per 32-bit word; the reductions of the update kernels chain six of them per
// value.  Here each exchange is a DPP move or a gfx950 permlane swap, a VALU instruction:
//   m = 1, 2      quad_perm [1,0,3,2] / [2,3,0,1]
//   m = 4         row_half_mirror (i -> i ^ 7 within 8 lanes), then quad_perm [3,2,1,0] (^ 3)
//   m = 8         row_ror:8 (a rotation by 8 within 16 lanes is i -> i ^ 8)
//   m = 16, 32    v_permlane16_swap / v_permlane32_swap of the value with itself: the two
//                 results hold the value with rows 1 and 3 (rows 2 and 3) exchanged for rows 0
//                 and 2 (rows 0 and 1), and each lane takes the one that holds its partner.
// The data movement is that of the LDS form, so results are bitwise the same.  Every lane of
// the wave must be active (no call sits in lane-divergent control flow).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dtx {

template <int M>
__device__ __forceinline__ uint32_t xor_lane_u32(uint32_t v) {
    static_assert(M == 1 || M == 2 || M == 4 || M == 8 || M == 16 || M == 32, "xor distance");
    const int x = (int)v;
    if constexpr (M == 1) {
        return (uint32_t)__builtin_amdgcn_mov_dpp(x, 0xB1, 0xF, 0xF, false);
    } else if constexpr (M == 2) {
        return (uint32_t)__builtin_amdgcn_mov_dpp(x, 0x4E, 0xF, 0xF, false);
    } else if constexpr (M == 4) {
        const int t = __builtin_amdgcn_mov_dpp(x, 0x141, 0xF, 0xF, false);
        return (uint32_t)__builtin_amdgcn_mov_dpp(t, 0x1B, 0xF, 0xF, false);
    } else if constexpr (M == 8) {
        return (uint32_t)__builtin_amdgcn_mov_dpp(x, 0x128, 0xF, 0xF, false);
    } else if constexpr (M == 16) {
        const auto p = __builtin_amdgcn_permlane16_swap(v, v, false, false);
        return (threadIdx.x & 16) ? p[0] : p[1];
    } else {
        const auto p = __builtin_amdgcn_permlane32_swap(v, v, false, false);
        return (threadIdx.x & 32) ? p[0] : p[1];
    }
}

template <int Q>
__device__ __forceinline__ uint32_t quad_lane_u32(uint32_t v) {
    static_assert(Q >= 0 && Q < 4, "lane within a quad");
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, Q * 0x55, 0xF, 0xF, false);
}

template <int M>
__device__ __forceinline__ double xor_lane(double v) {
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint64_t lo = xor_lane_u32<M>((uint32_t)u), hi = xor_lane_u32<M>((uint32_t)(u >> 32));
    return __builtin_bit_cast(double, lo | (hi << 32));
}

template <int Q>
__device__ __forceinline__ double quad_lane(double v) {
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint64_t lo = quad_lane_u32<Q>((uint32_t)u), hi = quad_lane_u32<Q>((uint32_t)(u >> 32));
    return __builtin_bit_cast(double, lo | (hi << 32));
}

template <int Q>
__device__ __forceinline__ int quad_lane(int v) {
    return (int)quad_lane_u32<Q>((uint32_t)v);
}

template <int M>
__device__ __forceinline__ unsigned long long xor_lane(unsigned long long v) {
    const uint64_t lo = xor_lane_u32<M>((uint32_t)v), hi = xor_lane_u32<M>((uint32_t)(v >> 32));
    return lo | (hi << 32);
}

// xor_lane with a distance known only at run time (wave-uniform: a uniform branch), or folded
// to one case after unrolling.
template <typename T>
__device__ __forceinline__ T xor_lane_rt(T v, int m) {
    switch (m) {
        case 1: return xor_lane<1>(v);
        case 2: return xor_lane<2>(v);
        case 4: return xor_lane<4>(v);
        case 8: return xor_lane<8>(v);
        case 16: return xor_lane<16>(v);
        default: return xor_lane<32>(v);
    }
}

template <typename T>
__device__ __forceinline__ T quad_lane_rt(T v, int q) {
    switch (q) {
        case 0: return quad_lane<0>(v);
        case 1: return quad_lane<1>(v);
        case 2: return quad_lane<2>(v);
        default: return quad_lane<3>(v);
    }
}

}  // namespace dtx
