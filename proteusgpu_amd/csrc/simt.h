// Wave-level primitives used by the kernels, as thin inline wrappers so the
// kernel bodies read the same whatever the lane grouping.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pg {

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & 63u); }
__device__ __forceinline__ uint64_t wave_ballot(bool p) { return (uint64_t)__ballot(p); }
__device__ __forceinline__ void block_sync() { __syncthreads(); }

template <class T>
__device__ __forceinline__ T shfl(T v, int src, int width) {
    return __shfl(v, src, width);
}
template <class T>
__device__ __forceinline__ T shfl_xor(T v, int mask, int width) {
    return __shfl_xor(v, mask, width);
}
template <class T>
__device__ __forceinline__ T shfl_up(T v, int d, int width) {
    return __shfl_up(v, d, width);
}

// Lane l's value of v at lane l ^ M, every lane of the wave active.  Inside
// rows of 16 lanes the exchange is a DPP move (no LDS traffic): quad
// permutes for M = 1, 2, 3, the half-row mirror for 7, the row mirror for
// 15, a row rotation for 8, and 4 = 7 then 3.  M = 16 and 31 stay inside
// 32 lanes (ds_swizzle, bit-mask mode); 32 and 63 go through ds_bpermute.
template <int M>
__device__ __forceinline__ uint32_t lane_xor(uint32_t v) {
    const int x = (int)v;
    if constexpr (M == 1) return (uint32_t)__builtin_amdgcn_mov_dpp(x, 0xB1, 0xF, 0xF, false);
    else if constexpr (M == 2) return (uint32_t)__builtin_amdgcn_mov_dpp(x, 0x4E, 0xF, 0xF, false);
    else if constexpr (M == 3) return (uint32_t)__builtin_amdgcn_mov_dpp(x, 0x1B, 0xF, 0xF, false);
    else if constexpr (M == 4) return lane_xor<3>(lane_xor<7>(v));
    else if constexpr (M == 7) return (uint32_t)__builtin_amdgcn_mov_dpp(x, 0x141, 0xF, 0xF, false);
    else if constexpr (M == 8) return (uint32_t)__builtin_amdgcn_mov_dpp(x, 0x128, 0xF, 0xF, false);
    else if constexpr (M == 15) return (uint32_t)__builtin_amdgcn_mov_dpp(x, 0x140, 0xF, 0xF, false);
    else if constexpr (M == 16) return (uint32_t)__builtin_amdgcn_ds_swizzle(x, 0x401F);
    else if constexpr (M == 31) return (uint32_t)__builtin_amdgcn_ds_swizzle(x, 0x7C1F);
    else {
        static_assert(M == 32 || M == 63, "lane_xor: unsupported lane mask");
        return (uint32_t)__builtin_amdgcn_ds_bpermute((lane_id() ^ M) << 2, x);
    }
}
template <int M>
__device__ __forceinline__ double lane_xor_d(double v) {
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint64_t lo = lane_xor<M>((uint32_t)u), hi = lane_xor<M>((uint32_t)(u >> 32));
    return __builtin_bit_cast(double, lo | hi << 32);
}
// v at a lane index the whole wave agrees on (a scalar result)
__device__ __forceinline__ uint32_t read_lane(uint32_t v, int lane) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, lane);
}
__device__ __forceinline__ double read_lane_d(double v, int lane) {
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint64_t lo = read_lane((uint32_t)u, lane), hi = read_lane((uint32_t)(u >> 32), lane);
    return __builtin_bit_cast(double, lo | hi << 32);
}
// Inclusive prefix sum over the 64 lanes, every lane active: row shifts
// inside rows of 16, then the row broadcasts of lanes 15 and 31.
__device__ __forceinline__ int wave_incl_sum(int x) {
    x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xF, 0xF, false);  // row_shr:1
    x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xF, 0xF, false);  // row_shr:2
    x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xF, 0xF, false);  // row_shr:4
    x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xF, 0xF, false);  // row_shr:8
    x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xA, 0xF, false);  // row_bcast:15 into rows 1, 3
    x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xC, 0xF, false);  // row_bcast:31 into rows 2, 3
    return x;
}

// Butterfly over groups of G <= 16 lanes (DPP), every lane active: x is
// combined with lane l ^ 1, then l ^ 2, .. l ^ G/2, as op(theirs, mine).
template <int G, class F>
__device__ __forceinline__ double group_butterfly_d(double x, F op) {
    static_assert(G == 1 || G == 2 || G == 4 || G == 8 || G == 16, "group of 1..16 lanes");
    if constexpr (G >= 2) x = op(lane_xor_d<1>(x), x);
    if constexpr (G >= 4) x = op(lane_xor_d<2>(x), x);
    if constexpr (G >= 8) x = op(lane_xor_d<4>(x), x);
    if constexpr (G >= 16) x = op(lane_xor_d<8>(x), x);
    return x;
}
// Inclusive prefix sum inside groups of G <= 16 lanes (gl = lane index in
// its group), every lane active: row shifts by DPP, kept only where the
// source lane lies in the same group.
template <int G>
__device__ __forceinline__ int group_incl_sum(int x, int gl) {
    static_assert(G == 1 || G == 2 || G == 4 || G == 8 || G == 16, "group of 1..16 lanes");
    if constexpr (G >= 2) {
        const int t = __builtin_amdgcn_update_dpp(0, x, 0x111, 0xF, 0xF, false);  // row_shr:1
        if (gl >= 1) x += t;
    }
    if constexpr (G >= 4) {
        const int t = __builtin_amdgcn_update_dpp(0, x, 0x112, 0xF, 0xF, false);  // row_shr:2
        if (gl >= 2) x += t;
    }
    if constexpr (G >= 8) {
        const int t = __builtin_amdgcn_update_dpp(0, x, 0x114, 0xF, 0xF, false);  // row_shr:4
        if (gl >= 4) x += t;
    }
    if constexpr (G >= 16) {
        const int t = __builtin_amdgcn_update_dpp(0, x, 0x118, 0xF, 0xF, false);  // row_shr:8
        if (gl >= 8) x += t;
    }
    return x;
}
// The value of the last lane of each group of G <= 16 lanes (ds_swizzle,
// bit-mask mode: source lane (l & ~(G - 1)) | (G - 1) inside 32 lanes).
template <int G>
__device__ __forceinline__ int group_last(int x) {
    static_assert(G == 1 || G == 2 || G == 4 || G == 8 || G == 16, "group of 1..16 lanes");
    constexpr int and_mask = 0x1F & ~(G - 1), or_mask = G - 1;
    return __builtin_amdgcn_ds_swizzle(x, and_mask | (or_mask << 5));
}

__device__ __forceinline__ int popc64(uint64_t x) { return __popcll(x); }
// index of the lowest set bit, x != 0
__device__ __forceinline__ int ffs64(uint64_t x) { return __ffsll((unsigned long long)x) - 1; }

}  // namespace pg
