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

__device__ __forceinline__ int popc64(uint64_t x) { return __popcll(x); }
// index of the lowest set bit, x != 0
__device__ __forceinline__ int ffs64(uint64_t x) { return __ffsll((unsigned long long)x) - 1; }

}  // namespace pg
