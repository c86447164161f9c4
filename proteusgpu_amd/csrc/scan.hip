// Device-wide exclusive prefix sum (three passes: tile sums, scan of tile
// sums, tile scan + offset).  Tiles of 1024 int32 per 256-thread block,
// wave-level scans with DPP-backed shuffles.
#include "common.h"
#include "kernels.h"

namespace pg {
namespace {

constexpr int TB = 256;
constexpr int PER = 4;
constexpr int TILE = TB * PER;

PG_DEV int32_t wave_incl_scan(int32_t v) {
    const int lane = threadIdx.x & (WAVE - 1);
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
        int32_t t = __shfl_up(v, d, WAVE);
        if (lane >= d) v += t;
    }
    return v;
}

// Exclusive scan of one value per thread across the block; returns the
// block total through *tot.
PG_DEV int32_t block_excl_scan(int32_t v, int32_t* tot) {
    __shared__ int32_t wsum[TB / WAVE];
    const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
    int32_t inc = wave_incl_scan(v);
    if (lane == WAVE - 1) wsum[w] = inc;
    __syncthreads();
    int32_t base = 0, all = 0;
#pragma unroll
    for (int i = 0; i < TB / WAVE; i++) {
        int32_t s = wsum[i];
        if (i < w) base += s;
        all += s;
    }
    __syncthreads();
    *tot = all;
    return base + inc - v;
}

__global__ void __launch_bounds__(TB) k_tile_sums(const int32_t* in, int64_t n, int32_t* sums) {
    int64_t base = (int64_t)blockIdx.x * TILE + threadIdx.x * PER;
    int32_t v = 0;
#pragma unroll
    for (int j = 0; j < PER; j++)
        if (base + j < n) v += in[base + j];
    int32_t tot;
    block_excl_scan(v, &tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// Exclusive scan of the tile sums in place (one block, sequential chunks).
__global__ void __launch_bounds__(TB) k_scan_sums(int32_t* sums, int64_t nt, int32_t* total) {
    int32_t carry = 0;
    for (int64_t c = 0; c < nt; c += TB) {
        int64_t i = c + threadIdx.x;
        int32_t v = i < nt ? sums[i] : 0;
        int32_t tot;
        int32_t ex = block_excl_scan(v, &tot);
        if (i < nt) sums[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0 && total) *total = carry;
}

__global__ void __launch_bounds__(TB) k_tile_apply(const int32_t* in, int32_t* out, int64_t n,
                                                   const int32_t* sums) {
    int64_t base = (int64_t)blockIdx.x * TILE + threadIdx.x * PER;
    int32_t v[PER];
    int32_t s = 0;
#pragma unroll
    for (int j = 0; j < PER; j++) {
        v[j] = (base + j < n) ? in[base + j] : 0;
        s += v[j];
    }
    int32_t tot;
    int32_t ex = block_excl_scan(s, &tot) + sums[blockIdx.x];
#pragma unroll
    for (int j = 0; j < PER; j++) {
        if (base + j < n) out[base + j] = ex;
        ex += v[j];
    }
}

}  // namespace

int64_t scan_tmp_elems(int64_t n) { return (n + TILE - 1) / TILE + 1; }

void scan_exclusive_i32(const int32_t* in, int32_t* out, int64_t n, int32_t* tmp, int32_t* total,
                        hipStream_t s) {
    int64_t nt = (n + TILE - 1) / TILE;
    if (nt == 0) {
        if (total) (void)hipMemsetAsync(total, 0, sizeof(int32_t), s);
        return;
    }
    k_tile_sums<<<dim3((unsigned)nt), dim3(TB), 0, s>>>(in, n, tmp);
    k_scan_sums<<<1, TB, 0, s>>>(tmp, nt, total);
    k_tile_apply<<<dim3((unsigned)nt), dim3(TB), 0, s>>>(in, out, n, tmp);
}

}  // namespace pg
