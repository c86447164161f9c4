// k-nearest neighbours on the uniform grid (knn::solve, knn.cpp:225-347).
//
// One thread per query point.  The query walks Chebyshev rings of grid cells
// around its own cell and stops once the largest kept distance is below the
// ring's minimum distance ((ring-1)/N_grid)^2 -- the reference's pruning
// rule.  The k best candidates sit in a max-heap in LDS, laid out
// entry-major ([entry][thread]) so the threads of a wave hit distinct banks.
// Keys are (FP64 dist^2, original index): identical to the reference's
// strict-< heap wherever distances differ, and a total order on ties.
// Unlike the reference's linear ring offsets (knn.cpp:58,77, which wrap
// across grid rows) cells outside the grid are skipped per axis; the
// neighbour set is the same.
#include "common.h"
#include "kernels.h"
#include "knn_core.h"

namespace pg {
namespace {

template <int DIM, int BLOCK>
__global__ void __launch_bounds__(BLOCK)
    k_knn(const double* __restrict__ sorted, const uint32_t* __restrict__ perm, int N,
          const int32_t* __restrict__ starts, const int32_t* __restrict__ counts,
          const uint32_t* __restrict__ queries, int64_t nq, int k, uint32_t* __restrict__ out) {
    extern __shared__ unsigned char smem_raw[];
    double* hd = (double*)smem_raw;
    uint32_t* hi = (uint32_t*)(hd + (size_t)k * BLOCK);
    const int64_t qi = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (qi >= nq) return;
    const uint32_t q = queries ? queries[qi] : (uint32_t)qi;
    knn_heap_query<DIM, BLOCK>(sorted, perm, N, 0, starts, counts, q, k, out + (size_t)qi * k, hd, hi);
}

// waves per SIMD the register budget is sized for: 6 for rows up to 128,
// 5 up to 256; the long restart rows (up to 1024) are rare
template <int M>
constexpr int knn_wave_occupancy() { return M <= 2 ? 6 : (M <= 4 ? 5 : 2); }

// SUB: the grid has crowded cells with sub-grids (a separate instance, so
// that the uniform case keeps its registers)
template <int DIM, int M, bool SUB>
__global__ void __launch_bounds__(64, knn_wave_occupancy<M>())
    k_knn_wave(const double* __restrict__ sorted, const uint32_t* __restrict__ perm, int N,
               const int32_t* __restrict__ starts, const int32_t* __restrict__ counts,
               const uint32_t* __restrict__ queries, int64_t nq, int k, uint32_t* __restrict__ out,
               KnnTail tail, SubGrid sg) {
    __shared__ int32_t c_start[64], c_pfx[65];
    __shared__ KnnPend pend;
    const int64_t qi = xcd_block(blockIdx.x, gridDim.x);
    if (qi >= nq) return;  // whole block
    const uint32_t q = queries ? queries[qi] : (uint32_t)qi;
    int kp = 0;
    uint32_t lo = 0;
    tail_of(tail, q, k, kp, lo);
    knn_wave_query<DIM, M, SUB>(sorted, perm, N, 0, starts, counts, q, k, out + (size_t)qi * k, c_start, c_pfx,
                                &pend, kp, lo, sg);
}

template <int DIM, int M>
void run_wave(const double* sorted, const uint32_t* perm, int N, const int32_t* starts,
              const int32_t* counts, const uint32_t* queries, int64_t nq, int k, uint32_t* out,
              KnnTail tail, SubGrid sg, hipStream_t s) {
    if (sg.m != nullptr)
        k_knn_wave<DIM, M, true><<<(unsigned)nq, 64, 0, s>>>(sorted, perm, N, starts, counts, queries, nq, k, out,
                                                             tail, sg);
    else
        k_knn_wave<DIM, M, false><<<(unsigned)nq, 64, 0, s>>>(sorted, perm, N, starts, counts, queries, nq, k, out,
                                                              tail, sg);
}

template <int DIM>
void dispatch_wave(const double* sorted, const uint32_t* perm, int N, const int32_t* starts,
                   const int32_t* counts, const uint32_t* queries, int64_t nq, int k, uint32_t* out,
                   KnnTail tail, SubGrid sg, hipStream_t s) {
    // the kept list holds 64 M entries; a resumed query keeps only its
    // k - kp missing neighbours, but M is sized for the full row
    if (k <= 64) run_wave<DIM, 1>(sorted, perm, N, starts, counts, queries, nq, k, out, tail, sg, s);
    else if (k <= 128) run_wave<DIM, 2>(sorted, perm, N, starts, counts, queries, nq, k, out, tail, sg, s);
    else if (k <= 192) run_wave<DIM, 3>(sorted, perm, N, starts, counts, queries, nq, k, out, tail, sg, s);
    else if (k <= 256) run_wave<DIM, 4>(sorted, perm, N, starts, counts, queries, nq, k, out, tail, sg, s);
    else if (k <= 512) run_wave<DIM, 8>(sorted, perm, N, starts, counts, queries, nq, k, out, tail, sg, s);
    else run_wave<DIM, 16>(sorted, perm, N, starts, counts, queries, nq, k, out, tail, sg, s);
}

template <int DIM, int BLOCK>
void run(const double* sorted, const uint32_t* perm, int N, const int32_t* starts,
         const int32_t* counts, const uint32_t* queries, int64_t nq, int k, uint32_t* out,
         hipStream_t s) {
    size_t smem = (size_t)k * BLOCK * (sizeof(double) + sizeof(uint32_t));
    if (smem > 64 * 1024)
        (void)hipFuncSetAttribute((const void*)k_knn<DIM, BLOCK>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)smem);
    unsigned nb = (unsigned)((nq + BLOCK - 1) / BLOCK);
    k_knn<DIM, BLOCK><<<nb, BLOCK, smem, s>>>(sorted, perm, N, starts, counts, queries, nq, k, out);
}

}  // namespace

void launch_knn(int dim, const double* sorted, const uint32_t* perm, int64_t n, int N,
                const int32_t* starts, const int32_t* counts, const uint32_t* queries, int64_t nq,
                int k, uint32_t* out, hipStream_t s, KnnTail tail, SubGrid sg) {
    (void)n;
    if (nq <= 0) return;
    // Long rows (and short query lists) go to the wave-per-query kernel; the
    // threshold can be moved with PGPU_KNN_WAVE_MIN_K for tuning runs.
    static const int wave_min_k = [] {
        const char* e = getenv("PGPU_KNN_WAVE_MIN_K");
        return e ? atoi(e) : 20;
    }();
    // the LDS heap holds at most 256 entries per query, and only the wave
    // kernel takes crowded cells through their sub-grids
    if (k >= wave_min_k || k > 256 || sg.m != nullptr) {
        if (dim == 2) dispatch_wave<2>(sorted, perm, N, starts, counts, queries, nq, k, out, tail, sg, s);
        else dispatch_wave<3>(sorted, perm, N, starts, counts, queries, nq, k, out, tail, sg, s);
        return;
    }
    // LDS per block = 12 B * k * BLOCK.  gfx950 gives a workgroup up to
    // 160 KB, so full 64-lane waves are used up to k = 208; beyond that a
    // 32-lane block (a half-filled wave) keeps the heap within 160 KB.
    if (k <= 208) {
        if (dim == 2) run<2, 64>(sorted, perm, N, starts, counts, queries, nq, k, out, s);
        else run<3, 64>(sorted, perm, N, starts, counts, queries, nq, k, out, s);
    } else {
        if (dim == 2) run<2, 32>(sorted, perm, N, starts, counts, queries, nq, k, out, s);
        else run<3, 32>(sorted, perm, N, starts, counts, queries, nq, k, out, s);
    }
}

}  // namespace pg
