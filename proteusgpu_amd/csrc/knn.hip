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
#include "knn_wave.h"

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
    const int t = threadIdx.x;
    const int64_t qi = (int64_t)blockIdx.x * BLOCK + t;
    if (qi >= nq) return;
    const uint32_t q = queries ? queries[qi] : (uint32_t)qi;
    double p[3] = {0, 0, 0};
#pragma unroll
    for (int d = 0; d < DIM; d++) p[d] = sorted[(size_t)DIM * q + d];
    const int c = grid_cell<DIM>(p, N);
    const int ci = c % N, cj = (c / N) % N, ck = DIM == 3 ? c / (N * N) : 0;

#define HD(e) hd[(size_t)(e) * BLOCK + t]
#define HI(e) hi[(size_t)(e) * BLOCK + t]
    auto less = [&](double da, uint32_t ia, double db, uint32_t ib) -> bool {
        return da < db || (da == db && perm[ia] < perm[ib]);
    };
    int size = 0;
    auto push = [&](double d, uint32_t id) {
        if (size < k) {
            int j = size++;
            while (j > 0) {
                int par = (j - 1) >> 1;
                double pd = HD(par);
                uint32_t pi = HI(par);
                if (!less(pd, pi, d, id)) break;
                HD(j) = pd;
                HI(j) = pi;
                j = par;
            }
            HD(j) = d;
            HI(j) = id;
        } else if (less(d, id, HD(0), HI(0))) {
            int j = 0;
            for (;;) {
                int l = 2 * j + 1, r = l + 1, m = -1;
                double md = d;
                uint32_t mi = id;
                if (l < k && less(md, mi, HD(l), HI(l))) {
                    m = l;
                    md = HD(l);
                    mi = HI(l);
                }
                if (r < k && less(md, mi, HD(r), HI(r))) {
                    m = r;
                    md = HD(r);
                    mi = HI(r);
                }
                if (m < 0) break;
                HD(j) = md;
                HI(j) = mi;
                j = m;
            }
            HD(j) = d;
            HI(j) = id;
        }
    };
    auto scan_cell = [&](int ii, int jj, int kk) {
        const int cc = ii + jj * N + kk * N * N;
        const int32_t b = starts[cc], e = b + counts[cc];
        for (int32_t s = b; s < e; s++) {
            if ((uint32_t)s == q) continue;
            double pc[3] = {0, 0, 0};
#pragma unroll
            for (int d = 0; d < DIM; d++) pc[d] = sorted[(size_t)DIM * s + d];
            push(dist2<DIM>(p, pc), (uint32_t)s);
        }
    };
    for (int ring = 0; ring < N; ring++) {
        if (ring >= 1 && size == k) {
            double md = (double)(ring - 1) / (double)N;
            if (HD(0) < md * md) break;
        }
        const int klo = DIM == 3 ? max(ck - ring, 0) : 0, khi = DIM == 3 ? min(ck + ring, N - 1) : 0;
        const int jlo = max(cj - ring, 0), jhi = min(cj + ring, N - 1);
        for (int kk = klo; kk <= khi; kk++) {
            const bool kedge = DIM == 3 && (kk == ck - ring || kk == ck + ring);
            for (int jj = jlo; jj <= jhi; jj++) {
                const bool jedge = (jj == cj - ring || jj == cj + ring);
                if (kedge || jedge || ring == 0) {
                    const int ilo = max(ci - ring, 0), ihi = min(ci + ring, N - 1);
                    for (int ii = ilo; ii <= ihi; ii++) scan_cell(ii, jj, kk);
                } else {
                    if (ci - ring >= 0) scan_cell(ci - ring, jj, kk);
                    if (ci + ring < N) scan_cell(ci + ring, jj, kk);
                }
            }
        }
    }
    // heapsort to ascending order (knn.cpp:420-426)
    for (int end = size - 1; end > 0; end--) {
        double d = HD(end);
        uint32_t id = HI(end);
        HD(end) = HD(0);
        HI(end) = HI(0);
        int j = 0;
        for (;;) {
            int l = 2 * j + 1, r = l + 1, m = -1;
            double md = d;
            uint32_t mi = id;
            if (l < end && less(md, mi, HD(l), HI(l))) {
                m = l;
                md = HD(l);
                mi = HI(l);
            }
            if (r < end && less(md, mi, HD(r), HI(r))) {
                m = r;
                md = HD(r);
                mi = HI(r);
            }
            if (m < 0) break;
            HD(j) = md;
            HI(j) = mi;
            j = m;
        }
        HD(j) = d;
        HI(j) = id;
    }
    uint32_t* o = out + (size_t)qi * k;
    for (int e = 0; e < k; e++) o[e] = e < size ? HI(e) : 0xffffffffu;
#undef HD
#undef HI
}

template <int DIM, int M>
void run_wave(const double* sorted, const uint32_t* perm, int N, const int32_t* starts,
              const int32_t* counts, const uint32_t* queries, int64_t nq, int k, uint32_t* out,
              hipStream_t s) {
    k_knn_wave<DIM, M><<<(unsigned)nq, 64, 0, s>>>(sorted, perm, N, starts, counts, queries, nq, k, out);
}

template <int DIM>
void dispatch_wave(const double* sorted, const uint32_t* perm, int N, const int32_t* starts,
                   const int32_t* counts, const uint32_t* queries, int64_t nq, int k, uint32_t* out,
                   hipStream_t s) {
    if (k <= 64) run_wave<DIM, 1>(sorted, perm, N, starts, counts, queries, nq, k, out, s);
    else if (k <= 128) run_wave<DIM, 2>(sorted, perm, N, starts, counts, queries, nq, k, out, s);
    else if (k <= 192) run_wave<DIM, 3>(sorted, perm, N, starts, counts, queries, nq, k, out, s);
    else run_wave<DIM, 4>(sorted, perm, N, starts, counts, queries, nq, k, out, s);
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
                int k, uint32_t* out, hipStream_t s) {
    (void)n;
    if (nq <= 0) return;
    // Long rows (and short query lists) go to the wave-per-query kernel; the
    // threshold can be moved with PGPU_KNN_WAVE_MIN_K for tuning runs.
    static const int wave_min_k = [] {
        const char* e = getenv("PGPU_KNN_WAVE_MIN_K");
        return e ? atoi(e) : 33;
    }();
    if (k >= wave_min_k) {
        if (dim == 2) dispatch_wave<2>(sorted, perm, N, starts, counts, queries, nq, k, out, s);
        else dispatch_wave<3>(sorted, perm, N, starts, counts, queries, nq, k, out, s);
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
