// Periodic ghosts and the uniform-grid counting sort (knn::init).
//
// Ghosts (periodic_mesh.cpp:63-126): a point inside the buff-wide layer of a
// face, edge or corner region of the unit box gets a copy shifted by +-1
// along those axes.  Ghosts are appended point-major in the reference's
// region order, then every point is scaled into [0,1]^d by 1/(1+2 buff).
//
// Grid (knn.cpp:119-222): N_grid^d cells of width 1/N_grid; per-cell counts,
// an exclusive scan of the counts in cell-id order gives each cell's range,
// and within a range the points are ordered by input index -- exactly the
// order of the reference's serial store loop.  The scatter uses atomics and
// the per-cell sort restores that order deterministically.
#include <algorithm>

#include "common.h"
#include "ghosts.h"
#include "kernels.h"
#include "subgrid.h"

namespace pg {
namespace {

constexpr int TB = 256;

template <int DIM>
__global__ void __launch_bounds__(TB) k_ghost_count(const double* pts, int64_t n, double buff,
                                                    int32_t* cnt) {
    int64_t i = (int64_t)blockIdx.x * TB + threadIdx.x;
    if (i >= n) return;
    double p[DIM];
#pragma unroll
    for (int d = 0; d < DIM; d++) p[d] = pts[DIM * i + d];
    int c = 0;
    constexpr int NR = DIM == 2 ? 8 : 26;
    for (int r = 0; r < NR; r++) c += in_region<DIM>(p, DIM == 2 ? REG2[r] : REG3[r], buff) ? 1 : 0;
    cnt[i] = c;
}

template <int DIM>
__global__ void __launch_bounds__(TB) k_ghost_write(const double* pts, int64_t n, double buff,
                                                    const int32_t* off, double* out,
                                                    int64_t* gsrc) {
    int64_t i = (int64_t)blockIdx.x * TB + threadIdx.x;
    if (i >= n) return;
    const double scale = 1. / (1. + (2 * buff));
    double p[DIM];
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        p[d] = pts[DIM * i + d];
        out[DIM * i + d] = scale * (p[d] - 0.5) + 0.5;
    }
    int64_t g = off ? off[i] : 0;
    constexpr int NR = DIM == 2 ? 8 : 26;
    for (int r = 0; r < NR; r++) {
        const uint8_t* code = DIM == 2 ? REG2[r] : REG3[r];
        if (!in_region<DIM>(p, code, buff)) continue;
#pragma unroll
        for (int d = 0; d < DIM; d++) {
            double sh = code[d] == 1 ? 1. : (code[d] == 2 ? -1. : 0.);
            double q = p[d] + sh;
            out[DIM * (n + g) + d] = scale * (q - 0.5) + 0.5;
        }
        gsrc[g] = i;
        g++;
    }
}

template <int DIM>
__global__ void __launch_bounds__(TB) k_cell_count(const double* pts, int64_t n, int N,
                                                   int32_t* cell, int32_t* counts) {
    int64_t i = (int64_t)blockIdx.x * TB + threadIdx.x;
    if (i >= n) return;
    double p[DIM];
#pragma unroll
    for (int d = 0; d < DIM; d++) p[d] = pts[DIM * i + d];
    int c = grid_cell<DIM>(p, N);
    cell[i] = c;
    atomicAdd(&counts[c], 1);
}

__global__ void __launch_bounds__(TB) k_cell_scatter(int64_t n, const int32_t* cell,
                                                     const int32_t* starts, int32_t* fill,
                                                     uint32_t* idx) {
    int64_t i = (int64_t)blockIdx.x * TB + threadIdx.x;
    if (i >= n) return;
    int c = cell[i];
    int pos = starts[c] + atomicAdd(&fill[c], 1);
    idx[pos] = (uint32_t)i;
}

// Sort each cell's index range ascending.  Uniform inputs hold ~3 points
// per cell: ranges up to 32 are insertion-sorted in registers by one
// thread; longer ranges (clustered inputs) are listed and sorted by a whole
// block (k_cell_sort_big).
__global__ void __launch_bounds__(TB) k_cell_sort(int64_t ncell, const int32_t* starts,
                                                  const int32_t* counts, uint32_t* idx, int32_t* big,
                                                  int32_t* nbig) {
    int64_t c = (int64_t)blockIdx.x * TB + threadIdx.x;
    if (c >= ncell) return;
    int len = counts[c];
    if (len < 2) return;
    if (len > 32) {
        big[atomicAdd(nbig, 1)] = (int32_t)c;
        return;
    }
    uint32_t* a = idx + starts[c];
    uint32_t v[32];
    for (int i = 0; i < len; i++) v[i] = a[i];
    for (int i = 1; i < len; i++) {
        uint32_t x = v[i];
        int j = i - 1;
        while (j >= 0 && v[j] > x) {
            v[j + 1] = v[j];
            j--;
        }
        v[j + 1] = x;
    }
    for (int i = 0; i < len; i++) a[i] = v[i];
}

// One block per listed cell (grid-stride over the list): bitonic sort of
// the range in place, in the form whose comparators all point the same way
// (the first step of every merge compares mirror positions), so a range of
// any length sorts as if padded with +inf: comparators whose partner lies
// beyond the range are skipped.  The values are distinct point indices.
__global__ void __launch_bounds__(TB) k_cell_sort_big(const int32_t* __restrict__ big, const int32_t* __restrict__ nbig,
                                                      const int32_t* __restrict__ starts,
                                                      const int32_t* __restrict__ counts, uint32_t* idx) {
    const int nb = *nbig;
    for (int bi = blockIdx.x; bi < nb; bi += gridDim.x) {
        const int c = big[bi];
        const int64_t len = counts[c];
        uint32_t* a = idx + starts[c];
        int64_t P = 1;
        while (P < len) P <<= 1;
        for (int64_t k = 2; k <= P; k <<= 1) {
            for (int64_t j = k >> 1; j > 0; j >>= 1) {
                for (int64_t t = threadIdx.x; t < P / 2; t += TB) {
                    // comparator t of this substep: position i in the lower half
                    // of its group of 2j, partner l
                    const int64_t i = (t / j) * 2 * j + (t % j);
                    const int64_t l = (j == (k >> 1)) ? (i / k) * k + (k - 1 - (i % k)) : i + j;
                    if (l < len && i < len) {
                        const uint32_t x = a[i], y = a[l];
                        if (x > y) {
                            a[i] = y;
                            a[l] = x;
                        }
                    }
                }
                __syncthreads();
            }
        }
    }
}

template <int DIM>
__global__ void __launch_bounds__(TB) k_gather_points(int64_t n, const uint32_t* perm,
                                                      const double* pts, double* sorted) {
    int64_t s = (int64_t)blockIdx.x * TB + threadIdx.x;
    if (s >= n) return;
    uint32_t o = perm[s];
#pragma unroll
    for (int d = 0; d < DIM; d++) sorted[DIM * s + d] = pts[DIM * (int64_t)o + d];
}

// Deterministic displacement (the degeneracy fallback of a mesh build):
// coordinate d of point i moves by amp * (2u - 1), u = the top 53 bits of
// SplitMix64 at counter salt + (i * DIM + d + 1) * 0x9E3779B97F4A7C15, over
// 2^53.  A coordinate that leaves [0, 1) wraps (periodic) or is clamped.
template <int DIM>
__global__ void __launch_bounds__(TB) k_displace(const double* in, int64_t n, double amp, uint64_t salt,
                                                 int periodic, double* out) {
    int64_t i = (int64_t)blockIdx.x * TB + threadIdx.x;
    if (i >= n) return;
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        uint64_t z = salt + (uint64_t)(i * DIM + d + 1) * 0x9E3779B97F4A7C15ull;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        const double u = (double)(z >> 11) * 0x1p-53;
        const double t = 2.0 * u - 1.0;
        double x = in[DIM * i + d] + amp * t;
        if (periodic) {
            if (x < 0.0) x += 1.0;
            if (x >= 1.0) x -= 1.0;
        } else {
            x = fmin(fmax(x, 0.0), 0x1.fffffffffffffp-1);
        }
        out[DIM * i + d] = x;
    }
}

inline unsigned nblk(int64_t n) { return (unsigned)((n + TB - 1) / TB); }

__global__ void k_fill_one(int32_t* p) {
    if (threadIdx.x == 0) *p = 0;
}

__global__ void __launch_bounds__(TB) k_fill_i32(int32_t* a, int64_t n, int32_t v) {
    for (int64_t i = (int64_t)blockIdx.x * TB + threadIdx.x; i < n; i += (int64_t)gridDim.x * TB) a[i] = v;
}

// One block per listed cell (grid-stride over the list of cells longer than
// 32): the cell's sub-cell histogram in LDS, its exclusive scan (each thread
// a contiguous run of sub-cells, then the runs' sums in order), the table
// of (first position, count) per sub-cell from a range taken off *alloc,
// and the scatter of the cell's sorted ids by sub-cell into idx.  The order
// inside a sub-cell follows the LDS atomics (searches do not depend on it).
// A cell whose table would not fit in cap keeps m = 0 (its searches scan
// the whole cell).
template <int DIM>
__global__ void __launch_bounds__(TB) k_sub_build(const int32_t* __restrict__ big, const int32_t* __restrict__ nbig,
                                                  const int32_t* __restrict__ starts,
                                                  const int32_t* __restrict__ counts,
                                                  const double* __restrict__ sorted, int N, int32_t* __restrict__ sm,
                                                  int32_t* __restrict__ sptr, int32_t* __restrict__ sstart,
                                                  int32_t* __restrict__ scnt, int64_t cap, uint32_t* __restrict__ idx,
                                                  int32_t* alloc) {
    __shared__ int32_t hist[4096];
    __shared__ int32_t part[TB];
    __shared__ int32_t base_s;
    const int nb = *nbig;
    const int tid = threadIdx.x;
    for (int bi = blockIdx.x; bi < nb; bi += gridDim.x) {
        const int c = big[bi];
        const int64_t st = starts[c];
        const int cnt = counts[c];
        const int m = sub_split<DIM>(cnt);
        const int nsub = DIM == 3 ? m * m * m : m * m;
        const int ix[3] = {c % N, (c / N) % N, DIM == 3 ? c / (N * N) : 0};
        for (int j = tid; j < nsub; j += TB) hist[j] = 0;
        __syncthreads();
        for (int i = tid; i < cnt; i += TB)
            atomicAdd(&hist[sub_index<DIM>(sorted + (size_t)DIM * (st + i), ix, m, N)], 1);
        __syncthreads();
        const int per = (nsub + TB - 1) / TB;
        const int q0 = min(nsub, tid * per), q1 = min(nsub, q0 + per);
        int lsum = 0;
        for (int q = q0; q < q1; q++) lsum += hist[q];
        part[tid] = lsum;
        __syncthreads();
        if (tid == 0) {
            int run = 0;
            for (int t = 0; t < TB; t++) {
                const int v = part[t];
                part[t] = run;
                run += v;
            }
            const int b = atomicAdd(alloc, nsub);
            base_s = ((int64_t)b + nsub <= cap) ? b : -1;
        }
        __syncthreads();
        const int base = base_s;
        if (base >= 0) {
            int run = part[tid];
            for (int q = q0; q < q1; q++) {
                const int h = hist[q];
                sstart[base + q] = (int32_t)(st + run);
                scnt[base + q] = h;
                hist[q] = run;
                run += h;
            }
            __syncthreads();
            for (int i = tid; i < cnt; i += TB) {
                const int64_t s = st + i;
                const int j = sub_index<DIM>(sorted + (size_t)DIM * s, ix, m, N);
                const int pos = atomicAdd(&hist[j], 1);
                idx[st + pos] = (uint32_t)s;
            }
            if (tid == 0) {
                sm[c] = m;
                sptr[c] = base;
            }
        }
        __syncthreads();
    }
}

}  // namespace

void launch_ghost_count(int dim, const double* pts, int64_t n, double buff, int32_t* cnt,
                        hipStream_t s) {
    if (n <= 0) return;
    if (dim == 2)
        k_ghost_count<2><<<nblk(n), TB, 0, s>>>(pts, n, buff, cnt);
    else
        k_ghost_count<3><<<nblk(n), TB, 0, s>>>(pts, n, buff, cnt);
}

void launch_ghost_write(int dim, const double* pts, int64_t n, double buff, const int32_t* off,
                        double* out, int64_t* gsrc, hipStream_t s) {
    if (n <= 0) return;
    if (dim == 2)
        k_ghost_write<2><<<nblk(n), TB, 0, s>>>(pts, n, buff, off, out, gsrc);
    else
        k_ghost_write<3><<<nblk(n), TB, 0, s>>>(pts, n, buff, off, out, gsrc);
}

void launch_cell_count(int dim, const double* pts, int64_t n, int N, int32_t* cell,
                       int32_t* counts, hipStream_t s) {
    if (n <= 0) return;
    if (dim == 2)
        k_cell_count<2><<<nblk(n), TB, 0, s>>>(pts, n, N, cell, counts);
    else
        k_cell_count<3><<<nblk(n), TB, 0, s>>>(pts, n, N, cell, counts);
}

void launch_cell_scatter(int64_t n, const int32_t* cell, const int32_t* starts, int32_t* fill,
                         uint32_t* idx, hipStream_t s) {
    if (n <= 0) return;
    k_cell_scatter<<<nblk(n), TB, 0, s>>>(n, cell, starts, fill, idx);
}

void launch_cell_sort(int64_t ncell, const int32_t* starts, const int32_t* counts, uint32_t* idx,
                      int32_t* big, int32_t* nbig, hipStream_t s) {
    if (ncell <= 0) return;
    k_fill_one<<<1, 64, 0, s>>>(nbig);
    k_cell_sort<<<nblk(ncell), TB, 0, s>>>(ncell, starts, counts, idx, big, nbig);
    k_cell_sort_big<<<(unsigned)std::min<int64_t>(ncell, 1024), TB, 0, s>>>(big, nbig, starts, counts, idx);
}

void launch_subgrid(int dim, const double* sorted, int64_t n, int N, int64_t ncell, const int32_t* starts,
                    const int32_t* counts, const int32_t* big, const int32_t* nbig, int32_t* m, int32_t* ptr,
                    int32_t* start, int32_t* cnt, int64_t cap, uint32_t* idx, int32_t* alloc, hipStream_t s) {
    if (ncell <= 0 || n <= 0) return;
    k_fill_i32<<<(unsigned)std::min<int64_t>(nblk(ncell), 4096), TB, 0, s>>>(m, ncell, 0);
    k_fill_one<<<1, 64, 0, s>>>(alloc);
    const unsigned nb = (unsigned)std::min<int64_t>(ncell, 1024);
    if (dim == 2)
        k_sub_build<2><<<nb, TB, 0, s>>>(big, nbig, starts, counts, sorted, N, m, ptr, start, cnt, cap, idx, alloc);
    else
        k_sub_build<3><<<nb, TB, 0, s>>>(big, nbig, starts, counts, sorted, N, m, ptr, start, cnt, cap, idx, alloc);
}

void launch_displace(int dim, const double* in, int64_t n, double amp, uint64_t salt, int periodic, double* out,
                     hipStream_t s) {
    if (n <= 0) return;
    if (dim == 2)
        k_displace<2><<<nblk(n), TB, 0, s>>>(in, n, amp, salt, periodic, out);
    else
        k_displace<3><<<nblk(n), TB, 0, s>>>(in, n, amp, salt, periodic, out);
}

void launch_gather_points(int dim, int64_t n, const uint32_t* perm, const double* pts,
                          double* sorted, hipStream_t s) {
    if (n <= 0) return;
    if (dim == 2)
        k_gather_points<2><<<nblk(n), TB, 0, s>>>(n, perm, pts, sorted);
    else
        k_gather_points<3><<<nblk(n), TB, 0, s>>>(n, perm, pts, sorted);
}

}  // namespace pg
