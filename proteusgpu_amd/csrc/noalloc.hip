// Allocation-free KNN build: the grid counting sort and the k-nearest rows
// of pg_knn_init + pg_knn_solve with every scratch buffer carved from one
// static device array, for callers that cannot allocate device memory (code
// under HIP graph capture, tight serving loops).  The kernels are the ones
// the allocating entry points launch; only where their scratch lives
// differs.
#include <algorithm>
#include <mutex>

#include "common.h"
#include "kernels.h"

namespace pg {

namespace {

// One copy per GPU, held until the process ends.  In host code this name is
// a host-side stand-in: every host use goes through the device address that
// hipGetSymbolAddress returns (kept per device in g_scratch_ptr).
__device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
void* g_scratch_ptr[MAX_DEVICES] = {};
std::mutex g_scratch_mu;

inline int64_t round256(int64_t b) { return (b + 255) & ~(int64_t)255; }

}  // namespace

hipError_t noalloc_setup(int device, void** ptr) {
    if (device < 0 || device >= MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        // the device must be current: each GPU holds the array at its own address
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[device] = p;
    }
    *ptr = g_scratch_ptr[device];
    return hipSuccess;
}

int64_t noalloc_knn_bytes(int dim, int64_t n, int N) {
    const int64_t ncell = (int64_t)N * N * (dim == 3 ? N : 1);
    return round256(4 * n)                                          // cell of every point
           + 3 * round256(4 * ncell)                                // counts, starts, fill
           + round256(4 * (ncell + 1))                              // long-cell list + its counter
           + round256(4 * scan_tmp_elems(std::max(ncell, n)))       // scan partials
           + round256(4);                                           // scan total
}

void noalloc_knn(void* base, int dim, const double* pts, int64_t n, int N, int k, double* sorted, uint32_t* perm,
                 uint32_t* rows, hipStream_t s) {
    const int64_t ncell = (int64_t)N * N * (dim == 3 ? N : 1);
    char* p = (char*)base;
    auto carve = [&](int64_t elems) {
        int32_t* a = (int32_t*)p;
        p += round256(4 * elems);
        return a;
    };
    int32_t* cell = carve(n);
    int32_t* counts = carve(ncell);
    int32_t* starts = carve(ncell);
    int32_t* fill = carve(ncell);
    int32_t* big = carve(ncell + 1);
    int32_t* tmp = carve(scan_tmp_elems(std::max(ncell, n)));
    int32_t* total = carve(1);
    // the array keeps the last call's contents: clear the counters by kernel
    // (no memset node, so the sequence can be captured into a graph)
    launch_fill_i32(counts, ncell, 0, s);
    launch_fill_i32(fill, ncell, 0, s);
    launch_cell_count(dim, pts, n, N, cell, counts, s);
    scan_exclusive_i32(counts, starts, ncell, tmp, total, s);
    launch_cell_scatter(n, cell, starts, fill, perm, s);
    launch_cell_sort(ncell, starts, counts, perm, big + 1, big, s);
    launch_gather_points(dim, n, perm, pts, sorted, s);
    launch_knn(dim, sorted, perm, n, N, starts, counts, nullptr, n, k, rows, s);
}

}  // namespace pg
