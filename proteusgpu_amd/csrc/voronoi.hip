// Kernel entry points for the Voronoi clip and the restart bookkeeping.
#include <atomic>
#include <cstdio>
#include <cstdlib>

#include "voronoi_kernel.h"

namespace pg {
namespace {

// Fast kernel: 16 lanes x 4 slots = 64 vertex slots per seed, 4 seeds per
// wave.  Large kernel: 16 lanes x 8 slots = 128 slots, for seeds whose cell
// outgrows 64 vertices before max_t (rare).
template <int DIM, int G, int S>
__global__ void __launch_bounds__(64, (VoroCfg<DIM, G, S>::WAVES)) k_voronoi(const double* __restrict__ pts, int64_t n_total,
                                                const uint32_t* __restrict__ seeds, int64_t nq,
                                                const uint32_t* __restrict__ knn, int k, int max_p,
                                                int max_t, FaceArrays fa) {
    using C = VoroCfg<DIM, G, S>;
    __shared__ SeedSmem<C::NS, C::MAXP> sm[C::SPW];
    voronoi_group<DIM, G, S>(pts, n_total, seeds, nq, knn, k, max_p, max_t, fa, sm,
                             xcd_block(blockIdx.x, gridDim.x));
}

// The same kernel under its own name for certificate restarts, so that
// profiles separate the first pass from the restarts.
template <int DIM, int G, int S>
__global__ void __launch_bounds__(64, (VoroCfg<DIM, G, S>::WAVES)) k_voronoi_restart(const double* __restrict__ pts, int64_t n_total,
                                                        const uint32_t* __restrict__ seeds, int64_t nq,
                                                        const uint32_t* __restrict__ knn, int k, int max_p,
                                                        int max_t, FaceArrays fa) {
    using C = VoroCfg<DIM, G, S>;
    __shared__ SeedSmem<C::NS, C::MAXP> sm[C::SPW];
    voronoi_group<DIM, G, S>(pts, n_total, seeds, nq, knn, k, max_p, max_t, fa, sm,
                             xcd_block(blockIdx.x, gridDim.x));
}

// Cells whose filtered predicate was degenerate, again with the exact
// predicate (128 vertex slots).
template <int DIM>
__global__ void __launch_bounds__(64, 1) k_voronoi_exact(const double* __restrict__ pts, int64_t n_total,
                                                         const uint32_t* __restrict__ seeds, int64_t nq,
                                                         const uint32_t* __restrict__ knn, int k, int max_p,
                                                         int max_t, FaceArrays fa) {
    using C = VoroCfg<DIM, 16, 8>;
    __shared__ SeedSmem<C::NS, C::MAXP> sm[C::SPW];
    voronoi_group<DIM, 16, 8, true>(pts, n_total, seeds, nq, knn, k, max_p, max_t, fa, sm, blockIdx.x);
}

// Vertex-ball completion of seeds no row certifies (128 vertex slots).
template <int DIM>
__global__ void __launch_bounds__(64, 1) k_voronoi_ball(const double* __restrict__ pts, int64_t n_total,
                                                        const uint32_t* __restrict__ seeds, int64_t nq,
                                                        const uint32_t* __restrict__ knn, int k, int max_p,
                                                        int max_t, FaceArrays fa, BallGrid bg) {
    using C = VoroCfg<DIM, 16, 8>;
    __shared__ SeedSmem<C::NS, C::MAXP> sm[C::SPW];
    __shared__ int32_t scr[C::SPW * (2 * 16 + 1)];
    bg.scr = scr;
    voronoi_group<DIM, 16, 8, false, true>(pts, n_total, seeds, nq, knn, k, max_p, max_t, fa, sm, blockIdx.x, bg);
}

template <int DIM, int G, int S>
void run(const double* pts, int64_t n_total, const uint32_t* seeds, int64_t nq, const uint32_t* knn, int k,
         int max_p, int max_t, FaceArrays fa, bool restart, hipStream_t s) {
    using C = VoroCfg<DIM, G, S>;
    int64_t nb = (nq + C::SPW - 1) / C::SPW;
    if (restart)
        k_voronoi_restart<DIM, G, S><<<dim3((unsigned)nb), dim3(64), 0, s>>>(pts, n_total, seeds, nq, knn, k,
                                                                              max_p, max_t, fa);
    else
        k_voronoi<DIM, G, S><<<dim3((unsigned)nb), dim3(64), 0, s>>>(pts, n_total, seeds, nq, knn, k, max_p,
                                                                      max_t, fa);
}

constexpr int TB = 256;

__global__ void __launch_bounds__(TB) k_flag_status(int64_t nq, const uint32_t* seeds,
                                                    const int32_t* status, int code, int32_t* flags) {
    int64_t i = (int64_t)blockIdx.x * TB + threadIdx.x;
    if (i >= nq) return;
    uint32_t s = seeds ? seeds[i] : (uint32_t)i;
    int st = status[s];
    flags[i] = (code < 0) ? (st != SUCCESS) : (st == code);
}

__global__ void __launch_bounds__(TB) k_compact(int64_t nq, const uint32_t* seeds,
                                                const int32_t* flags, const int32_t* offs,
                                                uint32_t* list) {
    int64_t i = (int64_t)blockIdx.x * TB + threadIdx.x;
    if (i >= nq || !flags[i]) return;
    list[offs[i]] = seeds ? seeds[i] : (uint32_t)i;
}

}  // namespace

// Vertex slots of the fast kernel per seed: 16 lanes x 4 (64 slots) or
// 16 lanes x 3 (48 slots; less work per clip, more seeds overflow into the
// 128-slot kernel).  Chosen by the autotuner or PGPU_CLIP_SLOTS.
// Fast-kernel shape: lanes per seed x vertex slots per lane.  16 x 4 (64
// slots, 4 seeds per wave), 16 x 3 (48 slots; less work per clip, a few
// more seeds overflow into the 128-slot kernel), 8 x 6 and 8 x 8 (8 seeds
// per wave: the group-wide work of a clip is shared by twice as many seeds,
// at twice the LDS per wave).  Chosen by the autotuner or PGPU_CLIP_CFG
// ("16x4", "16x3", "8x6", "8x8").
static std::atomic<int> g_clip_lanes{16}, g_clip_slots{64};
static const bool g_clip_env = [] {
    const char* e = getenv("PGPU_CLIP_CFG");
    int l = 0, sl = 0;
    if (e && sscanf(e, "%dx%d", &l, &sl) == 2) set_clip_config(l, l * sl);
    const char* e2 = getenv("PGPU_CLIP_SLOTS");
    if (e2 && atoi(e2) == 48) set_clip_config(16, 48);
    return true;
}();
bool valid_clip_config(int lanes, int slots) {
    return (lanes == 16 || lanes == 8) && (slots == 48 || slots == 64);
}
void set_clip_config(int lanes, int slots) {
    if (!valid_clip_config(lanes, slots)) return;
    g_clip_lanes = lanes;
    g_clip_slots = slots;
}
int clip_lanes() { return g_clip_lanes; }
int clip_slots() { return g_clip_slots; }

// Certificate restarts resume the saved cell (default) or clip again from
// the first neighbour (PGPU_CLIP_RESUME=0); the results are the same.
static std::atomic<bool> g_clip_resume{[] {
    const char* e = getenv("PGPU_CLIP_RESUME");
    return !(e && atoi(e) == 0);
}()};
bool clip_resume() { return g_clip_resume; }
void set_clip_resume(bool on) { g_clip_resume = on; }

template <int DIM>
void launch_dim(const double* sorted, int64_t n, const uint32_t* seeds, int64_t nq, const uint32_t* knn, int k,
                int max_p, int max_t, int big, FaceArrays fa, hipStream_t s, bool restart, int lanes, int slots,
                bool exact) {
    if (exact) {
        using C = VoroCfg<DIM, 16, 8>;
        const int64_t nb = (nq + C::SPW - 1) / C::SPW;
        k_voronoi_exact<DIM><<<dim3((unsigned)nb), dim3(64), 0, s>>>(sorted, n, seeds, nq, knn, k, max_p, max_t,
                                                                      fa);
    } else if (big) run<DIM, 16, 8>(sorted, n, seeds, nq, knn, k, max_p, max_t, fa, true, s);
    else if (lanes == 8 && slots == 48) run<DIM, 8, 6>(sorted, n, seeds, nq, knn, k, max_p, max_t, fa, restart, s);
    else if (lanes == 8) run<DIM, 8, 8>(sorted, n, seeds, nq, knn, k, max_p, max_t, fa, restart, s);
    else if (slots == 48) run<DIM, 16, 3>(sorted, n, seeds, nq, knn, k, max_p, max_t, fa, restart, s);
    else run<DIM, 16, 4>(sorted, n, seeds, nq, knn, k, max_p, max_t, fa, restart, s);
}

void launch_voronoi(int dim, const double* sorted, int64_t n, const uint32_t* seeds, int64_t nq,
                    const uint32_t* knn, int k, int max_p, int max_t, int big, FaceArrays fa, int lanes, int slots,
                    hipStream_t s, bool restart, bool exact) {
    if (nq <= 0) return;
    (void)g_clip_env;
    if (dim == 2)
        launch_dim<2>(sorted, n, seeds, nq, knn, k, max_p, max_t, big, fa, s, restart, lanes, slots, exact);
    else
        launch_dim<3>(sorted, n, seeds, nq, knn, k, max_p, max_t, big, fa, s, restart, lanes, slots, exact);
}

void launch_voronoi_ball(int dim, const double* sorted, int64_t n, const uint32_t* seeds, int64_t nq,
                         const uint32_t* knn, int k, int max_p, int max_t, FaceArrays fa, int N,
                         const int32_t* starts, const int32_t* counts, const uint32_t* perm, int kmark,
                         hipStream_t s, SubGrid sg) {
    if (nq <= 0) return;
    BallGrid bg;
    bg.N = N;
    bg.starts = starts;
    bg.counts = counts;
    bg.perm = perm;
    bg.kmark = kmark;
    bg.sg = sg;
    const int64_t nb = (nq + 3) / 4;  // 4 seeds per wave
    if (dim == 2)
        k_voronoi_ball<2><<<dim3((unsigned)nb), dim3(64), 0, s>>>(sorted, n, seeds, nq, knn, k, max_p, max_t, fa, bg);
    else
        k_voronoi_ball<3><<<dim3((unsigned)nb), dim3(64), 0, s>>>(sorted, n, seeds, nq, knn, k, max_p, max_t, fa, bg);
}

void launch_collect_status(int64_t nq, const uint32_t* seeds, const int32_t* status, int code,
                           uint32_t* list, int32_t* flags, int32_t* offs, int32_t* tmp,
                           int32_t* nout, hipStream_t s) {
    if (nq <= 0) return;
    unsigned nb = (unsigned)((nq + TB - 1) / TB);
    k_flag_status<<<nb, TB, 0, s>>>(nq, seeds, status, code, flags);
    scan_exclusive_i32(flags, offs, nq, tmp, nout, s);
    k_compact<<<nb, TB, 0, s>>>(nq, seeds, flags, offs, list);
}

void launch_collect_failed(int64_t n, const int32_t* status, uint32_t* list, int32_t* flags,
                           int32_t* offs, int32_t* tmp, int32_t* nfail, hipStream_t s) {
    launch_collect_status(n, nullptr, status, -1, list, flags, offs, tmp, nfail, s);
}

}  // namespace pg
