// Batched boxes: many independent boxes advanced by one captured step.
//
// The boxes share every array, box-major: hydro cells [hoff[b], hoff[b+1]),
// seeds (hydro then ghosts, the single-box order) [soff[b], soff[b+1]),
// grid cells [coff[b], coff[b+1]).  Seed and cell counts change from step to
// step with the ghosts, so they live in device memory and every kernel below
// reads them there; launch shapes depend only on capacities, which makes the
// whole step capturable in one hipGraph.  Each box sees exactly the
// arithmetic of a single-box run (its own grid resolution, neighbour rows,
// clip, restart schedule, CFL step), so its results are bit-identical to
// running it alone.  No counterpart in the reference (one box per run).
#include <algorithm>
#include <type_traits>

#include "common.h"
#include "ghosts.h"
#include "hydro_core.h"
#include "kernels.h"
#include "knn_core.h"
#include "voronoi_kernel.h"

namespace pg {
namespace {

constexpr int TB = 256;
inline unsigned nblk(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + TB - 1) / TB); }
inline unsigned nblk_dt(int64_t n) {
    return (unsigned)std::max<int64_t>(1, (n + (int64_t)TB * DT_CELLS - 1) / ((int64_t)TB * DT_CELLS));
}

// ------------------------------------------------------------------ setup
// Per box: ghosts, seed count, grid resolution (knn.cpp:22, the same
// arithmetic as the host's ngrid), cell count; then offsets.  One block.
__global__ void __launch_bounds__(TB)
    k_box_setup(int nbox, int dim, const int64_t* __restrict__ hoff, const int32_t* __restrict__ goff,
                const int32_t* __restrict__ gtot, int64_t seed_cap, int64_t cell_cap, int64_t* soff,
                int32_t* N, int64_t* coff, int32_t* err) {
    __shared__ int64_t ns[4096], nc[4096];
    for (int b = threadIdx.x; b < nbox; b += TB) {
        int64_t g = 0;
        if (goff) {
            const int64_t h0 = hoff[b], h1 = hoff[b + 1];
            const int64_t g0 = goff[h0];
            const int64_t g1 = b + 1 < nbox ? (int64_t)goff[h1] : (int64_t)*gtot;
            g = g1 - g0;
        }
        const int64_t n = hoff[b + 1] - hoff[b] + g;
        const int Nb = ngrid_of(dim, n);
        N[b] = Nb;
        ns[b] = n;
        nc[b] = dim == 2 ? (int64_t)Nb * Nb : (int64_t)Nb * Nb * Nb;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t s = 0, c = 0;
        for (int b = 0; b < nbox; b++) {
            soff[b] = s;
            coff[b] = c;
            s += ns[b];
            c += nc[b];
        }
        soff[nbox] = s;
        coff[nbox] = c;
        if (s > seed_cap || c > cell_cap) {
            // over capacity: report, and leave nothing for the kernels after
            atomicOr(err, s > seed_cap ? 1 : 2);
            soff[nbox] = 0;
            coff[nbox] = 0;
        }
    }
}

// Seeds of every box in the single-box order: the box's hydro points
// (scaled for periodic boxes), then its ghosts point-major in region order.
// orig_h[seed] = hydro id (>= 0) or -1 - source hydro id for a ghost.
template <int DIM>
__global__ void __launch_bounds__(TB)
    k_bseeds(const double* __restrict__ pts, int64_t n_h, double buff, int periodic,
             const int32_t* __restrict__ hbox, const int64_t* __restrict__ hoff,
             const int32_t* __restrict__ goff, const int64_t* __restrict__ soff, const int32_t* err,
             double* __restrict__ out, int32_t* __restrict__ sbox, int32_t* __restrict__ orig_h) {
    const int64_t i = (int64_t)blockIdx.x * TB + threadIdx.x;
    if (i >= n_h || *err) return;
    const int b = hbox[i];
    const int64_t nb = hoff[b + 1] - hoff[b];
    const int64_t base = soff[b];
    const int64_t l = i - hoff[b];
    double p[DIM];
#pragma unroll
    for (int d = 0; d < DIM; d++) p[d] = pts[(size_t)DIM * i + d];
    if (!periodic) {
#pragma unroll
        for (int d = 0; d < DIM; d++) out[(size_t)DIM * (base + l) + d] = p[d];
        sbox[base + l] = b;
        orig_h[base + l] = (int32_t)i;
        return;
    }
    const double scale = 1. / (1. + (2 * buff));
#pragma unroll
    for (int d = 0; d < DIM; d++) out[(size_t)DIM * (base + l) + d] = scale * (p[d] - 0.5) + 0.5;
    sbox[base + l] = b;
    orig_h[base + l] = (int32_t)i;
    int64_t g = base + nb + (goff[i] - goff[hoff[b]]);
    constexpr int NR = DIM == 2 ? 8 : 26;
    for (int r = 0; r < NR; r++) {
        const uint8_t* code = DIM == 2 ? REG2[r] : REG3[r];
        if (!in_region<DIM>(p, code, buff)) continue;
#pragma unroll
        for (int d = 0; d < DIM; d++) {
            const double sh = code[d] == 1 ? 1. : (code[d] == 2 ? -1. : 0.);
            out[(size_t)DIM * g + d] = scale * ((p[d] + sh) - 0.5) + 0.5;
        }
        sbox[g] = b;
        orig_h[g] = -1 - (int32_t)i;
        g++;
    }
}

template <int DIM>
__global__ void __launch_bounds__(TB)
    k_bcell_count(const double* __restrict__ pts, const int64_t* __restrict__ n_dev, int64_t cap,
                  const int32_t* __restrict__ sbox, const int32_t* __restrict__ N, const int64_t* __restrict__ coff,
                  int32_t* __restrict__ cell, int32_t* __restrict__ counts) {
    const int64_t n = *n_dev;
    for (int64_t i = (int64_t)blockIdx.x * TB + threadIdx.x; i < n && i < cap; i += (int64_t)gridDim.x * TB) {
        double p[DIM];
#pragma unroll
        for (int d = 0; d < DIM; d++) p[d] = pts[(size_t)DIM * i + d];
        const int b = sbox[i];
        const int64_t c = coff[b] + grid_cell<DIM>(p, N[b]);
        cell[i] = (int32_t)c;
        atomicAdd(&counts[c], 1);
    }
}

__global__ void __launch_bounds__(TB)
    k_bcell_scatter(const int64_t* __restrict__ n_dev, int64_t cap, const int32_t* __restrict__ cell,
                    const int32_t* __restrict__ starts, int32_t* fill, uint32_t* idx) {
    const int64_t n = *n_dev;
    for (int64_t i = (int64_t)blockIdx.x * TB + threadIdx.x; i < n && i < cap; i += (int64_t)gridDim.x * TB) {
        const int c = cell[i];
        idx[starts[c] + atomicAdd(&fill[c], 1)] = (uint32_t)i;
    }
}

template <int DIM>
__global__ void __launch_bounds__(TB)
    k_bgather(const int64_t* __restrict__ n_dev, int64_t cap, const uint32_t* __restrict__ perm,
              const double* __restrict__ pts, const int32_t* __restrict__ sbox, double* __restrict__ sorted,
              int32_t* __restrict__ sorted_box) {
    const int64_t n = *n_dev;
    for (int64_t s = (int64_t)blockIdx.x * TB + threadIdx.x; s < n && s < cap; s += (int64_t)gridDim.x * TB) {
        const uint32_t o = perm[s];
#pragma unroll
        for (int d = 0; d < DIM; d++) sorted[(size_t)DIM * s + d] = pts[(size_t)DIM * o + d];
        sorted_box[s] = sbox[o];
    }
}

template <class T>
__global__ void __launch_bounds__(TB) k_fill(T* a, int64_t n, T v) {
    for (int64_t i = (int64_t)blockIdx.x * TB + threadIdx.x; i < n; i += (int64_t)gridDim.x * TB) a[i] = v;
}

// Work queue of the one-wave persistent kernels: a block's first chunk of
// items is its own (chunk blockIdx.x), the later ones come from a counter
// (cleared before the launch) counting past the grid's first chunks, so
// seeds of uneven cost balance over the resident blocks.  Atomics on the
// one counter serialize (about 10 ns each): the first chunks take none (a
// block of an empty or short list leaves without one), and chunks keep the
// rest few.  `first` starts true in every block.
PG_DEV int64_t next_ticket(unsigned long long* ticket, int chunk, bool& first) {
    if (first) {
        first = false;
        return (int64_t)blockIdx.x * chunk;
    }
    unsigned long long t = 0;
    if (threadIdx.x == 0) t = atomicAdd(ticket, (unsigned long long)chunk);
    // lane 0's value as a scalar: the compiler then knows the whole wave
    // works on the same items (scalar loads, no per-lane copies of them)
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)t);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(t >> 32));
    return (int64_t)gridDim.x * chunk + (int64_t)((uint64_t)hi << 32 | lo);
}
// Items per ticket: up to `most`, fewer for short lists (each block should
// see several chunks, so that the last ones balance).
PG_DEV int ticket_chunk(int64_t items, int most) {
    const int64_t c = items / (4 * (int64_t)gridDim.x);
    return (int)(c < 1 ? 1 : (c > most ? most : c));
}
constexpr int KNN_CHUNK = 16;  // queries per ticket, at most
constexpr int CLIP_CHUNK = 2;  // seed groups per ticket, at most

// Resident one-wave blocks of a kernel on the device (its occupancy per CU,
// capped at blocks_per_cu, times the CUs).
int device_cus() {
    static const int ncu = [] {
        int dev = 0, n = 0;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev);
        return n > 0 ? n : 256;
    }();
    return ncu;
}
template <class K>
int occupancy(K kernel) {
    int occ = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kernel, 64, 0) != hipSuccess || occ <= 0) occ = 1;
    return occ;
}
unsigned resident_blocks(int occ, int blocks_per_cu) {
    return (unsigned)(device_cus() * std::max(1, std::min(occ, blocks_per_cu)));
}

// ------------------------------------------------------------------ KNN
template <int DIM, int BLOCK>
__global__ void __launch_bounds__(BLOCK)
    k_bknn(const double* __restrict__ sorted, const uint32_t* __restrict__ perm, const int32_t* __restrict__ sbox,
           const int32_t* __restrict__ N, const int64_t* __restrict__ coff, const int32_t* __restrict__ starts,
           const int32_t* __restrict__ counts, const uint32_t* __restrict__ queries, const int64_t* __restrict__ nq_dev,
           int k, uint32_t* __restrict__ out) {
    extern __shared__ unsigned char smem_raw[];
    double* hd = (double*)smem_raw;
    uint32_t* hi = (uint32_t*)(hd + (size_t)k * BLOCK);
    const int64_t nq = *nq_dev;
    for (int64_t qi = (int64_t)blockIdx.x * BLOCK + threadIdx.x; qi < nq; qi += (int64_t)gridDim.x * BLOCK) {
        const uint32_t q = queries ? queries[qi] : (uint32_t)qi;
        const int b = sbox[q];
        knn_heap_query<DIM, BLOCK>(sorted, perm, N[b], coff[b], starts, counts, q, k, out + (size_t)qi * k, hd, hi);
    }
}

// The persistent KNN kernel below takes one struct of arguments and, for every
// item, read it again from the kernel-argument segment through a pointer
// the compiler must treat as new each time: loop-invariant arguments would
// otherwise be hoisted out of the item loop and stay live across it, and
// the item's own registers would spill.  The struct is the kernel's only
// explicit argument, so it starts the segment.
template <class A>
PG_DEV const __attribute__((address_space(4))) A* fresh_args() {
    const __attribute__((address_space(4))) A* p =
        (const __attribute__((address_space(4))) A*)__builtin_amdgcn_kernarg_segment_ptr();
    asm volatile("" : "+s"(p));
    return p;
}

struct BKnnArgs {
    const double* sorted;
    const uint32_t* perm;
    const int32_t* sbox;
    const int32_t* N;
    const int64_t* coff;
    const int32_t* starts;
    const int32_t* counts;
    const uint32_t* queries;
    const int64_t* nq_dev;
    uint32_t* out;
    unsigned long long* ticket;
    KnnTail tail;
    int k;
};

template <int DIM, int M>
__global__ void __launch_bounds__(64, (M <= 2 ? 6 : 5)) k_bknn_wave(BKnnArgs args) {
    __shared__ int32_t c_start[64], c_pfx[65];
    __shared__ KnnPend pend;
    const int64_t nq = *args.nq_dev;
    const int chunk = ticket_chunk(nq, KNN_CHUNK);
    bool first = true;
    for (;;) {
        const int64_t q0 = next_ticket(fresh_args<BKnnArgs>()->ticket, chunk, first);
        if (q0 >= nq) break;
        const int64_t q1 = q0 + chunk < nq ? q0 + chunk : nq;
        for (int64_t qi = q0; qi < q1; qi++) {
            const auto* A = fresh_args<BKnnArgs>();
            const uint32_t q = A->queries ? A->queries[qi] : (uint32_t)qi;
            const int b = A->sbox[q];
            const int k = A->k;
            KnnTail tail;
            tail.resume = A->tail.resume;
            tail.lastnb = A->tail.lastnb;
            tail.tag = A->tail.tag;
            int kp = 0;
            uint32_t lo = 0;
            tail_of(tail, q, k, kp, lo);
            knn_wave_query<DIM, M>(A->sorted, A->perm, A->N[b], A->coff[b], A->starts, A->counts, q, k,
                                   A->out + (size_t)qi * k, c_start, c_pfx, &pend, kp, lo);
            __syncthreads();
        }
    }
}

// ------------------------------------------------------------------ clip
// Two forms.  k_bvoronoi: one block per group of SPW listed seeds, over the
// list's capacity; blocks past the device-side count end at once (each
// costs a dispatch, so it suits lists that fill most of their capacity).
// k_bvoronoi_loop: persistent blocks taking groups from a counter, the
// arguments re-read per group (in a loop the compiler still keeps values
// across groups, and the clip's registers spill: it suits the short lists
// of the certificate restarts, where capacity-sized grids would be mostly
// empty blocks).
struct BClipArgs {
    const double* pts;
    const uint32_t* seeds;
    const int64_t* nq_dev;
    const uint32_t* knn;
    unsigned long long* ticket;
    FaceArrays fa;
    int k, max_p, max_t;
};

// a FaceArrays out of the argument segment, field by field
PG_DEV FaceArrays face_arrays_of(const __attribute__((address_space(4))) FaceArrays& a) {
    FaceArrays f;
    f.stride = a.stride;
    f.nbr = a.nbr;
    f.area = a.area;
    f.nrm = a.nrm;
    f.cen = a.cen;
    f.count = a.count;
    f.vol = a.vol;
    f.ccen = a.ccen;
    f.status = a.status;
    f.kcert = a.kcert;
    f.cstate = a.cstate;
    f.resume = a.resume;
    f.lastnb = a.lastnb;
    f.key = a.key;
    f.nbr_cm = a.nbr_cm;
    return f;
}

// FRESH: the arguments are re-read per group (3D: fewer spills); else they
// stay in registers across groups (2D, whose clip has room for them).
template <int DIM, int G, int S, bool FRESH>
__global__ void __launch_bounds__(64, (VoroCfg<DIM, G, S>::WAVES)) k_bvoronoi_loop(BClipArgs args) {
    using C = VoroCfg<DIM, G, S>;
    __shared__ SeedSmem<C::NS, C::MAXP> sm[C::SPW];
    if constexpr (FRESH) {
        bool first = true;
        for (;;) {
            const auto* A0 = fresh_args<BClipArgs>();
            const int64_t nvb = (*A0->nq_dev + C::SPW - 1) / C::SPW;
            const int64_t vb = next_ticket(A0->ticket, 1, first);
            if (vb >= nvb) break;
            const auto* A = fresh_args<BClipArgs>();
            voronoi_group<DIM, G, S>(A->pts, INT64_MAX, A->seeds, *A->nq_dev, A->knn, A->k, A->max_p, A->max_t,
                                     face_arrays_of(A->fa), sm, vb);
            __syncthreads();
        }
    } else {
        const int64_t nq = *args.nq_dev;
        const int64_t nvb = (nq + C::SPW - 1) / C::SPW;
        const int chunk = ticket_chunk(nvb, CLIP_CHUNK);
        bool first = true;
        for (;;) {
            const int64_t b0 = next_ticket(args.ticket, chunk, first);
            if (b0 >= nvb) break;
            const int64_t b1 = b0 + chunk < nvb ? b0 + chunk : nvb;
            for (int64_t vb = b0; vb < b1; vb++) {
                voronoi_group<DIM, G, S>(args.pts, INT64_MAX, args.seeds, nq, args.knn, args.k, args.max_p,
                                         args.max_t, args.fa, sm, vb);
                __syncthreads();
            }
        }
    }
}

template <int DIM, int G, int S>
__global__ void __launch_bounds__(64, (VoroCfg<DIM, G, S>::WAVES))
    k_bvoronoi(const double* __restrict__ pts, const uint32_t* __restrict__ seeds, const int64_t* __restrict__ nq_dev,
               const uint32_t* __restrict__ knn, int k, int max_p, int max_t, FaceArrays fa) {
    using C = VoroCfg<DIM, G, S>;
    __shared__ SeedSmem<C::NS, C::MAXP> sm[C::SPW];
    const int64_t nq = *nq_dev;
    const int64_t vb = blockIdx.x;
    if (vb * C::SPW >= nq) return;  // whole block
    // no box holds k + 1 seeds or fewer: the exhaustive-row rule never applies
    voronoi_group<DIM, G, S>(pts, INT64_MAX, seeds, nq, knn, k, max_p, max_t, fa, sm, vb);
}

// status == code (code < 0: != SUCCESS) among the listed seeds
__global__ void __launch_bounds__(TB)
    k_bflag(const int64_t* __restrict__ nq_dev, int64_t cap, const uint32_t* __restrict__ seeds,
            const int32_t* __restrict__ status, int code, int32_t* __restrict__ flags) {
    const int64_t nq = *nq_dev;
    for (int64_t i = (int64_t)blockIdx.x * TB + threadIdx.x; i < cap; i += (int64_t)gridDim.x * TB) {
        int f = 0;
        if (i < nq) {
            const int st = status[seeds ? seeds[i] : (uint32_t)i];
            f = code < 0 ? (st != SUCCESS) : (st == code);
        }
        flags[i] = f;
    }
}

__global__ void __launch_bounds__(TB)
    k_bcompact(const int64_t* __restrict__ nq_dev, const uint32_t* __restrict__ seeds,
               const int32_t* __restrict__ flags, const int32_t* __restrict__ offs, const int32_t* __restrict__ total,
               uint32_t* __restrict__ list, int64_t* __restrict__ nout) {
    const int64_t nq = *nq_dev;
    if (blockIdx.x == 0 && threadIdx.x == 0) *nout = *total;
    for (int64_t i = (int64_t)blockIdx.x * TB + threadIdx.x; i < nq; i += (int64_t)gridDim.x * TB)
        if (flags[i]) list[offs[i]] = seeds ? seeds[i] : (uint32_t)i;
}

// ------------------------------------------------------------------ hydro
// Hydro tables of every box (k_finalize of hydro.hip with the box-major
// numbering): bad[0] hydro cells without faces, bad[1] periodic hydro faces
// on the box plane.
__global__ void __launch_bounds__(TB)
    k_bfinalize(const int64_t* __restrict__ n_dev, int64_t stride, const uint32_t* __restrict__ perm,
                const int32_t* __restrict__ orig_h, const uint8_t* __restrict__ count, const int32_t* __restrict__ nbr,
                const double* __restrict__ area, const double* __restrict__ vol, double ascale, double vscale,
                int periodic, int32_t* __restrict__ hid, double* __restrict__ area_u, double* __restrict__ vol_u,
                int32_t* __restrict__ o2s, int32_t* __restrict__ hsorted, int32_t* __restrict__ hcode, int32_t* bad) {
    const int64_t n = *n_dev;
    for (int64_t s = (int64_t)blockIdx.x * TB + threadIdx.x; s < n; s += (int64_t)gridDim.x * TB) {
        const int32_t h = orig_h[perm[s]];
        hsorted[s] = h;
        hcode[s] = h;  // ghosts: -1 - source, as orig_h
        if (h < 0) continue;
        o2s[h] = (int32_t)s;
        vol_u[s] = vscale * vol[s];
        const int c = count[s];
        if (c == 0) atomicAdd(bad, 1);
        for (int f = 0; f < c; f++) {
            const int64_t at = (int64_t)f * stride + s;
            const int32_t j = nbr[at];
            int32_t hj = -1;
            if (j >= 0) {
                const int32_t oj = orig_h[perm[j]];
                hj = oj >= 0 ? oj : -1 - oj;
            } else if (periodic) {
                atomicAdd(bad + 1, 1);
            }
            hid[at] = hj;
            area_u[at] = ascale * area[at];
        }
    }
}

// Per-box CFL step (finite_volume_solver.cpp:151-186 per box).  A wave
// takes 64 x DT_CELLS consecutive cells; a lane keeps the minimum of its
// current box (boxes hold consecutive cells) and reports a box it leaves;
// the lanes holding lane 0's box reduce together, the others report alone,
// and the waves of a block that end in one box report it once.
template <int DIM>
__global__ void __launch_bounds__(TB)
    k_bdt(int64_t n_h, const int32_t* __restrict__ hbox, const int32_t* __restrict__ o2s,
          const double* __restrict__ vol_u, double gam, double cfl, const double* __restrict__ rho,
          const double* __restrict__ vel, const double* __restrict__ E, unsigned long long* dt_box,
          const double* __restrict__ wgen, CflMesh cm) {
    const int64_t base = ((int64_t)blockIdx.x * TB + (threadIdx.x & ~63u)) * DT_CELLS + (threadIdx.x & 63u);
    double acc = 1e100;
    int bacc = -1;
    for (int it = 0; it < DT_CELLS; it++) {
        const int64_t i = base + 64 * it;
        if (i >= n_h) break;
        const int b = hbox[i];
        const double d = cell_cfl_dt<DIM>(i, o2s, vol_u, gam, cfl, rho, vel, E, wgen, cm);
        if (b != bacc) {
            if (bacc >= 0) report_dt(&dt_box[bacc], acc);
            bacc = b;
            acc = d;
        } else {
            acc = smin(d, acc);
        }
    }
    const int b0 = __shfl(bacc, 0, 64);
    double m = (bacc == b0) ? acc : 1e100;
    for (int d = 32; d > 0; d >>= 1) m = smin(__shfl_xor(m, d, 64), m);
    if (bacc >= 0 && bacc != b0) report_dt(&dt_box[bacc], acc);
    __shared__ double wm[TB / 64];
    __shared__ int wb[TB / 64];
    if ((threadIdx.x & 63) == 0) {
        wm[threadIdx.x >> 6] = m;
        wb[threadIdx.x >> 6] = b0;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int cb = wb[0];
        double cmin = wm[0];
        for (int w = 1; w < TB / 64; w++) {
            if (wb[w] == cb) {
                cmin = smin(wm[w], cmin);
            } else {
                if (cb >= 0) report_dt(&dt_box[cb], cmin);
                cb = wb[w];
                cmin = wm[w];
            }
        }
        if (cb >= 0) report_dt(&dt_box[cb], cmin);
    }
}

template <int DIM>
__global__ void __launch_bounds__(TB)
    k_bdrift(int64_t n_h, const int32_t* __restrict__ hbox, double* x, const double* __restrict__ vel,
             const double* __restrict__ dt_box, int periodic) {
    const int64_t i = (int64_t)blockIdx.x * TB + threadIdx.x;
    if (i >= n_h) return;
    const double dt = dt_box[hbox[i]];
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        double p = x[(size_t)DIM * i + d] + vel[(size_t)DIM * i + d] * dt;
        if (periodic) {
            p -= floor(p);
            if (p >= 1.0) p = 0.0;
        } else {
            if (p < 0.0) p = -p;
            if (p >= 1.0) p = 2.0 - p;
            p = fmin(fmax(p, 0.0), 0.9999999999999999);
        }
        x[(size_t)DIM * i + d] = p;
    }
}

// FP64 totals per box, one block per box, fixed summation order.
template <int DIM>
__global__ void __launch_bounds__(TB) k_btotals(const int64_t* __restrict__ hoff, SimView sv, double* out) {
    __shared__ double red[5][TB];
    const int b = blockIdx.x;
    double v[5] = {0, 0, 0, 0, 0};
    for (int64_t i = hoff[b] + threadIdx.x; i < hoff[b + 1]; i += TB) {
        v[0] += sv.mass[i];
#pragma unroll
        for (int d = 0; d < DIM; d++) v[1 + d] += sv.mom[(size_t)DIM * i + d];
        v[4] += sv.en[i];
    }
#pragma unroll
    for (int q = 0; q < 5; q++) red[q][threadIdx.x] = v[q];
    __syncthreads();
    for (int w = TB / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w)
#pragma unroll
            for (int q = 0; q < 5; q++) red[q][threadIdx.x] += red[q][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
#pragma unroll
        for (int q = 0; q < 5; q++) out[5 * b + q] = red[q][0];
}

template <int DIM>
__global__ void __launch_bounds__(TB)
    k_bcount_failed(const int64_t* __restrict__ n_dev, const int32_t* __restrict__ status, int32_t* nfail) {
    const int64_t n = *n_dev;
    int c = 0;
    for (int64_t s = (int64_t)blockIdx.x * TB + threadIdx.x; s < n; s += (int64_t)gridDim.x * TB)
        c += status[s] != SUCCESS;
    for (int d = 32; d > 0; d >>= 1) c += __shfl_xor(c, d, 64);
    if ((threadIdx.x & 63) == 0 && c) atomicAdd(nfail, c);
}

}  // namespace

// ------------------------------------------------------------------ launchers
void launch_box_setup(int nbox, int dim, const int64_t* hoff, const int32_t* goff, const int32_t* gtot,
                      int64_t seed_cap, int64_t cell_cap, int64_t* soff, int32_t* N, int64_t* coff, int32_t* err,
                      hipStream_t s) {
    k_box_setup<<<1, TB, 0, s>>>(nbox, dim, hoff, goff, gtot, seed_cap, cell_cap, soff, N, coff, err);
}

void launch_bseeds(int dim, const double* pts, int64_t n_h, double buff, int periodic, const int32_t* hbox,
                   const int64_t* hoff, const int32_t* goff, const int64_t* soff, const int32_t* err, double* out,
                   int32_t* sbox, int32_t* orig_h, hipStream_t s) {
    if (dim == 2)
        k_bseeds<2><<<nblk(n_h), TB, 0, s>>>(pts, n_h, buff, periodic, hbox, hoff, goff, soff, err, out, sbox, orig_h);
    else
        k_bseeds<3><<<nblk(n_h), TB, 0, s>>>(pts, n_h, buff, periodic, hbox, hoff, goff, soff, err, out, sbox, orig_h);
}

void launch_bgrid(int dim, const double* pts, const int64_t* n_dev, int64_t cap, int64_t cell_cap,
                  const int32_t* sbox, const int32_t* N, const int64_t* coff, int32_t* cell, int32_t* counts,
                  int32_t* starts, int32_t* fill, int32_t* tmp, int32_t* total, uint32_t* perm, double* sorted,
                  int32_t* sorted_box, int32_t* big, hipStream_t s) {
    const unsigned gb = std::min<unsigned>(nblk(cap), 2048);
    const unsigned gc = std::min<unsigned>(nblk(cell_cap), 2048);
    k_fill<int32_t><<<gc, TB, 0, s>>>(counts, cell_cap, 0);
    k_fill<int32_t><<<gc, TB, 0, s>>>(fill, cell_cap, 0);
    if (dim == 2) k_bcell_count<2><<<gb, TB, 0, s>>>(pts, n_dev, cap, sbox, N, coff, cell, counts);
    else k_bcell_count<3><<<gb, TB, 0, s>>>(pts, n_dev, cap, sbox, N, coff, cell, counts);
    scan_exclusive_i32(counts, starts, cell_cap, tmp, total, s);
    k_bcell_scatter<<<gb, TB, 0, s>>>(n_dev, cap, cell, starts, fill, perm);
    launch_cell_sort(cell_cap, starts, counts, perm, big + 1, big, s);
    if (dim == 2) k_bgather<2><<<gb, TB, 0, s>>>(n_dev, cap, perm, pts, sbox, sorted, sorted_box);
    else k_bgather<3><<<gb, TB, 0, s>>>(n_dev, cap, perm, pts, sbox, sorted, sorted_box);
}

namespace {
template <int DIM, int BLOCK>
void bknn_heap(const double* sorted, const uint32_t* perm, const int32_t* sbox, const int32_t* N,
               const int64_t* coff, const int32_t* starts, const int32_t* counts, const uint32_t* q,
               const int64_t* nq_dev, int64_t cap, int k, uint32_t* out, int blocks_per_cu, hipStream_t s) {
    const size_t smem = (size_t)k * BLOCK * (sizeof(double) + sizeof(uint32_t));
    // raise the LDS limit once (the first call is an eager launch, never a
    // captured one)
    static const bool raised = [] {
        (void)hipFuncSetAttribute((const void*)k_bknn<DIM, BLOCK>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  160 * 1024);
        return true;
    }();
    (void)raised;
    const unsigned nb = (unsigned)std::min<int64_t>((cap + BLOCK - 1) / BLOCK, 256LL * blocks_per_cu);
    k_bknn<DIM, BLOCK><<<std::max(nb, 1u), BLOCK, smem, s>>>(sorted, perm, sbox, N, coff, starts, counts, q, nq_dev,
                                                             k, out);
}
template <int DIM, int M>
void bknn_wave(const double* sorted, const uint32_t* perm, const int32_t* sbox, const int32_t* N,
               const int64_t* coff, const int32_t* starts, const int32_t* counts, const uint32_t* q,
               const int64_t* nq_dev, int64_t cap, int k, uint32_t* out, int blocks_per_cu,
               unsigned long long* ticket, hipStream_t s, KnnTail tail) {
    static const int occ = occupancy(k_bknn_wave<DIM, M>);
    const unsigned nb = (unsigned)std::min<int64_t>(cap, resident_blocks(occ, 8 * blocks_per_cu));
    k_fill<unsigned long long><<<1, TB, 0, s>>>(ticket, 1, 0ull);
    const BKnnArgs args{sorted, perm, sbox, N, coff, starts, counts, q, nq_dev, out, ticket, tail, k};
    k_bknn_wave<DIM, M><<<std::max(nb, 1u), 64, 0, s>>>(args);
}
template <int DIM>
void bknn_dim(const double* sorted, const uint32_t* perm, const int32_t* sbox, const int32_t* N,
              const int64_t* coff, const int32_t* starts, const int32_t* counts, const uint32_t* q,
              const int64_t* nq_dev, int64_t cap, int k, uint32_t* out, int wave_min_k, int blocks_per_cu,
              unsigned long long* ticket, hipStream_t s, KnnTail tail) {
    if (k >= wave_min_k) {
        if (k <= 64) bknn_wave<DIM, 1>(sorted, perm, sbox, N, coff, starts, counts, q, nq_dev, cap, k, out, blocks_per_cu, ticket, s, tail);
        else if (k <= 128) bknn_wave<DIM, 2>(sorted, perm, sbox, N, coff, starts, counts, q, nq_dev, cap, k, out, blocks_per_cu, ticket, s, tail);
        else if (k <= 192) bknn_wave<DIM, 3>(sorted, perm, sbox, N, coff, starts, counts, q, nq_dev, cap, k, out, blocks_per_cu, ticket, s, tail);
        else bknn_wave<DIM, 4>(sorted, perm, sbox, N, coff, starts, counts, q, nq_dev, cap, k, out, blocks_per_cu, ticket, s, tail);
    } else if (k <= 208) {
        bknn_heap<DIM, 64>(sorted, perm, sbox, N, coff, starts, counts, q, nq_dev, cap, k, out, blocks_per_cu, s);
    } else {
        bknn_heap<DIM, 32>(sorted, perm, sbox, N, coff, starts, counts, q, nq_dev, cap, k, out, blocks_per_cu, s);
    }
}
template <int DIM, int G, int S>
void bclip(const double* pts, const uint32_t* seeds, const int64_t* nq_dev, int64_t cap, const uint32_t* knn, int k,
           int max_p, int max_t, FaceArrays fa, int blocks_per_cu, unsigned long long* ticket, bool loop,
           hipStream_t s) {
    using C = VoroCfg<DIM, G, S>;
    if (loop) {
        constexpr bool FRESH = DIM == 3;
        static const int occ = occupancy(k_bvoronoi_loop<DIM, G, S, FRESH>);
        const unsigned nb =
            (unsigned)std::min<int64_t>((cap + C::SPW - 1) / C::SPW, resident_blocks(occ, blocks_per_cu));
        k_fill<unsigned long long><<<1, TB, 0, s>>>(ticket, 1, 0ull);
        const BClipArgs args{pts, seeds, nq_dev, knn, ticket, fa, k, max_p, max_t};
        k_bvoronoi_loop<DIM, G, S, FRESH><<<std::max(nb, 1u), 64, 0, s>>>(args);
        return;
    }
    const unsigned nb = (unsigned)std::max<int64_t>((cap + C::SPW - 1) / C::SPW, 1);
    k_bvoronoi<DIM, G, S><<<nb, 64, 0, s>>>(pts, seeds, nq_dev, knn, k, max_p, max_t, fa);
}
}  // namespace

void launch_bknn(int dim, const double* sorted, const uint32_t* perm, const int32_t* sbox, const int32_t* N,
                 const int64_t* coff, const int32_t* starts, const int32_t* counts, const uint32_t* queries,
                 const int64_t* nq_dev, int64_t cap, int k, uint32_t* out, int wave_min_k, int blocks_per_cu,
                 unsigned long long* ticket, hipStream_t s, KnnTail tail) {
    if (dim == 2)
        bknn_dim<2>(sorted, perm, sbox, N, coff, starts, counts, queries, nq_dev, cap, k, out, wave_min_k,
                    blocks_per_cu, ticket, s, tail);
    else
        bknn_dim<3>(sorted, perm, sbox, N, coff, starts, counts, queries, nq_dev, cap, k, out, wave_min_k,
                    blocks_per_cu, ticket, s, tail);
}

void launch_bvoronoi(int dim, const double* sorted, const uint32_t* seeds, const int64_t* nq_dev, int64_t cap,
                     const uint32_t* knn, int k, int max_p, int max_t, int big, int lanes, int slots, FaceArrays fa,
                     int blocks_per_cu, unsigned long long* ticket, bool loop, hipStream_t s) {
    auto go = [&](auto dimc) {
        constexpr int D = decltype(dimc)::value;
        if (big) bclip<D, 16, 8>(sorted, seeds, nq_dev, cap, knn, k, max_p, max_t, fa, blocks_per_cu, ticket, loop, s);
        else if (lanes == 8 && slots == 48) bclip<D, 8, 6>(sorted, seeds, nq_dev, cap, knn, k, max_p, max_t, fa, blocks_per_cu, ticket, loop, s);
        else if (lanes == 8) bclip<D, 8, 8>(sorted, seeds, nq_dev, cap, knn, k, max_p, max_t, fa, blocks_per_cu, ticket, loop, s);
        else if (slots == 48) bclip<D, 16, 3>(sorted, seeds, nq_dev, cap, knn, k, max_p, max_t, fa, blocks_per_cu, ticket, loop, s);
        else bclip<D, 16, 4>(sorted, seeds, nq_dev, cap, knn, k, max_p, max_t, fa, blocks_per_cu, ticket, loop, s);
    };
    if (dim == 2) go(std::integral_constant<int, 2>{});
    else go(std::integral_constant<int, 3>{});
}

void launch_bcollect(const int64_t* nq_dev, int64_t cap, const uint32_t* seeds, const int32_t* status, int code,
                     uint32_t* list, int32_t* flags, int32_t* offs, int32_t* tmp, int32_t* total, int64_t* nout,
                     hipStream_t s) {
    const unsigned gb = std::min<unsigned>(nblk(cap), 2048);
    k_bflag<<<gb, TB, 0, s>>>(nq_dev, cap, seeds, status, code, flags);
    scan_exclusive_i32(flags, offs, cap, tmp, total, s);
    k_bcompact<<<gb, TB, 0, s>>>(nq_dev, seeds, flags, offs, total, list, nout);
}

void launch_bfinalize(const int64_t* n_dev, int64_t cap, int64_t stride, const uint32_t* perm, const int32_t* orig_h,
                      const uint8_t* count, const int32_t* nbr, const double* area, const double* vol, double ascale,
                      double vscale, int periodic, int32_t* hid, double* area_u, double* vol_u, int32_t* o2s,
                      int32_t* hsorted, int32_t* hcode, int32_t* bad, hipStream_t s) {
    k_bfinalize<<<std::min<unsigned>(nblk(cap), 2048), TB, 0, s>>>(n_dev, stride, perm, orig_h, count, nbr, area, vol,
                                                                   ascale, vscale, periodic, hid, area_u, vol_u, o2s,
                                                                   hsorted, hcode, bad);
}

void launch_bdt(int dim, int64_t n_h, int nbox, const int32_t* hbox, const int32_t* o2s, const double* vol_u,
                double gam, double cfl, const double* rho, const double* vel, const double* E, double* dt_box,
                hipStream_t s, const double* wgen, CflMesh cm) {
    k_fill<double><<<nblk(nbox), TB, 0, s>>>(dt_box, nbox, 1e100);
    if (dim == 2)
        k_bdt<2><<<nblk_dt(n_h), TB, 0, s>>>(n_h, hbox, o2s, vol_u, gam, cfl, rho, vel, E, (unsigned long long*)dt_box,
                                          wgen, cm);
    else
        k_bdt<3><<<nblk_dt(n_h), TB, 0, s>>>(n_h, hbox, o2s, vol_u, gam, cfl, rho, vel, E, (unsigned long long*)dt_box,
                                          wgen, cm);
}

void launch_bdrift(int dim, int64_t n_h, const int32_t* hbox, double* x, const double* vel, const double* dt_box,
                   int periodic, hipStream_t s) {
    if (dim == 2) k_bdrift<2><<<nblk(n_h), TB, 0, s>>>(n_h, hbox, x, vel, dt_box, periodic);
    else k_bdrift<3><<<nblk(n_h), TB, 0, s>>>(n_h, hbox, x, vel, dt_box, periodic);
}

void launch_btotals(int dim, int nbox, const int64_t* hoff, SimView sv, double* out, hipStream_t s) {
    if (dim == 2) k_btotals<2><<<nbox, TB, 0, s>>>(hoff, sv, out);
    else k_btotals<3><<<nbox, TB, 0, s>>>(hoff, sv, out);
}

void launch_bcount_failed(const int64_t* n_dev, int64_t cap, const int32_t* status, int32_t* nfail, hipStream_t s) {
    k_bcount_failed<2><<<std::min<unsigned>(nblk(cap), 1024), TB, 0, s>>>(n_dev, status, nfail);
}

void launch_fill_i32(int32_t* a, int64_t n, int32_t v, hipStream_t s) {
    k_fill<int32_t><<<std::min<unsigned>(nblk(n), 1024), TB, 0, s>>>(a, n, v);
}

}  // namespace pg
