// k-nearest-neighbour search of one query point (device code shared by the
// single-box and batched kernels of knn.hip and batch.hip).
#pragma once

#include "common.h"
#include "kernels.h"
#include "simt.h"
#include "subgrid.h"

namespace pg {

// Lower bound of the squared distance from a query to any point stored in
// grid cell (ii, jj, kk).  P = the query's coordinates times N, the same
// product grid_cell() floors, so a point of cell i has P_q in [i, i + 1)
// per axis; boundary cells are open towards the outside (points beyond
// [0, 1) are clamped into them).  A slack of 1e-9 cell widths and a factor
// 1 - 1e-12 keep the bound below every candidate's rounded distance, so a
// cell whose bound exceeds the current k-th distance holds no candidate
// that could enter the row: skipping it leaves the row unchanged (the
// reference's ring rule, knn.cpp:272, still decides when the walk stops).
template <int DIM>
PG_DEV double cell_lower_d2(const double (&P)[3], int ii, int jj, int kk, int N) {
    const int idx[3] = {ii, jj, kk};
    double s = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        const double lo = idx[d] == 0 ? -1e300 : (double)idx[d];
        const double hi = idx[d] == N - 1 ? 1e300 : (double)(idx[d] + 1);
        const double g = fmax(fmax(lo - P[d], P[d] - hi) - 1e-9, 0.0);
        s += g * g;
    }
    const double n = (double)N;
    return s / (n * n) * (1.0 - 1e-12);
}

// One query of the per-thread heap search; the heap lives in LDS at
// hd/hi[entry * BLOCK + threadIdx.x].  Grid cells of this query's box are
// cbase + i + j N + k N^2.
template <int DIM, int BLOCK>
PG_DEV void knn_heap_query(const double* __restrict__ sorted, const uint32_t* __restrict__ perm, int N,
                           int64_t cbase, const int32_t* __restrict__ starts,
                           const int32_t* __restrict__ counts, uint32_t q, int k, uint32_t* __restrict__ o,
                           double* hd, uint32_t* hi) {
    const int t = threadIdx.x;
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
    double P[3] = {0, 0, 0};
#pragma unroll
    for (int d = 0; d < DIM; d++) P[d] = p[d] * (double)N;
    auto scan_cell = [&](int ii, int jj, int kk) {
        if (size == k && cell_lower_d2<DIM>(P, ii, jj, kk, N) > HD(0)) return;
        const int64_t cc = cbase + ii + jj * N + (int64_t)kk * N * N;
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
    // heapsort to ascending order (knn.cpp:341-347)
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
    for (int e = 0; e < k; e++) o[e] = e < size ? HI(e) : 0xffffffffu;
#undef HD
#undef HI
}

// ------------------------------------------------------------------------
// Wave-per-query variant.  The 64 lanes of a wave cooperate on one query:
// the best 64 M (dist^2, sorted id) are kept sorted in registers,
// slot-major (rank r at slot r / 64 of lane r % 64, M = ceil(k/64) slots).
// Ring cells are enumerated 64 at a time, their point ranges prefix-summed
// across the wave, and candidates processed 64 at a time: those inside the
// query's ball (a radius guessed from the local density, so that one or
// two sorts usually make the row) and ahead of the current k-th entry
// (broadcast) gather in LDS; every 64 of them are sorted across the wave
// by a bitonic network and merged into the kept list, one 64-entry slot at
// a time (lower half of slot + reversed batch stays, the upper half carries
// into the next slot).
// Same rows and same total order as k_knn, no LDS heap, full waves: used
// for rows of 20 and more, where the per-thread heap would leave the CUs
// nearly empty.

PG_DEV bool cand_less(double da, uint32_t oa, double db, uint32_t ob) {
    return da < db || (da == db && oa < ob);
}

// The wave kernel's entries are (FP64 d^2, sorted id); the order is
// (d^2, original id), and the original id (perm of the sorted id) is read
// only for equal distances.  Padding entries are (+inf, 0xffffffff).
PG_DEV uint32_t orig_id(const uint32_t* __restrict__ perm, uint32_t s) {
    return s == 0xffffffffu ? 0xffffffffu : perm[s];
}
PG_DEV bool ent_less(const uint32_t* __restrict__ perm, double da, uint32_t sa, double db, uint32_t sb) {
    if (da != db) return da < db;
    return orig_id(perm, sa) < orig_id(perm, sb);
}

// Compare-exchange with lane l ^ X: the lane whose bit B is clear keeps the
// smaller entry of the pair under ent_less, the other the larger.
template <int X, int B>
PG_DEV void wave_cx(const uint32_t* __restrict__ perm, double& d, uint32_t& s, int l) {
    const double pd = lane_xor_d<X>(d);
    const uint32_t ps = lane_xor<X>(s);
    const bool take = ent_less(perm, pd, ps, d, s) == ((l & B) == 0);
    d = take ? pd : d;
    s = take ? ps : s;
}

// Half-cleaners J, J/2, .., 1: sorts a bitonic sequence held in blocks of
// 2 J lanes (ascending with the lane index).
template <int J>
PG_DEV void wave_half_clean(const uint32_t* __restrict__ perm, double& d, uint32_t& s, int l) {
    if constexpr (J >= 1) {
        wave_cx<J, J>(perm, d, s, l);
        wave_half_clean<J / 2>(perm, d, s, l);
    }
}

// Bitonic sort with one direction throughout: blocks of K lanes hold two
// ascending halves; the mirror exchange (lane l with l ^ (K - 1)) leaves
// two bitonic halves with the lower below the upper, and the half-cleaners
// sort each.  K = 2, 4, .., 64.
template <int K>
PG_DEV void wave_sort_phase(const uint32_t* __restrict__ perm, double& d, uint32_t& s, int l) {
    wave_cx<K - 1, K / 2>(perm, d, s, l);
    wave_half_clean<K / 4>(perm, d, s, l);
    if constexpr (K < 64) wave_sort_phase<2 * K>(perm, d, s, l);
}

// t / d for 0 <= t < 2^22, 1 <= d, rd = 1/d rounded: the float quotient is
// within one of the exact one (relative error below 2^-22 on a value below
// 2^22), and one step corrects it.
PG_DEV int idiv_small(int t, int d, float rd) {
    int q = (int)((float)t * rd);
    const int r = t - q * d;
    if (r < 0) q--;
    else if (r >= d) q++;
    return q;
}

// Bitonic sort of one entry per lane, ascending with the lane index.
PG_DEV void wave_sort64(const uint32_t* __restrict__ perm, double& d, uint32_t& s, int l) {
    wave_sort_phase<2>(perm, d, s, l);
}

// Sort a bitonic sequence held one entry per lane (ascending).
PG_DEV void wave_merge64(const uint32_t* __restrict__ perm, double& d, uint32_t& s, int l) {
    wave_half_clean<32>(perm, d, s, l);
}

// Candidates that beat the k-th kept entry wait in LDS until 64 have
// gathered; only then are they sorted and merged into the kept list.
struct KnnPend {
    double d[128];
    uint32_t s[128];
};

// Merge 64 candidates (one per lane, unsorted, +inf padded) into the kept
// list (slot-major, ascending): sort them across the wave (bitonic), then,
// slot by slot, the lower half of (slot, batch) stays and the upper half
// carries on.
template <int M>
PG_DEV void wave_merge_batch(const uint32_t* __restrict__ perm, double (&kd)[M], uint32_t (&ks)[M], double dd,
                             uint32_t cs, int l) {
    wave_sort64(perm, dd, cs, l);
#pragma unroll
    for (int jj = 0; jj < M; jj++) {
        const double rd = lane_xor_d<63>(dd);  // lane 63 - l
        const uint32_t rs = lane_xor<63>(cs);
        const bool mine = ent_less(perm, kd[jj], ks[jj], rd, rs);
        double hd = mine ? rd : kd[jj];
        uint32_t hs = mine ? rs : ks[jj];
        if (!mine) {
            kd[jj] = rd;
            ks[jj] = rs;
        }
        wave_merge64(perm, kd[jj], ks[jj], l);
        if (jj + 1 < M) {  // the upper half carries into the next slot
            wave_merge64(perm, hd, hs, l);
            dd = hd;
            cs = hs;
        }
    }
}

// Rows of a resumed seed start after its saved row length (KnnTail).
PG_DEV void tail_of(const KnnTail& t, uint32_t q, int k, int& kp, uint32_t& lo) {
    kp = 0;
    lo = 0;
    if (t.resume == nullptr) return;
    const int32_t r = t.resume[q];
    if (r > 0 && (r & 0x7fff0000) == t.tag && (r & 0xffff) < k) {
        kp = r & 0xffff;
        lo = t.lastnb[q];
    }
}

// Squared radius of the first pass of a wave query: the ball expected to
// hold lam points at the density of the 3^DIM block of grid cells around
// the query (cnt points, the query included, in nin in-grid cells of side
// 1/N), widened where the unit box cuts it (the share of the cube
// [p - R, p + R] inside the box stands for the ball's).  A heuristic: the
// rows do not depend on it, only the work to find them.
template <int DIM>
PG_DEV double knn_guess_r2(const double (&p)[3], int N, int cnt, int nin, float lam) {
    // float arithmetic and the hardware's log2 / exp2: a heuristic needs
    // no more
    const float h = 1.0f / (float)N;
    const float vcell = DIM == 3 ? h * h * h : h * h;
    const float rho = (float)max(cnt - 1, 1) / ((float)max(nin, 1) * vcell);
    // ball volume lam / rho = (4 pi / 3) R^3 in 3D, pi R^2 in 2D
    const float v = lam / rho * (DIM == 3 ? 0.2387324146f : 0.3183098862f);
    float R = __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(v) * (1.0f / DIM));
    float f = 1.0f;
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        const float x = (float)p[d];
        f *= (fminf(x, R) + fminf(1.0f - x, R)) / (2.0f * R);
    }
    f = fmaxf(f, 1.0f / 64.0f);
    R *= __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(f) * (-1.0f / DIM));
    return (double)R * (double)R;
}

// One query by the 64 lanes of a one-wave block; c_start / c_pfx are the
// block's LDS tables (64 and 65 entries), pend its candidate buffer.  Grid
// cells of the query's box are cbase + i + j N + k N^2.  kp > 0: the first
// kp neighbours are known, lo_s (sorted id) the last of them; only
// neighbours kp..k-1 are found (candidates up to lo_s in the order are
// skipped) and written to o[kp..k).
//
// The walk is the reference's (knn.cpp:237-347): rings of cells in
// Chebyshev distance, stopped before ring r once the k-th distance is below
// ((r - 1) / N)^2.  A pass also carries a ball of squared radius r2 (from
// knn_guess_r2): only candidates with d^2 <= r2 are kept, cells whose lower
// bound exceeds r2 are skipped, and the walk ends before the first ring r
// with ((r - 1) / N)^2 > r2.  If at least k candidates lie in the ball,
// their k first in the (d^2, original id) order are the reference's row:
//  * every point the reference's row could hold within the ball was seen
//    (pruned cells and rings hold none), and none outside it can rank
//    before the ball's k-th;
//  * the reference stops no later than the ring after the ball's last
//    (its k-th distance there is at most the ball's k-th <= r2 <
//    ((r - 1) / N)^2), and no earlier ring can stop it with a kept
//    candidate beyond it, as long as every kept candidate of ring r has
//    d^2 >= ((r - 1) / N)^2, the bound its ring gives.  A rounded distance
//    below that bound (a point within rounding of a ring's edge) is checked
//    for, and sends the query to a pass with no ball: the reference's walk
//    itself.
// Fewer than k candidates in the ball: a new pass from scratch with a wider
// one (after three, none).
template <int DIM, int M, bool SUB = false>
PG_DEV void knn_wave_query(const double* __restrict__ sorted, const uint32_t* __restrict__ perm, int N,
                           int64_t cbase, const int32_t* __restrict__ starts,
                           const int32_t* __restrict__ counts, uint32_t q, int k, uint32_t* __restrict__ o,
                           int32_t* c_start, int32_t* c_pfx, KnnPend* pend, int kp = 0, uint32_t lo_s = 0,
                           SubGrid sg = {}) {
    const int l = threadIdx.x;
    double p[3] = {0, 0, 0};
#pragma unroll
    for (int d = 0; d < DIM; d++) p[d] = sorted[(size_t)DIM * q + d];
    const int kfull = k;
    double lo_d = -1.0;  // kp = 0: no lower bound
    if (kp > 0) {
        double pc[3] = {0, 0, 0};
#pragma unroll
        for (int d = 0; d < DIM; d++) pc[d] = sorted[(size_t)DIM * lo_s + d];
        lo_d = dist2<DIM>(p, pc);
        o += kp;
        k -= kp;
    }
    double P[3] = {0, 0, 0};
#pragma unroll
    for (int d = 0; d < DIM; d++) P[d] = p[d] * (double)N;
    const int c = grid_cell<DIM>(p, N);
    const int ci = c % N, cj = (c / N) % N, ck = DIM == 3 ? c / (N * N) : 0;

    double kd[M];
    uint32_t ks[M];
    int size = 0, npend = 0;
    const int kl = (k - 1) & 63, kj = (k - 1) >> 6;  // lane and slot of rank k-1
    auto kth_entry = [&](double& d, uint32_t& s) {
#pragma unroll
        for (int j = 0; j < M; j++)
            if (j == kj) {
                d = read_lane_d(kd[j], kl);
                s = read_lane(ks[j], kl);
            }
    };
    // merge the first n (<= 64) waiting candidates into the kept list; into
    // an empty list the sorted batch is the list
    auto flush = [&](int n) {
        __syncthreads();
        double dd = __builtin_inf();
        uint32_t cs = 0xffffffffu;
        if (l < n) {
            dd = pend->d[l];
            cs = pend->s[l];
        }
        if (size == 0) {
            wave_sort64(perm, dd, cs, l);
            kd[0] = dd;
            ks[0] = cs;
        } else {
            wave_merge_batch<M>(perm, kd, ks, dd, cs, l);
        }
        size = min(k, size + n);
    };

    // points the first ball should hold: k and a margin for the spread of
    // the count (about 1% of passes come out short on uniform seeds)
    const float lam = 1.25f * (float)kfull + 3.0f * __builtin_sqrtf((float)kfull) + 6.0f;
    double r2;  // +inf: no ball (the reference's walk alone)
    {
        // the first ball from the density of the 3^DIM block of cells
        // around the query
        int cnt = 0;
        bool in = false;
        if (l < (DIM == 3 ? 27 : 9)) {
            const int di = l % 3 - 1, dj = (l / 3) % 3 - 1, dk = DIM == 3 ? l / 9 - 1 : 0;
            const int ii = ci + di, jj = cj + dj, kk = ck + dk;
            in = ii >= 0 && jj >= 0 && kk >= 0 && ii < N && jj < N && (DIM == 2 || kk < N);
            if (in) cnt = counts[cbase + ii + jj * N + (int64_t)kk * N * N];
        }
        const int tot = (int)read_lane((uint32_t)wave_incl_sum(cnt), 63);
        r2 = knn_guess_r2<DIM>(p, N, tot, __popcll((unsigned long long)__ballot(in)), lam);
    }
    bool odd_l = false;  // a kept candidate below its ring's distance bound
    // Candidates 0 .. total-1 of the cells (or sub-cells) in c_start / c_pfx
    // (64 entries: first position, exclusive prefix of the counts), 64 at a
    // time; positions are sorted ids, or indices into indir.  Those inside
    // the ball and ahead of the current k-th entry join the LDS buffer.
    auto scan = [&](int total, const uint32_t* __restrict__ indir, double md2) {
        for (int cb = 0; cb < total; cb += 64) {
            const int j = cb + l;
            bool take = false;
            double dd = 0;
            uint32_t cs = 0;
            if (j < total) {
                // cell holding candidate j: largest a with c_pfx[a] <= j
                int lo = 0, hi = 63;
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (c_pfx[mid] <= j) lo = mid;
                    else hi = mid - 1;
                }
                cs = (uint32_t)(c_start[lo] + (j - c_pfx[lo]));
                if (indir != nullptr) cs = indir[cs];
                if (cs != q) {
                    double pc[3] = {0, 0, 0};
#pragma unroll
                    for (int d = 0; d < DIM; d++) pc[d] = sorted[(size_t)DIM * cs + d];
                    dd = dist2<DIM>(p, pc);
                    take = dd <= r2 && (kp == 0 || ent_less(perm, lo_d, lo_s, dd, cs));
                }
            }
            if (size == k) {
                double kth;
                uint32_t kts;
                kth_entry(kth, kts);
                take = take && ent_less(perm, dd, cs, kth, kts);
            }
            // a kept candidate nearer than its ring's bound (rounding at a
            // ring's edge): the ball's argument fails
            odd_l = odd_l || (take && dd < md2);
            const uint64_t surv = (uint64_t)__ballot(take);
            if (surv == 0ull) continue;
            // survivors join the LDS buffer in lane order (npend < 64, at
            // most 64 join: positions stay below 128)
            if (take) {
                const int at = npend + __popcll((unsigned long long)(surv & ((1ull << l) - 1ull)));
                pend->d[at] = dd;
                pend->s[at] = cs;
            }
            npend += __popcll((unsigned long long)surv);
            if (npend >= 64) {
                flush(64);
                // move the rest (fewer than 64) to the front
                double rd = 0;
                uint32_t rs = 0;
                const int rest = npend - 64;
                if (l < rest) {
                    rd = pend->d[64 + l];
                    rs = pend->s[64 + l];
                }
                __syncthreads();
                if (l < rest) {
                    pend->d[l] = rd;
                    pend->s[l] = rs;
                }
                npend = rest;
            }
        }
    };
    for (int pass = 0;; pass++) {
#pragma unroll
        for (int j = 0; j < M; j++) {
            kd[j] = __builtin_inf();
            ks[j] = 0xffffffffu;
        }
        size = 0;
        npend = 0;
        odd_l = false;
        bool stop = false;
        // the first step takes the cells at Chebyshev distance 0 and 1
        // together (the reference's ring-1 test cannot stop the walk: its
        // bound is 0), then ring by ring
        for (int ring = 0; ring < N; ring = ring == 0 ? 2 : ring + 1) {
            const double md = (double)(ring - 1) / (double)N;
            const double md2 = ring >= 2 ? md * md : 0.0;  // ring's distance bound
            if (ring >= 2) {
                // the ring test (and the cell pruning) need the exact k-th
                // entry; while fewer than k candidates are known there is
                // none: the waiting candidates stay in LDS
                if (npend > 0 && size + npend >= k) {
                    flush(npend);
                    npend = 0;
                }
                if (size == k) {
                    double kth;
                    uint32_t kts;
                    kth_entry(kth, kts);
                    if (kth < md2) {
                        stop = true;
                        break;
                    }
                }
                if (md2 > r2) break;  // the ring lies beyond the ball
            }
            const int reach = ring == 0 ? 1 : ring;  // cube half-side
            const int side = 2 * reach + 1;
            const int ncube = DIM == 3 ? side * side * side : side * side;
            const float rside = 1.0f / (float)side;
            for (int base = 0; base < ncube; base += 64) {
                // this lane's cell of the cube, if it lies on the ring and in the grid
                const int t = base + l;
                int cnt = 0, st = 0;
                // cells wholly beyond the ball or the current k-th distance are skipped
                double prune = r2;
                if (size == k) {
                    double kth;
                    uint32_t kts;
                    kth_entry(kth, kts);
                    prune = fmin(prune, kth);
                }
                bool crowded = false;  // a cell with a sub-grid: taken sub-cell by sub-cell below
                int ii = 0, jj = 0, kk = 0;
                if (t < ncube) {
                    // t = (dk + reach) side^2 + (dj + reach) side + (di + reach)
                    int q1, q2;
                    if (ncube < (1 << 22)) {  // float quotient, one correction step
                        q1 = idiv_small(t, side, rside);
                        q2 = idiv_small(q1, side, rside);
                    } else {
                        q1 = t / side;
                        q2 = q1 / side;
                    }
                    const int di = t - q1 * side - reach, dj = q1 - q2 * side - reach;
                    const int dk = DIM == 3 ? q2 - reach : 0;
                    const int cheb = max(abs(di), max(abs(dj), abs(dk)));
                    ii = ci + di;
                    jj = cj + dj;
                    kk = ck + dk;
                    const bool cell_in = (cheb == ring || ring == 0) && ii >= 0 && jj >= 0 && kk >= 0 && ii < N &&
                                         jj < N && (DIM == 2 || kk < N);
                    if (cell_in && cell_lower_d2<DIM>(P, ii, jj, kk, N) <= prune) {
                        const int64_t cc = cbase + ii + jj * N + (int64_t)kk * N * N;
                        if (SUB && sg.m[cc] > 0) {
                            crowded = true;
                        } else {
                            st = starts[cc];
                            cnt = counts[cc];
                        }
                    }
                }
                const int incl = wave_incl_sum(cnt);
                const int total = (int)read_lane((uint32_t)incl, 63);
                if (total > 0) {
                    c_start[l] = st;
                    c_pfx[l + 1] = incl;
                    if (l == 0) c_pfx[0] = 0;
                    __syncthreads();
                    scan(total, nullptr, md2);
                    __syncthreads();
                }
                // the crowded cells of this batch, one at a time: their
                // sub-cells that the pruning distance can reach
                uint64_t crowd = SUB ? (uint64_t)__ballot(crowded) : 0ull;
                while (SUB && crowd != 0ull) {
                    const int a = __builtin_ctzll(crowd);
                    crowd &= crowd - 1ull;
                    const int ix[3] = {(int)read_lane((uint32_t)ii, a), (int)read_lane((uint32_t)jj, a),
                                       (int)read_lane((uint32_t)kk, a)};
                    const int64_t cc = cbase + ix[0] + ix[1] * N + (int64_t)ix[2] * N * N;
                    const int m = sg.m[cc];
                    const int64_t sbase = sg.ptr[cc];
                    double pd = r2;
                    if (size == k) {
                        double kth;
                        uint32_t kts;
                        kth_entry(kth, kts);
                        pd = fmin(pd, kth);
                    }
                    int lo[3], hi[3];
                    sub_box<DIM>(P, ix, m, N, pd, lo, hi);
                    const int nx = hi[0] - lo[0] + 1, ny = hi[1] - lo[1] + 1, nz = hi[2] - lo[2] + 1;
                    const int nbox = nx * ny * nz;
                    for (int b0 = 0; b0 < nbox; b0 += 64) {
                        const int u = b0 + l;
                        int sst = 0, scnt = 0;
                        if (u < nbox) {
                            const int sx[3] = {lo[0] + u % nx, lo[1] + (u / nx) % ny, lo[2] + u / (nx * ny)};
                            if (sub_lower_d2<DIM>(P, ix, sx, m, N) <= pd) {
                                const int64_t j = sbase + sx[0] + (int64_t)m * (sx[1] + (int64_t)m * sx[2]);
                                sst = sg.start[j];
                                scnt = sg.cnt[j];
                            }
                        }
                        const int sincl = wave_incl_sum(scnt);
                        const int stotal = (int)read_lane((uint32_t)sincl, 63);
                        if (stotal == 0) continue;
                        c_start[l] = sst;
                        c_pfx[l + 1] = sincl;
                        if (l == 0) c_pfx[0] = 0;
                        __syncthreads();
                        scan(stotal, sg.idx, md2);
                        __syncthreads();
                    }
                }
            }
        }
        if (r2 < __builtin_inf() && __ballot(odd_l) != 0ull) {  // walk again, no ball
            r2 = __builtin_inf();
            continue;
        }
        if (stop || size + npend >= k || !(r2 < __builtin_inf())) break;
        // fewer than k candidates in the ball: a wider one, from scratch
        const float have = (float)(kp + size + npend) + 1.0f;
        const float g = __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(lam / have) * (2.0f / DIM));
        r2 = pass >= 2 ? __builtin_inf() : r2 * (double)fminf(fmaxf(g, 1.5f), 16.0f);
    }
    if (npend > 0) flush(npend);
#pragma unroll
    for (int jj = 0; jj < M; jj++) {
        const int r = jj * 64 + l;
        if (r < k) o[r] = r < size ? ks[jj] : 0xffffffffu;
    }
}

}  // namespace pg
