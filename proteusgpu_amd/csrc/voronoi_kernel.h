// Meshless Voronoi clipping on CDNA4: one 16-lane group per seed.
//
// Algorithm (cell.cpp:42-411, voronoi.cpp:115-156, after Ray et al. 2018):
// a seed's cell starts as the box [-1e-14, 1+1e-14]^d and is clipped by the
// bisector planes of its neighbours in KNN order; after every clip the cell
// is certified once the just-clipped neighbour is farther than twice the
// largest seed-vertex distance.  Vertices are plane triples (pairs in 2D).
//
// Layout on the GPU (differs from the reference, results do not):
//   * a wave64 holds SPW = 64/G seeds, G lanes each; a lane owns S vertex
//     slots, slot id gs = s*G + lane.  Per slot the lane keeps in registers
//     the vertex position (px,py,pz), the Cramer denominator W, the squared
//     seed distance d2 and two float bounds (q, m) used by the filter below.
//   * planes (FP64), plane->neighbour ids, vertex plane triples and the
//     vertex adjacency (slot across each edge) live in LDS.
//   * a clip tests every live vertex in parallel.  The reference decides
//     with a 4x4 (3x3 in 2D) FP64 determinant per vertex; here
//     det ~= W * (e . p + e.w) is evaluated from the cached vertex and used
//     only when |det| exceeds a proven bound (48u |e|_1 q + eps bound, see
//     conflict()); otherwise the reference determinant and its eps test are
//     evaluated exactly as written.  Decisions and needs_exact_predicates
//     flags are therefore identical to the reference.
//   * the cavity boundary is found from the adjacency (an edge of a removed
//     vertex whose neighbour is kept), in parallel over removed vertices,
//     instead of the reference's serial walk; new vertices are the same
//     plane triples (cur, a, b) with the same orientation, so vertex
//     positions and the certificate are bit-identical to the reference.
//   * extraction walks each face's vertex cycle through the adjacency; face
//     order is plane order as in the reference (box planes, then KNN order).
#pragma once

#include "common.h"
#include "exact.h"
#include "kernels.h"
#include "simt.h"
#include "subgrid.h"

namespace pg {

struct Plane {
    double x, y, z, w;
};

// Fast kernels (at most 64 vertex slots) hold 48 planes per seed, which
// keeps a seed's LDS below 3.3 KB (three waves per SIMD); a seed that needs
// more planes before max_p, or more vertex slots, is redone by the 128-slot
// kernel, which holds the full 64 planes.
template <int DIM, int G, int S>
struct VoroCfg {
    static constexpr int NS = G * S;               // vertex slots per seed
    static constexpr int SPW = 64 / G;             // seeds per wave (= block)
    static constexpr int MAXP = NS > 64 ? 64 : 48; // plane slots per seed
    static constexpr int NW = (NS + 63) / 64;      // 64-bit words of a slot mask
    static constexpr int PPL = MAXP / G;           // planes per lane at extraction
    // waves per SIMD the register budget is sized for: the 16-lane fast
    // kernels fit three in LDS (under 3.4 KB per seed)
    static constexpr int WAVES = (G == 16 && NS <= 64) ? 3 : 1;
    static_assert(NS <= 128, "slot ids are bytes below 255");
    static_assert(G <= 64 && 64 % G == 0, "groups tile the wave");
    static_assert(MAXP % G == 0, "planes tile the group");
};

constexpr int NB_CHUNK = 12;

template <int NS, int MAXP>
struct SeedSmem {
    union {
        Plane plane[MAXP];      // planes while clipping
        double vpos[3 * NS];    // vertex positions during extraction
    };
    float pn1[MAXP];        // |x|+|y|+|z|+|w| rounded up
    float pmx[MAXP];        // largest coefficient entering the eps test, rounded up
    int32_t pvid[MAXP];     // neighbour (sorted id) of each plane, -1 = box
    uint32_t tri[NS];       // planes of each vertex: a | b << 8 | c << 16
    uint8_t adj[3][NS];     // neighbour slot across edge/line e
    union {
        struct {
            uint32_t edge[NS];        // boundary edges of the current clip
            uint8_t newslot[NS];      // slot of the j-th new vertex
            uint8_t by_start[MAXP];   // 3D: boundary edge starting at a plane
            uint8_t by_end[MAXP];     // 3D: boundary edge ending at a plane
            uint32_t nbi[NB_CHUNK];   // next chunk of the neighbour row (prefetched)
            double nbp[NB_CHUNK][3];  // and the neighbours' positions
        };
        uint32_t first[MAXP];  // extraction: vertex of each plane with the lowest plane key
    };
};

// Round a non-negative double up to float (next float above on inexact).
PG_DEV float f2f_up(double d) {
    float f = (float)d;
    if ((double)f < d) f = __int_as_float(__float_as_int(f) + 1);
    return f;
}
PG_DEV int tri_p(uint32_t t, int e) { return (int)((t >> (8 * e)) & 0xffu); }

// Slot-independent identity of a vertex: its planes, sorted, packed in 24
// bits (16 in 2D).
template <int DIM>
PG_DEV uint32_t vertex_key(uint32_t t) {
    uint32_t a = t & 0xffu, b = (t >> 8) & 0xffu;
    if (DIM == 2) return a < b ? (a << 8) | b : (b << 8) | a;
    uint32_t c = (t >> 16) & 0xffu, x;
    if (a > b) { x = a; a = b; b = x; }
    if (b > c) { x = b; b = c; c = x; }
    if (a > b) { x = a; a = b; b = x; }
    return (a << 16) | (b << 8) | c;
}

template <int NW>
PG_DEV bool mask_has(const uint64_t (&m)[NW], int gs) {
    if (gs >= 64 * NW) return false;
    return (m[gs >> 6] >> (gs & 63)) & 1ull;
}
template <int NW>
PG_DEV int mask_rank(const uint64_t (&m)[NW], int gs) {
    int r = 0;
#pragma unroll
    for (int w = 0; w < NW; w++) {
        if (w < (gs >> 6)) r += popc64(m[w]);
        else if (w == (gs >> 6)) r += popc64(m[w] & ((1ull << (gs & 63)) - 1ull));
    }
    return r;
}

// Vertex of three planes (two lines in 2D) by the reference's Cramer rule
// (cell.cpp:388-411), homogeneous part (X, Y, Z, W).
template <int DIM>
PG_DEV void cramer(const Plane& a, const Plane& b, const Plane& c, double& X, double& Y, double& Z,
                   double& W) {
    if (DIM == 2) {
        X = -det2(a.w, a.y, b.w, b.y);
        Y = -det2(a.x, a.w, b.x, b.w);
        Z = 0.0;
        W = det2(a.x, a.y, b.x, b.y);
    } else {
        X = -det3(a.w, a.y, a.z, b.w, b.y, b.z, c.w, c.y, c.z);
        Y = -det3(a.x, a.w, a.z, b.x, b.w, b.z, c.x, c.w, c.z);
        Z = -det3(a.x, a.y, a.w, b.x, b.y, b.w, c.x, c.y, c.w);
        W = det3(a.x, a.y, a.z, b.x, b.y, b.z, c.x, c.y, c.z);
    }
}

// The reference conflict predicate (cell.cpp:175-221), evaluated exactly.
template <int DIM>
PG_DEV bool conflict_exact(const Plane& p1, const Plane& p2, const Plane& p3, const Plane& e,
                           bool& flag, double xb = 0.0) {
    double det, eps;
    if (DIM == 2) {
        det = det3(p1.x, p2.x, e.x, p1.y, p2.y, e.y, p1.w, p2.w, e.w);
        double mx = fmax(fmax(fabs(p1.x), fabs(p2.x)), fabs(e.x));
        double my = fmax(fmax(fabs(p1.y), fabs(p2.y)), fabs(e.y));
        double mw = fmax(fmax(fabs(p1.w), fabs(p2.w)), fabs(e.w));
        double mm = fmax(fmax(mx, my), mw);
        eps = 1e-14 * mx * my * mw;
        eps *= mm;
    } else {
        det = det4(p1.x, p2.x, p3.x, e.x, p1.y, p2.y, p3.y, e.y, p1.z, p2.z, p3.z, e.z, p1.w, p2.w,
                   p3.w, e.w);
        double mx = fmax(fmax(fmax(fabs(p1.x), fabs(p2.x)), fabs(p3.x)), fabs(e.x));
        double my = fmax(fmax(fmax(fabs(p1.y), fabs(p2.y)), fabs(p3.y)), fabs(e.y));
        double mz = fmax(fmax(fmax(fabs(p1.z), fabs(p2.z)), fabs(p3.z)), fabs(e.z));
        eps = 1e-14 * mx * my * mz;
        double mm = fmax(fmax(mx, my), mz);
        eps *= (mm * mm);
    }
    if (fabs(det) < eps || fabs(det) < xb) flag = true;
    return det > 0.0;
}

// n1: |x| + |y| + |z| + wm rounded up to float, wm = |w| (or, in exact
// mode, the magnitude that bounds the rounding error of w, see XTOL)
template <int DIM>
PG_DEV void plane_bounds(const Plane& p, double wm, float& n1, float& mx) {
    n1 = f2f_up(fabs(p.x) + fabs(p.y) + fabs(p.z) + wm);
    double m = DIM == 2 ? fmax(fmax(fabs(p.x), fabs(p.y)), fabs(p.w))
                        : fmax(fmax(fabs(p.x), fabs(p.y)), fabs(p.z));
    mx = f2f_up(m);
}

// Per-seed register state of one lane.  Per slot: q bounds the product of
// the vertex's plane 1-norms; e5 is the vertex's share of the reference's
// eps, an upper bound (float, at least the smallest normal float) of
// 1.0000001e-14 m^5 (m^4 in 2D), m the largest normal component among its
// planes, made once with the vertex instead of at every conflict test.
template <int S>
struct SlotRegs {
    double px[S], py[S], pz[S], pw[S], d2[S];
    float q[S], e5[S];
    uint32_t live;  // bit s = slot s holds a vertex
};

// 1.0000001e-14 M^5 (M^4 in 2D), the bound of the reference's conflict eps
// (cell.cpp:175-221: 1e-14 times a product of five, or four, component
// maxima each at most M), computed in double.  Rounding of the products is
// monotone, so the bound of max(a, b) is the max of the bounds.
template <int DIM>
PG_DEV double eps_bound(double M) {
    const double M2 = M * M;
    return DIM == 3 ? 1.0000001e-14 * M2 * M2 * M : 1.0000001e-14 * M2 * M2;
}

template <int DIM, int S>
PG_DEV void slot_write(SlotRegs<S>& r, int s, double px, double py, double pz, double pw, double d2,
                       float q, float m) {
    // In float: the constant, 1.000001e-14 as a float (relative error below
    // 6e-8), and at most four products, each within 2^-24 of the exact one,
    // stay above 1.0000001e-14 m^5.  When that bound is a normal float every
    // intermediate is normal too (m^k >= m^5 / 1e-14 for m < 1); a result
    // below the smallest normal float, subnormal or flushed to zero, is
    // raised to it by the integer max on its bit pattern (non-negative).
    const float m2 = m * m;
    const float mp = DIM == 3 ? m2 * m2 * m : m2 * m2;
    const float e5 = __uint_as_float(max(__float_as_uint(mp * 1.000001e-14f), 0x00800000u));
#pragma unroll
    for (int i = 0; i < S; i++) {
        if (i == s) {
            r.px[i] = px;
            r.py[i] = py;
            r.pz[i] = pz;
            r.pw[i] = pw;
            r.d2[i] = d2;
            r.q[i] = q;
            r.e5[i] = e5;
        }
    }
    r.live |= 1u << s;
}

// Build vertex numerics from its planes (reference arithmetic for position
// and squared seed distance, cell.cpp:369-411).
template <int DIM>
PG_DEV void vertex_numerics(const Plane& a, const Plane& b, const Plane& c, double sx, double sy,
                            double sz, double& px, double& py, double& pz, double& pw,
                            double& d2) {
    double X, Y, Z, W;
    cramer<DIM>(a, b, c, X, Y, Z, W);
    px = X / W;
    py = Y / W;
    pz = DIM == 3 ? Z / W : 0.0;
    pw = W;
    double dx = px - sx, dy = py - sy, dz = pz - sz;
    d2 = dx * dx + dy * dy + dz * dz;
}

// Saved clip state (CLIP_STATE_BYTES per seed): planes, their bounds and
// neighbour ids, then per vertex slot its planes (0xffffffff: empty slot)
// and adjacency, then the plane and live-vertex counts.
struct ClipState {
    Plane plane[64];
    float pn1[64], pmx[64];
    int32_t pvid[64];
    uint32_t tri[64];
    uint8_t adj[3][64];
    int32_t nb_v, nlive;
};
static_assert(sizeof(ClipState) <= CLIP_STATE_BYTES, "clip state layout");

// Exact decision of a degenerate conflict test (exact.h) from the vertex's
// plane triple tt, the new plane's point b (sorted id z) and the seed.
template <int DIM, int NS, int MAXP>
__device__ __noinline__ int exact_conflict(uint32_t tt, const SeedSmem<NS, MAXP>& sm, const double* pts,
                                           uint32_t sid, const double* sp, uint32_t z, const double* bp,
                                           const uint32_t* key) {
    XRow rows[DIM + 1];
    for (int r = 0; r < DIM; r++) {
        const int pid = (int)tri_p(tt, r);
        XRow& x = rows[r];
        x.box = pid < 2 * DIM;
        x.p[0] = x.p[1] = x.p[2] = 0.0;
        x.key = 0;
        x.axis = pid >> 1;
        x.sg = (pid & 1) ? -1.0 : 1.0;
        x.c = (pid & 1) ? 1.0 + 1e-14 : -1e-14;  // the box bounds of the initial cell
        if (!x.box) {
            const uint32_t q = (uint32_t)sm.pvid[pid];
            for (int d = 0; d < DIM; d++) x.p[d] = pts[(size_t)DIM * q + d];
            x.key = key ? key[q] : q;
        }
    }
    XRow& e = rows[DIM];
    e.box = false;
    e.axis = 0;
    e.sg = 1.0;
    e.c = 0.0;
    e.p[0] = bp[0];
    e.p[1] = bp[1];
    e.p[2] = bp[2];
    e.key = key ? key[z] : z;
    return conflict_sos<DIM>(sp, key ? key[sid] : sid, rows);
}

// Vertex-ball completion (BALL kernels; no reference counterpart): a seed
// whose certificate no row of k_max neighbours reaches is finished by the
// empty-sphere test.  After the first kb neighbours of its row, while the
// ball of some vertex v (centre v, radius |v - s|) strictly holds a point
// that is not one of the cell's planes, the cell is clipped by the
// candidate nearest the seed ((distance, input index) order) among the
// vertices' candidates, each the point nearest its vertex (smallest input
// index on ties); when no ball holds one, every vertex is a vertex of the
// Voronoi diagram and the cell is certified (kcert = kmark).  A candidate
// stays valid while its vertex lives: a clip by p removes every vertex whose
// ball holds p.  A ball search visits grid rings out to BALL_RINGS; one that
// would need more leaves the seed uncertified.
constexpr int BALL_RINGS = 32;

struct BallGrid {
    int N = 0;                        // grid cells per axis of the sorted points
    const int32_t* starts = nullptr;  // first sorted id per grid cell
    const int32_t* counts = nullptr;  // points per grid cell
    const uint32_t* perm = nullptr;   // input index per sorted id (ties)
    int kmark = 0;                    // kcert of a cell the completion certifies
    int32_t* scr = nullptr;           // LDS, 2 G + 1 ints per seed group (the group search's tables)
    SubGrid sg;                       // sub-grids of the crowded cells (none: sg.m == nullptr)
};

// Lower bound of the squared distance from v to any point of grid cell ix:
// boundary cells are open outwards (points beyond the unit box are clamped
// into them); a slack of 1e-9 cell widths per axis and a factor 1 - 1e-12
// keep it below the rounded distance of every point of the cell, so a cell
// whose bound is at least the best distance so far holds no point that
// would be picked.
template <int DIM>
PG_DEV double ball_cell_lb2(int N, const double (&v)[3], const int (&ix)[3]) {
    const double h = 1.0 / N;
    double s = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        const double lo = ix[d] == 0 ? -1e300 : ix[d] * h;
        const double hi = ix[d] == N - 1 ? 1e300 : (ix[d] + 1) * h;
        const double e = fmax(fmax(lo - v[d], v[d] - hi) - 1e-9 * h, 0.0);
        s += e * e;
    }
    return s * (1.0 - 1e-12);
}

// The point nearest v strictly inside the ball (v, r2) that is neither the
// seed nor a plane of the cell ((distance, input index) order); -1: none;
// -2: beyond BALL_RINGS.  Found by the whole group: the cells of a grid
// ring are taken G at a time, their points spread over the group's lanes
// (prefix sum of the cell counts, a binary search per lane), and the
// group's best is reduced at the end of every ring.  With the bounds of
// ball_cell_lb2 the answer is the argmin over all points, and the ring at
// which the walk stops depends only on the rings before it: the order the
// points are visited in does not matter.  A crowded cell (one with a
// sub-grid) is taken sub-cell by sub-cell, only the sub-cells the ball of
// the best so far reaches (sub_lower_d2 bounds them the same way).  v, r2,
// sid: the same on every lane of the group, every lane of which is active;
// shift / gmask: the group's lanes in a ballot; cst / cpf: the group's LDS
// tables (G and G + 1 entries).
template <int DIM, int G>
PG_DEV int32_t ball_search_group(const BallGrid& bg, const double* __restrict__ pts, const double (&v)[3], double r2,
                                 uint32_t sid, const int32_t* pvid, int nb_v, int gl, int shift, uint64_t gmask,
                                 int32_t* cst, int32_t* cpf) {
    const int N = bg.N;
    const double h = 1.0 / N;
    int cc[3] = {0, 0, 0};
#pragma unroll
    for (int d = 0; d < DIM; d++) cc[d] = max(0, min(N - 1, (int)floor(v[d] * N)));
    double best = r2;
    int32_t pick = -1;
    for (int ring = 0;; ring++) {
        if (ring > 1) {
            // a point of ring r lies more than r - 1 cell widths away
            const double lb = ((ring - 1) - 1e-9) * h;
            if (lb * lb * (1.0 - 1e-12) >= best) break;
        }
        if (ring > N) break;
        if (ring > BALL_RINGS) return -2;
        const int side = 2 * ring + 1;
        const int ncube = DIM == 3 ? side * side * side : side * side;
        double lbest = best;
        int32_t lpick = pick;
        // points 0 .. total-1 of the cells (or sub-cells) in cst / cpf
        // (first position, exclusive prefix), G at a time; positions are
        // sorted ids, or indices into indir
        auto scan = [&](int total, const uint32_t* __restrict__ indir) {
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
            for (int cb = 0; cb < total; cb += G) {
                const int j = cb + gl;
                if (j >= total) continue;
                // cell holding point j: largest a with cpf[a] <= j
                int lo = 0, hi = G - 1;
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (cpf[mid] <= j) lo = mid;
                    else hi = mid - 1;
                }
                int32_t q = cst[lo] + (j - cpf[lo]);
                if (indir != nullptr) q = (int32_t)indir[q];
                if ((uint32_t)q == sid) continue;
                double d2 = 0.0;
#pragma unroll
                for (int d = 0; d < DIM; d++) {
                    const double df = pts[(size_t)DIM * q + d] - v[d];
                    d2 += df * df;
                }
                // strictly inside; on ties the smaller input index
                if (d2 > lbest || (d2 == lbest && (lpick < 0 || bg.perm[q] > bg.perm[lpick]))) continue;
                bool plane = false;
                for (int pl = 0; pl < nb_v && !plane; pl++) plane = pvid[pl] == q;
                if (plane) continue;
                lbest = d2;
                lpick = q;
            }
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
        };
        for (int base = 0; base < ncube; base += G) {
            const int t = base + gl;
            int32_t st = 0, cnt = 0;
            bool crowded = false;  // a cell with a sub-grid: taken sub-cell by sub-cell below
            int ix[3] = {0, 0, 0};
            if (t < ncube) {
                const int di = t % side - ring, dj = (t / side) % side - ring;
                const int dk = DIM == 3 ? t / (side * side) - ring : 0;
                const int cheb = max(abs(di), max(abs(dj), abs(dk)));
                ix[0] = cc[0] + di;
                ix[1] = cc[1] + dj;
                ix[2] = cc[2] + dk;
                if (cheb == ring && ix[0] >= 0 && ix[1] >= 0 && ix[2] >= 0 && ix[0] < N && ix[1] < N &&
                    (DIM == 2 || ix[2] < N) && ball_cell_lb2<DIM>(N, v, ix) < best) {
                    const int64_t cell = ix[0] + (int64_t)ix[1] * N + (int64_t)ix[2] * N * N;
                    if (bg.sg.m != nullptr && bg.sg.m[cell] > 0) {
                        crowded = true;
                    } else {
                        st = bg.starts[cell];
                        cnt = bg.counts[cell];
                    }
                }
            }
            const int incl = group_incl_sum<G>(cnt, gl);
            const int total = group_last<G>(incl);
            if (total > 0) {
                cst[gl] = st;
                cpf[gl + 1] = incl;
                if (gl == 0) cpf[0] = 0;
                scan(total, nullptr);
            }
            // the crowded cells of this batch, one at a time: their
            // sub-cells that the ball of the ring-start best can reach
            uint64_t crowd = (wave_ballot(crowded) >> shift) & gmask;
            while (crowd != 0ull) {
                const int a = ffs64(crowd);
                crowd &= crowd - 1ull;
                const int cx[3] = {shfl(ix[0], a, G), shfl(ix[1], a, G), shfl(ix[2], a, G)};
                const int64_t cell = cx[0] + (int64_t)cx[1] * N + (int64_t)cx[2] * N * N;
                const int m = bg.sg.m[cell];
                const int64_t sbase = bg.sg.ptr[cell];
                const double P[3] = {v[0] * N, v[1] * N, DIM == 3 ? v[2] * N : 0.0};
                int lo[3], hi[3];
                sub_box<DIM>(P, cx, m, N, best, lo, hi);
                const int nx = hi[0] - lo[0] + 1, ny = hi[1] - lo[1] + 1, nz = hi[2] - lo[2] + 1;
                const int nbox = nx * ny * nz;
                for (int b0 = 0; b0 < nbox; b0 += G) {
                    const int u = b0 + gl;
                    int32_t sst = 0, scnt = 0;
                    if (u < nbox) {
                        const int sx[3] = {lo[0] + u % nx, lo[1] + (u / nx) % ny, lo[2] + u / (nx * ny)};
                        if (sub_lower_d2<DIM>(P, cx, sx, m, N) < best) {
                            const int64_t j = sbase + sx[0] + (int64_t)m * (sx[1] + (int64_t)m * sx[2]);
                            sst = bg.sg.start[j];
                            scnt = bg.sg.cnt[j];
                        }
                    }
                    const int sincl = group_incl_sum<G>(scnt, gl);
                    const int stotal = group_last<G>(sincl);
                    if (stotal == 0) continue;
                    cst[gl] = sst;
                    cpf[gl + 1] = sincl;
                    if (gl == 0) cpf[0] = 0;
                    scan(stotal, bg.sg.idx);
                }
            }
        }
        // the group's best after this ring
        best = lbest;
        pick = lpick;
#pragma unroll
        for (int d = 1; d < G; d <<= 1) {
            const double od = shfl_xor(best, d, G);
            const int32_t oq = shfl_xor(pick, d, G);
            if (oq >= 0 && (pick < 0 || od < best || (od == best && bg.perm[oq] < bg.perm[pick]))) {
                best = od;
                pick = oq;
            }
        }
    }
    return pick;
}

// Exact mode (fa.key set, or EXACT): a conflict test is trusted in floating
// point only above XTOL times the product of its rows' 1-norms (w measured
// by the magnitude its rounding is relative to), a valid bound on how far the
// rounding of the planes and of the determinant can move it; below, EXACT
// kernels decide it with exact_conflict and the others flag the cell
// (needs_exact_predicates) for them.  The certificate of EXACT kernels keeps
// a relative margin (CERT_TIE): a neighbour tied with twice the largest
// vertex distance can still cut a zero-area face under the symbolic
// perturbation, so only neighbours clearly beyond it certify the cell; the
// other kernels flag a near tie.
// BALL: the row is clipped whole, then the vertex-ball completion above
// certifies the cell.
template <int DIM, int G, int S, bool EXACT = false, bool BALL = false>
__device__ void voronoi_group(const double* __restrict__ pts, int64_t n_total, const uint32_t* __restrict__ seeds,
                              int64_t nq, const uint32_t* __restrict__ knn, int k, int max_p,
                              int max_t, FaceArrays fa,
                              SeedSmem<VoroCfg<DIM, G, S>::NS, VoroCfg<DIM, G, S>::MAXP>* smem,
                              int64_t vblock, BallGrid bg = {}) {
    using C = VoroCfg<DIM, G, S>;
    constexpr int NS = C::NS, NW = C::NW, MAXP = C::MAXP;
    constexpr int D = DIM;
    const int tid = (int)threadIdx.x;
    const int grp = tid / G, gl = tid % G;
    const int shift = grp * G;
    const uint64_t gmask = (G == 64) ? ~0ull : ((1ull << G) - 1ull);
    const int64_t qi = vblock * C::SPW + grp;
    const bool valid = qi < nq;
    SeedSmem<NS, MAXP>& sm = smem[grp];
    if (max_p > FACE_SLOTS) max_p = FACE_SLOTS;
    // exact mode (a build with the exact-predicate fallback): every test
    // whose sign the rounding of the planes could flip is decided exactly
    // (EXACT kernels) or flags the cell for them (the others)
    const bool xmode = EXACT || fa.key != nullptr;

    const uint32_t sid = valid ? (seeds ? seeds[qi] : (uint32_t)qi) : 0u;
    const uint32_t* row = knn + (valid ? (size_t)qi * (size_t)k : 0);
    const double sx = pts[(size_t)D * sid], sy = pts[(size_t)D * sid + 1];
    const double sz = D == 3 ? pts[(size_t)D * sid + 2] : 0.0;

    // ---------------------------------------------------------------- init
    // A seed whose previous pass (same shape, shorter row) stopped
    // uncertified resumes from its saved cell at neighbour v0: clipping the
    // same planes in the same order is deterministic, so the cell equals a
    // pass over the longer row from scratch.
    const int32_t shape_tag = (G << 24) | (S << 16);
    int v0 = 0;
    if (valid && fa.resume != nullptr && NS <= 64) {
        const int32_t r = fa.resume[sid];
        if (r > 0 && (r & 0x7fff0000) == shape_tag && (r & 0xffff) < k) v0 = r & 0xffff;
    }
    const ClipState* cst = v0 ? reinterpret_cast<const ClipState*>(fa.cstate + (size_t)sid * CLIP_STATE_BYTES)
                              : nullptr;
    int nb_v0 = 2 * D, nlive0 = D == 3 ? 8 : 4;
    if (v0) {
        nb_v0 = cst->nb_v;
        nlive0 = cst->nlive;
        for (int p = gl; p < nb_v0; p += G) {
            sm.plane[p] = cst->plane[p];
            sm.pn1[p] = cst->pn1[p];
            sm.pmx[p] = cst->pmx[p];
            sm.pvid[p] = cst->pvid[p];
        }
#pragma unroll
        for (int i = 0; i < S; i++) {
            const int gs = i * G + gl;
            sm.tri[gs] = cst->tri[gs];
#pragma unroll
            for (int e = 0; e < D; e++) sm.adj[e][gs] = cst->adj[e][gs];
        }
    }
    // Box planes (cell.cpp:44-82): eps = 1e-14 margin.
    if (!v0 && gl < 2 * D) {
        const double eps = 1e-14, lo = -eps, hi = 1.0 + eps;
        Plane p;
        switch (gl) {
            case 0: p = {1.0, 0.0, 0.0, -lo}; break;
            case 1: p = {-1.0, 0.0, 0.0, hi}; break;
            case 2: p = {0.0, 1.0, 0.0, -lo}; break;
            case 3: p = {0.0, -1.0, 0.0, hi}; break;
            case 4: p = {0.0, 0.0, 1.0, -lo}; break;
            default: p = {0.0, 0.0, -1.0, hi}; break;
        }
        sm.plane[gl] = p;
        plane_bounds<D>(p, fabs(p.w), sm.pn1[gl], sm.pmx[gl]);
        sm.pvid[gl] = -1;
    }
    // Initial vertices (cell.cpp:84-103).
    constexpr int NV0 = D == 3 ? 8 : 4;
    if (!v0 && gl < NV0) {
        uint32_t t;
        if (D == 3) {
            const uint32_t T3[8] = {2 | 5 << 8 | 0 << 16, 5 | 3 << 8 | 0 << 16, 1 | 5 << 8 | 2 << 16,
                                    5 | 1 << 8 | 3 << 16, 4 | 2 << 8 | 0 << 16, 4 | 0 << 8 | 3 << 16,
                                    2 | 4 << 8 | 1 << 16, 4 | 3 << 8 | 1 << 16};
            t = T3[gl];
        } else {
            const uint32_t T2[4] = {2 | 0 << 8, 1 | 2 << 8, 3 | 1 << 8, 0 | 3 << 8};
            t = T2[gl];
        }
        sm.tri[gl] = t;
    }
    block_sync();
    SlotRegs<S> R;
    R.live = 0;
#pragma unroll
    for (int s = 0; s < S; s++) {
        R.px[s] = R.py[s] = R.pz[s] = R.pw[s] = R.d2[s] = 0.0;
        R.q[s] = R.e5[s] = 0.f;
    }
    // resumed cell: vertex numerics from the saved planes, with the same
    // arithmetic (and operand order) as when the vertices were made
    if (v0) {
#pragma unroll
        for (int i = 0; i < S; i++) {
            const uint32_t t = sm.tri[i * G + gl];
            if (t == 0xffffffffu) continue;
            const Plane& a = sm.plane[tri_p(t, 0)];
            const Plane& b = sm.plane[tri_p(t, 1)];
            const Plane& c = sm.plane[D == 3 ? tri_p(t, 2) : tri_p(t, 1)];
            double px, py, pz, pw, d2;
            vertex_numerics<D>(a, b, c, sx, sy, sz, px, py, pz, pw, d2);
            float q = D == 3 ? sm.pn1[tri_p(t, 0)] * sm.pn1[tri_p(t, 1)] * sm.pn1[tri_p(t, 2)]
                             : sm.pn1[tri_p(t, 0)] * sm.pn1[tri_p(t, 1)];
            float m = fmaxf(sm.pmx[tri_p(t, 0)], sm.pmx[tri_p(t, 1)]);
            if (D == 3) m = fmaxf(m, sm.pmx[tri_p(t, 2)]);
            slot_write<D, S>(R, i, px, py, pz, pw, d2, q * 1.000001f, m);
        }
    }
    // NV0 <= G: initial vertex j lives in slot 0 of lane j.
    if (!v0 && gl < NV0) {
        uint32_t t = sm.tri[gl];
        const Plane& a = sm.plane[tri_p(t, 0)];
        const Plane& b = sm.plane[tri_p(t, 1)];
        const Plane& c = sm.plane[D == 3 ? tri_p(t, 2) : tri_p(t, 1)];
        double px, py, pz, pw, d2;
        vertex_numerics<D>(a, b, c, sx, sy, sz, px, py, pz, pw, d2);
        float q = D == 3 ? sm.pn1[tri_p(t, 0)] * sm.pn1[tri_p(t, 1)] * sm.pn1[tri_p(t, 2)]
                         : sm.pn1[tri_p(t, 0)] * sm.pn1[tri_p(t, 1)];
        float m = fmaxf(sm.pmx[tri_p(t, 0)], sm.pmx[tri_p(t, 1)]);
        if (D == 3) m = fmaxf(m, sm.pmx[tri_p(t, 2)]);
        slot_write<D, S>(R, 0, px, py, pz, pw, d2, q * 1.000001f, m);
        // adjacency of the initial polytope
        for (int e = 0; e < D; e++) {
            uint8_t nb = END;
            if (D == 3) {
                int a0 = tri_p(t, e), b0 = tri_p(t, (e + 1) % 3);
                for (int j = 0; j < NV0; j++) {
                    if (j == gl) continue;
                    uint32_t u = sm.tri[j];
                    for (int f = 0; f < 3; f++)
                        if (tri_p(u, f) == b0 && tri_p(u, (f + 1) % 3) == a0) nb = (uint8_t)j;
                }
            } else {
                int L = tri_p(t, e);
                for (int j = 0; j < NV0; j++) {
                    if (j == gl) continue;
                    uint32_t u = sm.tri[j];
                    if (tri_p(u, 0) == L || tri_p(u, 1) == L) nb = (uint8_t)j;
                }
            }
            sm.adj[e][gl] = nb;
        }
    }
    block_sync();

    int nb_v = nb_v0;
    int nlive = nlive0;
    int status = SUCCESS;
    int kc = 0;
    bool done = !valid;
    double R2 = 0.0;  // largest squared seed-vertex distance of the cell
#pragma unroll
    for (int s = 0; s < S; s++)
        if ((R.live >> s) & 1u) R2 = smax(R.d2[s], R2);
    R2 = group_butterfly_d<G>(R2, [](double a, double b) { return smax(a, b); });

    // One clip of the cell by the bisector of neighbour z at (bx, by, bz)
    // (act: this group clips in this step; every lane of the wave calls it):
    // new plane, conflict test of every live vertex, cavity, consistency and
    // new vertices.  Returns whether any seed of the wave clipped.
    auto clip_step = [&](bool act, uint32_t z, double bx, double by, double bz) -> bool {
            // -- new half plane (cell.cpp:158-172)
            Plane e = {0, 0, 0, 0};
            double ewm = 0.0;  // magnitude of e.w for the bounds (see plane_bounds)
            bool plane_ok = false;
            int cur = 0;
            if (act) {
                if (nb_v >= max_p) {
                    status = VERTEX_OVERFLOW;
                } else if (nb_v >= MAXP) {
                    status = SLOT_OVERFLOW;  // more planes than this kernel holds
                } else {
                    double dx = sx - bx, dy = sy - by, dz = sz - bz;
                    double ax = sx + bx, ay = sy + by, az = sz + bz;
                    double dot = ax * dx + ay * dy + az * dz;
                    e = {dx, dy, dz, -dot / 2.0};
                    ewm = xmode ? 0.5 * (fabs(ax * dx) + fabs(ay * dy) + fabs(az * dz)) : fabs(e.w);
                    cur = nb_v;
                    plane_ok = true;
                    if (gl == 0) {
                        sm.plane[cur] = e;
                        plane_bounds<D>(e, ewm, sm.pn1[cur], sm.pmx[cur]);
                        sm.pvid[cur] = (int32_t)z;
                    }
                    nb_v++;
                }
            }
            // -- conflict test of every live vertex
            uint32_t rem = 0;
            bool flag = false;
            if (act && plane_ok) {
                const double en1 = (fabs(e.x) + fabs(e.y) + fabs(e.z) + ewm) * (1.0 + 1e-15);
                const double emx = D == 2 ? fmax(fmax(fabs(e.x), fabs(e.y)), fabs(e.w))
                                          : fmax(fmax(fabs(e.x), fabs(e.y)), fabs(e.z));
                // per plane: the eps bound of its own components, and the
                // rounding bound's factor (exact mode: below XTOL en1 q the
                // rounding of the planes themselves may flip the sign, which
                // the reference's eps does not bound; such tests go to the
                // exact predicate)
                const uint32_t e_eps = __float_as_uint(fmaxf(f2f_up(eps_bound<D>(emx)), 1.17549435e-38f));
                const double c1 = (5.4e-15 + (xmode ? XTOL : 0.0)) * en1;
    #pragma unroll
                for (int s = 0; s < S; s++) {
                    if (!((R.live >> s) & 1u)) continue;
                    double t;
                    if (D == 3)
                        t = R.pw[s] * fma(e.x, R.px[s], fma(e.y, R.py[s], fma(e.z, R.pz[s], e.w)));
                    else
                        t = R.pw[s] * fma(e.x, R.px[s], fma(e.y, R.py[s], e.w));
                    // the larger of two non-negative floats by their bit
                    // patterns (ordered like the values)
                    const double epsu = (double)__uint_as_float(max(__float_as_uint(R.e5[s]), e_eps));
                    const double bnd = fma(c1, (double)R.q[s], epsu);
                    bool c;
                    if (fabs(t) > bnd) {
                        c = t > 0.0;
                    } else {
                        const double xb = xmode ? XTOL * en1 * (double)R.q[s] : 0.0;
                        uint32_t tt = sm.tri[s * G + gl];
                        const Plane p1 = sm.plane[tri_p(tt, 0)];
                        const Plane p2 = sm.plane[tri_p(tt, 1)];
                        const Plane p3 = sm.plane[D == 3 ? tri_p(tt, 2) : tri_p(tt, 1)];
                        if (EXACT) {
                            bool fl = false;
                            c = conflict_exact<D>(p1, p2, p3, e, fl, xb);
                            if (fl) {
                                const double sp[3] = {sx, sy, sz}, bp[3] = {bx, by, bz};
                                const int r = exact_conflict<D>(tt, sm, pts, sid, sp, z, bp, fa.key);
                                if (r == 0) flag = true;  // outside the exact predicate's range
                                else c = r > 0;
                            }
                        } else {
                            c = conflict_exact<D>(p1, p2, p3, e, flag, xb);
                        }
                    }
                    if (c) rem |= 1u << s;
                }
            }
            // -- group masks
            uint64_t Rm[NW], Lm[NW];
    #pragma unroll
            for (int w = 0; w < NW; w++) Rm[w] = Lm[w] = 0ull;
    #pragma unroll
            for (int s = 0; s < S; s++) {
                uint64_t b = (wave_ballot((rem >> s) & 1u) >> shift) & gmask;
                int pos = s * G;
                Rm[pos >> 6] |= b << (pos & 63);
            }
            const bool gflag = ((wave_ballot(flag) >> shift) & gmask) != 0ull;
            int nb_r = 0;
    #pragma unroll
            for (int w = 0; w < NW; w++) nb_r += popc64(Rm[w]);
            bool doclip = false;
            if (act && plane_ok) {
                R.live &= ~rem;  // removed vertices leave the cell (also on failure)
                if (gflag) {
                    status = NEEDS_EXACT_PREDICATES;
                } else if (nb_r == 0) {
                    nb_v--;  // the plane clipped nothing (cell.cpp:132-136)
                } else {
                    doclip = true;
                }
            }
            // Most clip steps clip nothing (about 30 % of a uniform 3D seed's
            // steps clip, and in half of a wave's steps none of its seeds does):
            // the cavity, consistency and new-vertex machinery runs only when a
            // seed of the wave clips (a wave-uniform branch; a block is one wave).
            const bool anyclip = wave_ballot(doclip) != 0ull;
            if (anyclip) {
                // -- cavity boundary: edges of removed vertices whose neighbour is
                //    kept; bit 3 s + e of bm = edge e of my slot s (a lane walks only
                //    its removed slots and then only its boundary edges)
                uint32_t bm = 0;
                if (doclip) {
                    uint32_t rr = rem;
                    while (rr) {
                        const int s = __builtin_ctz(rr);
                        rr &= rr - 1u;
                        const int gs = s * G + gl;
    #pragma unroll
                        for (int ee = 0; ee < D; ee++)
                            if (!mask_has<NW>(Rm, sm.adj[ee][gs])) bm |= 1u << (3 * s + ee);
                    }
                }
                const int mycnt = __builtin_popcount(bm);
                const int incl = group_incl_sum<G>(mycnt, gl);
                const int Bt = group_last<G>(incl);
                int j = incl - mycnt;
                if (doclip) {
                    // in (slot, edge) order, as counted
                    uint32_t b2 = bm;
                    while (b2) {
                        const int bit = __builtin_ctz(b2);
                        b2 &= b2 - 1u;
                        if (j >= NS) {  // more boundary edges than slots: overflow below
                            j++;
                            continue;
                        }
                        const int s = bit / 3, ee = bit - 3 * s;
                        const int gs = s * G + gl;
                        const uint32_t t = sm.tri[gs];
                        const int K = sm.adj[ee][gs];
                        if (D == 3) {
                            const int a = tri_p(t, ee), b = tri_p(t, ee == 2 ? 0 : ee + 1);
                            sm.edge[j] = (uint32_t)a | (uint32_t)b << 8 | (uint32_t)K << 16;
                            sm.by_start[a] = (uint8_t)j;
                            sm.by_end[b] = (uint8_t)j;
                        } else {
                            const int L = tri_p(t, ee);
                            sm.edge[j] = (uint32_t)L | (uint32_t)K << 16;
                        }
                        j++;
                    }
                }
                // live mask of the group (after removal)
    #pragma unroll
                for (int s = 0; s < S; s++) {
                    uint64_t b = (wave_ballot((R.live >> s) & 1u) >> shift) & gmask;
                    int pos = s * G;
                    Lm[pos >> 6] |= b << (pos & 63);
                }
                block_sync();
                // -- consistency and capacity
                bool ok = doclip;
                bool bad = false;
                int nxt = gl, lab = 0;  // 3D: successor edge and orbit label of edge gl
                if (doclip) {
                    if (D == 3) {
                        // every boundary edge owns its start and end plane, and the
                        // edge continuing it (the one starting at its end plane)
                        // exists: the successor map is a permutation of the Bt edges
                        for (int jj = gl; jj < Bt; jj += G) {
                            uint32_t ed = sm.edge[jj];
                            int a = ed & 0xff, b = (ed >> 8) & 0xff;
                            if (sm.by_start[a] != jj || sm.by_end[b] != jj) bad = true;
                            const int nx = sm.by_start[b];
                            if (nx >= Bt || (int)(sm.edge[nx] & 0xffu) != b) bad = true;
                        }
                        if (Bt <= G && gl < Bt) {
                            lab = gl;
                            if (!bad) nxt = sm.by_start[(sm.edge[gl] >> 8) & 0xffu];
                        }
                    } else {
                        bad = (Bt != 2);
                    }
                }
                if (D == 3) {
                    // ... and one cycle through all Bt edges: the smallest edge index
                    // on each orbit, by pointer jumping over the group's lanes (all
                    // groups of the wave take part; lanes without an edge point to
                    // themselves), must be 0 everywhere
                    if (nxt >= G) nxt = gl;  // only after a failed check above
    #pragma unroll
                    for (int r = 1; r < 16 && r < G; r <<= 1) {
                        const int l2 = shfl(lab, nxt, G);
                        nxt = shfl(nxt, nxt, G);
                        lab = min(lab, l2);
                    }
                    if (doclip) {
                        if (Bt <= G) {
                            if (lab != 0) bad = true;
                        } else if (gl == 0 && !bad) {
                            int jj = 0, steps = 0;
                            do {
                                jj = sm.by_start[(sm.edge[jj] >> 8) & 0xffu];
                                steps++;
                            } while (jj != 0 && jj < Bt && steps <= Bt);
                            if (steps != Bt || jj != 0) bad = true;
                        }
                    }
                }
                bad = ((wave_ballot(bad) >> shift) & gmask) != 0ull;
                if (doclip) {
                    const int nl_after = nlive - nb_r + Bt;
                    if (bad) {
                        status = INCONSISTENT_BOUNDARY;
                        ok = false;
                    } else if (nl_after >= max_t) {
                        status = TRIANGLE_OVERFLOW;  // cell.cpp:347-350
                        ok = false;
                    } else if (nl_after > NS) {
                        status = SLOT_OVERFLOW;
                        ok = false;
                    }
                    nlive -= nb_r;
                }
                // -- slots of the new vertices: the j-th boundary edge takes the
                //    j-th free slot (slot-id order)
                uint32_t asg = 0;
                if (ok) {
    #pragma unroll
                    for (int s = 0; s < S; s++) {
                        const int gs = s * G + gl;
                        if ((R.live >> s) & 1u) continue;
                        // free-slot rank = gs - (live slots below gs)
                        int r = gs - mask_rank<NW>(Lm, gs);
                        if (r < Bt) {
                            sm.newslot[r] = (uint8_t)gs;
                            asg |= 1u << s;
                        }
                    }
                }
                block_sync();
                // -- new vertices (cell.cpp:346-364 orientation, :388-411 position)
                if (ok) {
                    float cn1, cmx;
                    plane_bounds<D>(e, ewm, cn1, cmx);
                    while (asg) {
                        const int s = ffs64(asg);
                        asg &= asg - 1;
                        const int gs = s * G + gl;
                        const int r = gs - mask_rank<NW>(Lm, gs);
                        const uint32_t ed = sm.edge[r];
                        uint32_t t;
                        Plane pa, pb;
                        float q, m;
                        if (D == 3) {
                            const int a = ed & 0xff, b = (ed >> 8) & 0xff, K = (ed >> 16) & 0xff;
                            t = (uint32_t)cur | (uint32_t)a << 8 | (uint32_t)b << 16;
                            pa = sm.plane[a];
                            pb = sm.plane[b];
                            q = cn1 * sm.pn1[a] * sm.pn1[b] * 1.000001f;
                            m = fmaxf(cmx, fmaxf(sm.pmx[a], sm.pmx[b]));
                            sm.tri[gs] = t;
                            sm.adj[0][gs] = sm.newslot[sm.by_end[a]];
                            sm.adj[1][gs] = (uint8_t)K;
                            sm.adj[2][gs] = sm.newslot[sm.by_start[b]];
                            const uint32_t tk = sm.tri[K];
                            for (int f = 0; f < 3; f++)
                                if (tri_p(tk, f) == b && tri_p(tk, (f + 1) % 3) == a) sm.adj[f][K] = (uint8_t)gs;
                            double px, py, pz, pw, d2;
                            vertex_numerics<D>(e, pa, pb, sx, sy, sz, px, py, pz, pw, d2);
                            slot_write<D, S>(R, s, px, py, pz, pw, d2, q, m);
                        } else {
                            const int L = ed & 0xff, K = (ed >> 16) & 0xff;
                            pa = sm.plane[L];
                            const double rw = det2(e.x, e.y, pa.x, pa.y);
                            const bool swap = rw > 0;
                            t = swap ? ((uint32_t)L | (uint32_t)cur << 8) : ((uint32_t)cur | (uint32_t)L << 8);
                            const uint8_t other = sm.newslot[1 - r];
                            sm.tri[gs] = t;
                            sm.adj[0][gs] = swap ? (uint8_t)K : other;
                            sm.adj[1][gs] = swap ? other : (uint8_t)K;
                            const uint32_t tk = sm.tri[K];
                            for (int f = 0; f < 2; f++)
                                if (tri_p(tk, f) == L) sm.adj[f][K] = (uint8_t)gs;
                            q = cn1 * sm.pn1[L] * 1.000001f;
                            m = fmaxf(cmx, sm.pmx[L]);
                            double px, py, pz, pw, d2;
                            if (swap)
                                vertex_numerics<D>(pa, e, e, sx, sy, sz, px, py, pz, pw, d2);
                            else
                                vertex_numerics<D>(e, pa, pa, sx, sy, sz, px, py, pz, pw, d2);
                            slot_write<D, S>(R, s, px, py, pz, pw, d2, q, m);
                        }
                    }
                    nlive += Bt;
                }
                block_sync();
            }
        return anyclip;
    };

    // --------------------------------------------------------- clip loop
    // Seeds of a wave may start at different neighbours (resumed cells):
    // the wave walks v from the smallest start, a group joins at its own.
    int vstart = v0;
#pragma unroll
    for (int d = G; d < 64; d <<= 1) vstart = min(vstart, __shfl_xor(vstart, d, 64));
    int cbase = -NB_CHUNK;  // first neighbour of the prefetched chunk
    for (int v = vstart; v < k; v++) {
        if (wave_ballot(!done) == 0ull) break;
        const bool act = !done && v >= v0;
        // the group prefetches the next NB_CHUNK neighbours and their
        // positions into LDS, so the clip loop issues no dependent global loads
        if (v >= cbase + NB_CHUNK) {
            cbase = v;
            for (int j = gl; j < NB_CHUNK; j += G) {
                // a resumed group reads only its row from v0 on: the entries
                // before are not written by a KNN launch that resumes too
                if (!done && v + j < k && v + j >= v0) {
                    const uint32_t zz = row[v + j];
                    sm.nbi[j] = zz;
#pragma unroll
                    for (int d = 0; d < D; d++) sm.nbp[j][d] = pts[(size_t)D * zz + d];
                }
            }
            block_sync();
        }
        // the step's neighbour: the prefetched row entry
        uint32_t z = 0;
        double bx = 0, by = 0, bz = 0;
        if (act) {
            const int jj = v - cbase;
            z = sm.nbi[jj];
            bx = sm.nbp[jj][0];
            by = sm.nbp[jj][1];
            bz = D == 3 ? sm.nbp[jj][2] : 0.0;
        }
        const bool anyclip = clip_step(act, z, bx, by, bz);
        // -- certificate (cell.cpp:369-384) with the just-clipped neighbour;
        //    the largest seed-vertex distance only changes with a clip
        double mx = R2;
        if (anyclip) {
            mx = 0.0;
#pragma unroll
            for (int s = 0; s < S; s++)
                if ((R.live >> s) & 1u) mx = smax(R.d2[s], mx);
            mx = group_butterfly_d<G>(mx, [](double a, double b) { return smax(a, b); });
        }
        if (act) {
            R2 = mx;
            const double dx = bx - sx, dy = by - sy, dz = bz - sz;
            const double dz2 = dx * dx + dy * dy + dz * dz;
            if (!BALL && dz2 > 4 * R2 * (EXACT ? CERT_TIE : 1.0)) {
                if (!EXACT && xmode && status == SUCCESS && !(dz2 > 4 * R2 * CERT_TIE)) {
                    status = NEEDS_EXACT_PREDICATES;  // a near tie: the exact kernel decides
                } else {
                    kc = v + 1;
                    done = true;
                }
            }
            if (status != SUCCESS) done = true;
        }
    }

    // -- vertex-ball completion
    const int first_ball = nb_v;
    if (BALL) {
        bool bdone = !valid || status != SUCCESS;
        int32_t cand[S];
        uint32_t ctri[S];
        uint32_t known = 0;  // bit s: cand[s] belongs to the vertex now in slot s
#pragma unroll
        for (int s = 0; s < S; s++) {
            cand[s] = -1;
            ctri[s] = 0xffffffffu;
        }
        for (int iter = 0;; iter++) {
            if (wave_ballot(!bdone) == 0ull) break;
            // candidates of the live vertices not searched yet
            double bd = 0.0;
            int32_t bq = -1;
            bool far = false;
            // the group searches its vertices one at a time, slot by slot
            // (the wave walks until every group has searched its own)
            int32_t* const cst = bg.scr + grp * (2 * G + 1);
            int32_t* const cpf = cst + G;
#pragma unroll
            for (int s = 0; s < S; s++) {
                const uint32_t t = sm.tri[s * G + gl];
                const bool need = !bdone && ((R.live >> s) & 1u) && (!((known >> s) & 1u) || ctri[s] != t);
                uint64_t m = (wave_ballot(need) >> shift) & gmask;
                while (wave_ballot(m != 0ull) != 0ull) {
                    if (m == 0ull) continue;
                    const int src = ffs64(m);
                    m &= m - 1ull;
                    const double vv[3] = {shfl(R.px[s], src, G), shfl(R.py[s], src, G), shfl(R.pz[s], src, G)};
                    const double r2v = shfl(R.d2[s], src, G);
                    const int32_t cq =
                        ball_search_group<D, G>(bg, pts, vv, r2v, sid, sm.pvid, nb_v, gl, shift, gmask, cst, cpf);
                    if (gl == src) cand[s] = cq;
                }
                if (need) {
                    ctri[s] = t;
                    known |= 1u << s;
                }
            }
            if (!bdone) {
#pragma unroll
                for (int s = 0; s < S; s++) {
                    if (!((R.live >> s) & 1u)) continue;
                    if (cand[s] == -2) far = true;
                    if (cand[s] < 0) continue;
                    const int32_t q = cand[s];
                    const double dx = pts[(size_t)D * q] - sx, dy = pts[(size_t)D * q + 1] - sy;
                    const double dz = D == 3 ? pts[(size_t)D * q + 2] - sz : 0.0;
                    const double d2 = dx * dx + dy * dy + dz * dz;
                    if (bq < 0 || d2 < bd || (d2 == bd && bg.perm[q] < bg.perm[bq])) {
                        bd = d2;
                        bq = q;
                    }
                }
            }
            // the group's candidate nearest the seed
#pragma unroll
            for (int d = 1; d < G; d <<= 1) {
                const double od = shfl_xor(bd, d, G);
                const int32_t oq = shfl_xor(bq, d, G);
                if (oq >= 0 && (bq < 0 || od < bd || (od == bd && bg.perm[oq] < bg.perm[bq]))) {
                    bd = od;
                    bq = oq;
                }
            }
            far = ((wave_ballot(far) >> shift) & gmask) != 0ull;
            bool act = false;
            if (!bdone) {
                if (far) {
                    status = SECURITY_RADIUS_NOT_REACHED;  // a ball beyond BALL_RINGS: gives up
                    bdone = true;
                } else if (bq < 0) {
                    kc = bg.kmark;  // every vertex ball is empty
                    bdone = true;
                } else if (iter >= 4 * max_p) {
                    status = SECURITY_RADIUS_NOT_REACHED;
                    bdone = true;
                } else {
                    act = true;
                }
            }
            const uint32_t zb = act ? (uint32_t)bq : 0u;
            const double bx = act ? pts[(size_t)D * zb] : 0.0, by = act ? pts[(size_t)D * zb + 1] : 0.0;
            const double bz = (act && D == 3) ? pts[(size_t)D * zb + 2] : 0.0;
            clip_step(act, zb, bx, by, bz);
            if (act && status != SUCCESS) bdone = true;
        }
        if (valid && status != SUCCESS) kc = 0;
    }

    // final certificate with the k-th neighbour (voronoi.cpp:145-147); a row
    // holding every other seed certifies the cell by exhaustion
    // (a slot overflow keeps its code: the 128-slot kernel redoes the seed
    // at the same row length and applies this test itself)
    const int loop_status = status;
    const uint32_t zl = valid ? row[k - 1] : 0u;
    if (valid && !BALL) {
        const double dx = pts[(size_t)D * zl] - sx, dy = pts[(size_t)D * zl + 1] - sy;
        const double dz = D == 3 ? pts[(size_t)D * zl + 2] - sz : 0.0;
        const double dl2 = dx * dx + dy * dy + dz * dz;
        if (!EXACT && xmode && status == SUCCESS && dl2 > 4 * R2 && !(dl2 > 4 * R2 * CERT_TIE))
            status = NEEDS_EXACT_PREDICATES;
        if (!(dl2 > 4 * R2 * (EXACT ? CERT_TIE : 1.0))) {
            if ((int64_t)k >= n_total - 1 && status == SUCCESS)
                kc = k;
            else if (status != SLOT_OVERFLOW)
                status = SECURITY_RADIUS_NOT_REACHED;
        }
        if (status != SUCCESS) kc = 0;
    }

    // save an uncertified cell that ran through its whole row, for a resume
    // at a longer row
    if (fa.resume != nullptr && NS <= 64) {
        const bool save = valid && loop_status == SUCCESS && status == SECURITY_RADIUS_NOT_REACHED;
        if (save) {
            ClipState* out = reinterpret_cast<ClipState*>(fa.cstate + (size_t)sid * CLIP_STATE_BYTES);
            for (int p = gl; p < nb_v; p += G) {
                out->plane[p] = sm.plane[p];
                out->pn1[p] = sm.pn1[p];
                out->pmx[p] = sm.pmx[p];
                out->pvid[p] = sm.pvid[p];
            }
#pragma unroll
            for (int i = 0; i < S; i++) {
                const int gs = i * G + gl;
                out->tri[gs] = ((R.live >> i) & 1u) ? sm.tri[gs] : 0xffffffffu;
#pragma unroll
                for (int e = 0; e < D; e++) out->adj[e][gs] = sm.adj[e][gs];
            }
            if (gl == 0) {
                out->nb_v = nb_v;
                out->nlive = nlive;
            }
        }
        if (valid && gl == 0) {
            fa.resume[sid] = save ? (shape_tag | k) : 0;
            if (save) fa.lastnb[sid] = zl;
        }
    }

    // ---------------------------------------------------------- extraction
    const bool ext = valid && status == SUCCESS;
    // normals and neighbour ids of my planes (cell.cpp:575-579, 646-669)
    double nx[C::PPL], ny[C::PPL], nz[C::PPL];
    int32_t pv[C::PPL];
#pragma unroll
    for (int i = 0; i < C::PPL; i++) {
        const int p = gl + G * i;
        nx[i] = ny[i] = nz[i] = 0.0;
        pv[i] = -1;
        if (ext && p < nb_v) {
            const Plane pe = sm.plane[p];
            const double nl = sqrt(pe.x * pe.x + pe.y * pe.y + pe.z * pe.z);
            nx[i] = pe.x / nl;
            ny[i] = pe.y / nl;
            nz[i] = pe.z / nl;
            pv[i] = sm.pvid[p];
        }
        if (p < MAXP) sm.first[p] = 0xffffffffu;
    }
    block_sync();
    // vertex positions to LDS (over the planes), lowest slot per plane
    if (ext) {
#pragma unroll
        for (int s = 0; s < S; s++) {
            if (!((R.live >> s) & 1u)) continue;
            const int gs = s * G + gl;
            const uint32_t t = sm.tri[gs];
            // a face's cycle starts at its vertex with the smallest sorted
            // plane tuple: the same vertex whatever slot it occupies, so the
            // fan sums round identically for every kernel shape
            const uint32_t key = vertex_key<D>(t);
#pragma unroll
            for (int ee = 0; ee < D; ee++) atomicMin(&sm.first[tri_p(t, ee)], (key << 8) | (uint32_t)gs);
        }
    }
    block_sync();
    if (ext) {
#pragma unroll
        for (int s = 0; s < S; s++) {
            if (!((R.live >> s) & 1u)) continue;
            const int gs = s * G + gl;
            sm.vpos[3 * gs] = R.px[s];
            sm.vpos[3 * gs + 1] = R.py[s];
            sm.vpos[3 * gs + 2] = R.pz[s];
        }
    }
    block_sync();
    // faces: walk each plane's vertex cycle through the adjacency; per plane
    // the volume and centroid moments of its fan
    double pvol[C::PPL], pcx[C::PPL], pcy[C::PPL], pcz[C::PPL];
    bool isface[C::PPL];
    double farea[C::PPL], fcx[C::PPL], fcy[C::PPL], fcz[C::PPL];
    const double* vpos = sm.vpos;
#pragma unroll
    for (int i = 0; i < C::PPL; i++) {
        const int p = gl + G * i;
        isface[i] = false;
        farea[i] = fcx[i] = fcy[i] = fcz[i] = 0.0;
        pvol[i] = pcx[i] = pcy[i] = pcz[i] = 0.0;
        double vol = 0.0, cxs = 0.0, cys = 0.0, czs = 0.0;
        if (!(ext && p < nb_v)) continue;
        if (sm.first[p] == 0xffffffffu) continue;
        const uint32_t v0 = sm.first[p] & 0xffu;
        const double x0 = vpos[3 * v0], y0 = vpos[3 * v0 + 1], z0 = vpos[3 * v0 + 2];
        if (D == 2) {
            const uint32_t t0 = sm.tri[v0];
            const int ee = tri_p(t0, 0) == p ? 0 : 1;
            const int v1 = sm.adj[ee][v0];
            const double x1 = vpos[3 * v1], y1 = vpos[3 * v1 + 1];
            const double dx = x1 - x0, dy = y1 - y0;
            farea[i] = sqrt(dx * dx + dy * dy);
            fcx[i] = 0.5 * (x0 + x1);
            fcy[i] = 0.5 * (y0 + y1);
            const double ta = 0.5 * fabs((x0 - sx) * (y1 - sy) - (x1 - sx) * (y0 - sy));
            vol += ta;
            cxs += ta * (sx + x0 + x1) / 3.0;
            cys += ta * (sy + y0 + y1) / 3.0;
            isface[i] = true;
            pvol[i] = vol;
            pcx[i] = cxs;
            pcy[i] = cys;
        } else {
            int cnt = 0;
            uint32_t cv = v0;
            double ax = 0, ay = 0, az = 0;  // area-weighted centroid sums
            double area = 0.0;
            double px = 0, py = 0, pz = 0;
            bool have_prev = false;
            do {
                const double x = vpos[3 * cv], y = vpos[3 * cv + 1], zz = vpos[3 * cv + 2];
                if (have_prev && cnt >= 2) {
                    const double e1x = px - x0, e1y = py - y0, e1z = pz - z0;
                    const double e2x = x - x0, e2y = y - y0, e2z = zz - z0;
                    const double crx = e1y * e2z - e1z * e2y;
                    const double cry = e1z * e2x - e1x * e2z;
                    const double crz = e1x * e2y - e1y * e2x;
                    const double ta = 0.5 * sqrt(crx * crx + cry * cry + crz * crz);
                    area += ta;
                    ax += ta * (x0 + px + x) / 3.0;
                    ay += ta * (y0 + py + y) / 3.0;
                    az += ta * (z0 + pz + zz) / 3.0;
                    // tetra (seed, v0, prev, cur): seed inside => same sign on a face
                    const double a1 = x0 - sx, a2 = y0 - sy, a3 = z0 - sz;
                    const double b1 = px - sx, b2 = py - sy, b3 = pz - sz;
                    const double c1 = x - sx, c2 = y - sy, c3 = zz - sz;
                    const double tv =
                        fabs(a1 * (b2 * c3 - b3 * c2) + a2 * (b3 * c1 - b1 * c3) + a3 * (b1 * c2 - b2 * c1)) /
                        6.0;
                    vol += tv;
                    cxs += tv * (sx + x0 + px + x) / 4.0;
                    cys += tv * (sy + y0 + py + y) / 4.0;
                    czs += tv * (sz + z0 + pz + zz) / 4.0;
                }
                px = x;
                py = y;
                pz = zz;
                have_prev = true;
                cnt++;
                const uint32_t t = sm.tri[cv];
                const int ee = tri_p(t, 0) == p ? 0 : (tri_p(t, 1) == p ? 1 : 2);
                cv = sm.adj[ee][cv];
            } while (cv != v0 && cnt <= NS);
            pvol[i] = vol;
            pcx[i] = cxs;
            pcy[i] = cys;
            pcz[i] = czs;
            if (cnt >= 3) {
                isface[i] = true;
                farea[i] = area;
                if (area > 0) {
                    fcx[i] = ax / area;
                    fcy[i] = ay / area;
                    fcz[i] = az / area;
                }
            }
        }
    }
    // face order = plane order among planes that form a face
    uint64_t fm = 0;
#pragma unroll
    for (int i = 0; i < C::PPL; i++) {
        uint64_t b = (wave_ballot(isface[i]) >> shift) & gmask;
        fm |= b << (G * i);
    }
    // cell volume and centroid moments: a pairwise sum over the plane index,
    // lowest bit first (lanes hold bits below log2 G, the planes of a lane
    // the bits above), i.e. the same tree for every kernel shape
    const auto add = [](double theirs, double mine) { return mine + theirs; };
#pragma unroll
    for (int i = 0; i < C::PPL; i++) {
        pvol[i] = group_butterfly_d<G>(pvol[i], add);
        pcx[i] = group_butterfly_d<G>(pcx[i], add);
        pcy[i] = group_butterfly_d<G>(pcy[i], add);
        pcz[i] = group_butterfly_d<G>(pcz[i], add);
    }
#pragma unroll
    for (int w = 1; w < C::PPL; w <<= 1)
#pragma unroll
        for (int i = 0; i + w < C::PPL; i += 2 * w) {  // absent planes add nothing
            pvol[i] += pvol[i + w];
            pcx[i] += pcx[i + w];
            pcy[i] += pcy[i + w];
            pcz[i] += pcz[i + w];
        }
    const double vol = pvol[0], cxs = pcx[0], cys = pcy[0], czs = pcz[0];
    if (valid) {
        const int64_t st = fa.stride;
#pragma unroll
        for (int i = 0; i < C::PPL; i++) {
            const int p = gl + G * i;
            if (!isface[i]) continue;
            int fi = popc64(fm & ((1ull << p) - 1ull));
            if (BALL && p >= first_ball) {
                // faces of the completion's planes in (distance, input
                // index) order, after the row's
                const uint32_t zp = (uint32_t)sm.pvid[p];
                const double kp = dist2<D>(pts + (size_t)D * zp, pts + (size_t)D * sid);
                fi = popc64(first_ball >= 64 ? fm : (fm & ((1ull << first_ball) - 1ull)));
                for (int q = first_ball; q < nb_v; q++) {
                    if (q == p || !((fm >> q) & 1ull)) continue;
                    const uint32_t zq = (uint32_t)sm.pvid[q];
                    const double kq = dist2<D>(pts + (size_t)D * zq, pts + (size_t)D * sid);
                    if (kq < kp || (kq == kp && bg.perm[zq] < bg.perm[zp])) fi++;
                }
            }
            const int64_t at = (int64_t)fi * st + sid;
            fa.nbr[at] = pv[i];
            if (fa.nbr_cm) fa.nbr_cm[(int64_t)sid * FACE_SLOTS + fi] = pv[i];
            fa.area[at] = farea[i];
            const int64_t plane_stride = st * FACE_SLOTS;
            fa.nrm[at] = nx[i];
            fa.nrm[plane_stride + at] = ny[i];
            fa.nrm[2 * plane_stride + at] = nz[i];
            fa.cen[at] = fcx[i];
            fa.cen[plane_stride + at] = fcy[i];
            fa.cen[2 * plane_stride + at] = D == 3 ? fcz[i] : 0.0;
        }
        if (gl == 0) {
            fa.count[sid] = ext ? (uint8_t)popc64(fm) : (uint8_t)0;
            fa.vol[sid] = ext ? vol : 0.0;
            const double iv = vol != 0.0 ? 1.0 / vol : 0.0;
            fa.ccen[sid] = ext ? cxs * iv : 0.0;
            fa.ccen[st + sid] = ext ? cys * iv : 0.0;
            fa.ccen[2 * st + sid] = ext ? czs * iv : 0.0;
            fa.status[sid] = status;
            fa.kcert[sid] = kc;
        }
    }
}

}  // namespace pg
