// ORACLE — test infrastructure, not product code.
//
// CPU FP64 restatement of the reference's moving-mesh hot path.  Only the
// tests, __graft_entry__.smoke() and bench.py's cpu_baseline leg load this
// library, and only as the checker.  Built with -ffp-contract=off so every
// FP64 expression rounds exactly as the reference's non-FMA x86-64 build.
//
// How each part is pinned (tests/test_oracle_*.py):
//   or_ngrid, or_grid_sort ........ golden vectors written by the reference
//                                   harness (N_grid, sorted points, permutation);
//                                   closed form: stable sort by (cell, index).
//   or_periodic_ghosts ............ golden ghost_ids (reference periodic mesh);
//                                   invariant: every ghost is a +-1 shift of its
//                                   source, regions match the buff tests.
//   or_knn ........................ golden knearest (reference knn::solve) and
//                                   brute force on small inputs.
//   clipping / extraction ......... golden meshes (topology bit-exact, volumes,
//                                   normals, areas); closed forms on BCC and
//                                   hexagonal lattices; sum of volumes = box;
//                                   sum_f A_f n_f = 0; face (i,j) <=> (j,i).
//   restart ....................... equals the direct build at the larger k.
//   or_dt_cfl, or_hydro_step(0) ... golden post-step states from the reference
//                                   hydro loop.
//   or_hydro_step(1), walls ....... two-cell known answer, F(U,U)=f(U),
//                                   zero mass flux through a wall, uniform
//                                   state stays uniform (parity against the
//                                   reference does not exist for them).
//   centroids ..................... closed forms on lattices and the
//                                   divergence identity (no reference output).
#include "oracle.h"

#include <omp.h>

#include <algorithm>
#include <array>
#include <numeric>
#include <cfloat>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

namespace {

enum Status {
    triangle_overflow = 0,
    vertex_overflow = 1,
    inconsistent_boundary = 2,
    security_radius_not_reached = 3,
    success = 4,
    needs_exact_predicates = 5
};

struct P4 {
    double x, y, z, w;
};

inline P4 sub4(const P4& a, const P4& b) { return {a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w}; }
inline P4 add4(const P4& a, const P4& b) { return {a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }
inline double dot3(const P4& a, const P4& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline P4 cross3(const P4& a, const P4& b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x, 0};
}
// allvars.h:168-195 — the evaluation order of these determinants is part of
// the bit pattern the reference produces.
inline double d2(double a11, double a12, double a21, double a22) { return a11 * a22 - a12 * a21; }
inline double d3(double a11, double a12, double a13, double a21, double a22, double a23, double a31,
                 double a32, double a33) {
    return a11 * d2(a22, a23, a32, a33) - a21 * d2(a12, a13, a32, a33) + a31 * d2(a12, a13, a22, a23);
}
inline double d4(double a11, double a12, double a13, double a14, double a21, double a22, double a23,
                 double a24, double a31, double a32, double a33, double a34, double a41, double a42,
                 double a43, double a44) {
    double m12 = a21 * a12 - a11 * a22;
    double m13 = a31 * a12 - a11 * a32;
    double m14 = a41 * a12 - a11 * a42;
    double m23 = a31 * a22 - a21 * a32;
    double m24 = a41 * a22 - a21 * a42;
    double m34 = a41 * a32 - a31 * a42;
    double m123 = m23 * a13 - m13 * a23 + m12 * a33;
    double m124 = m24 * a13 - m14 * a23 + m12 * a43;
    double m134 = m34 * a13 - m14 * a33 + m13 * a43;
    double m234 = m34 * a23 - m24 * a33 + m23 * a43;
    return m234 * a14 - m134 * a24 + m124 * a34 - m123 * a44;
}
inline double dmin(double a, double b) { return (b < a) ? b : a; }  // std::min
inline double dmax(double a, double b) { return (a < b) ? b : a; }  // std::max

const uint8_t END = 255;

// ------------------------------------------------------------------ grid
// knn.cpp:208-222
inline int cell_of(int dim, int N, const double* p) {
    int i = (int)std::floor(p[0] * (double)N);
    int j = (int)std::floor(p[1] * (double)N);
    i = std::max(0, std::min(i, N - 1));
    j = std::max(0, std::min(j, N - 1));
    if (dim == 2) return i + j * N;
    int k = (int)std::floor(p[2] * (double)N);
    k = std::max(0, std::min(k, N - 1));
    return i + j * N + k * N * N;
}

inline double dist2(int dim, const double* a, const double* b) {
    double dx = a[0] - b[0], dy = a[1] - b[1];
    if (dim == 2) return dx * dx + dy * dy;
    double dz = a[2] - b[2];
    return dx * dx + dy * dy + dz * dz;
}

// ------------------------------------------------------------- exact sign
// Degenerate conflict tests (the reference's needs_exact_predicates,
// cell.cpp:216-218, which leaves the cell failed, voronoi.cpp:104-110),
// decided exactly.  No reference counterpart; the rule (pinned against
// rational arithmetic in tests/test_oracle_exact.py):
//   sign of det[rows: the vertex's planes, then the new plane] over the
//   planes computed without rounding from the generators: bisector of seed s
//   and point b = (s - b, (|b|^2 - |s|^2) / 2), box plane = the initial
//   cell's; ties broken by Simulation of Simplicity on power weights
//   eps^rank(key): point x's weight adds (w_s - w_x) C_x / 2 to the
//   determinant (C_x the cofactor of its w entry), the first non-zero
//   coefficient in key order decides (the seed's: sum of the bisector
//   cofactors; a bisector point's: -C_x).
// Integers: every value is m 2^e; with E the smallest e involved, rows
// scaled to (2 (S - B), |B|^2 - |S|^2) and (sg e_a, -sg C) (S, B, C: values
// over 2^E) keep the sign of the determinant.  Determinants by the Leibniz
// sum over permutations, sign-magnitude integers in base 2^32.
struct BigZ {
    int sg = 0;                 // -1, 0, +1
    std::vector<uint32_t> mag;  // little endian, no leading zero limbs
};

void bz_trim(std::vector<uint32_t>& a) {
    while (!a.empty() && a.back() == 0) a.pop_back();
}
int bz_cmp_mag(const std::vector<uint32_t>& a, const std::vector<uint32_t>& b) {
    if (a.size() != b.size()) return a.size() < b.size() ? -1 : 1;
    for (size_t i = a.size(); i-- > 0;)
        if (a[i] != b[i]) return a[i] < b[i] ? -1 : 1;
    return 0;
}
std::vector<uint32_t> bz_add_mag(const std::vector<uint32_t>& a, const std::vector<uint32_t>& b) {
    std::vector<uint32_t> r(std::max(a.size(), b.size()) + 1, 0);
    uint64_t c = 0;
    for (size_t i = 0; i < r.size(); i++) {
        uint64_t t = c + (i < a.size() ? a[i] : 0) + (i < b.size() ? b[i] : 0);
        r[i] = (uint32_t)t;
        c = t >> 32;
    }
    bz_trim(r);
    return r;
}
std::vector<uint32_t> bz_sub_mag(const std::vector<uint32_t>& a, const std::vector<uint32_t>& b) {  // a >= b
    std::vector<uint32_t> r(a.size(), 0);
    int64_t br = 0;
    for (size_t i = 0; i < a.size(); i++) {
        int64_t t = (int64_t)a[i] - (i < b.size() ? b[i] : 0) - br;
        br = t < 0;
        r[i] = (uint32_t)(t + (br << 32));
    }
    bz_trim(r);
    return r;
}
BigZ bz_add(const BigZ& a, const BigZ& b) {
    if (a.sg == 0) return b;
    if (b.sg == 0) return a;
    BigZ r;
    if (a.sg == b.sg) {
        r.mag = bz_add_mag(a.mag, b.mag);
        r.sg = a.sg;
        return r;
    }
    const int c = bz_cmp_mag(a.mag, b.mag);
    if (c == 0) return r;
    r.mag = c > 0 ? bz_sub_mag(a.mag, b.mag) : bz_sub_mag(b.mag, a.mag);
    r.sg = c > 0 ? a.sg : b.sg;
    return r;
}
BigZ bz_neg(BigZ a) {
    a.sg = -a.sg;
    return a;
}
BigZ bz_mul(const BigZ& a, const BigZ& b) {
    BigZ r;
    if (a.sg == 0 || b.sg == 0) return r;
    std::vector<uint64_t> acc(a.mag.size() + b.mag.size() + 1, 0);
    std::vector<uint32_t> m(acc.size(), 0);
    for (size_t i = 0; i < a.mag.size(); i++) {
        uint64_t c = 0;
        for (size_t j = 0; j < b.mag.size(); j++) {
            uint64_t t = (uint64_t)a.mag[i] * b.mag[j] + m[i + j] + c;
            m[i + j] = (uint32_t)t;
            c = t >> 32;
        }
        for (size_t k = i + b.mag.size(); c; k++) {
            uint64_t t = (uint64_t)m[k] + c;
            m[k] = (uint32_t)t;
            c = t >> 32;
        }
    }
    bz_trim(m);
    r.mag = m;
    r.sg = a.sg * b.sg;
    return r;
}
// m 2^sh for an integer |m| < 2^53 and sh >= 0
BigZ bz_from(int64_t m, int sh) {
    BigZ r;
    if (m == 0) return r;
    r.sg = m < 0 ? -1 : 1;
    uint64_t a = (uint64_t)(m < 0 ? -m : m);
    std::vector<uint32_t> v(sh / 32 + 4, 0);
    const int q = sh / 32, b = sh % 32;
    // a < 2^53 shifted by b < 32 bits spans at most three limbs
    unsigned __int128 x = (unsigned __int128)a << b;
    for (int i = 0; i < 3; i++) v[q + i] = (uint32_t)(x >> (32 * i));
    bz_trim(v);
    r.mag = v;
    return r;
}

// Leibniz determinant of an n x n matrix (n <= 4), row major.
BigZ bz_det(const std::vector<BigZ>& a, int n) {
    int perm[4] = {0, 1, 2, 3};
    BigZ det;
    do {
        // parity by counting inversions
        int inv = 0;
        for (int i = 0; i < n; i++)
            for (int j = i + 1; j < n; j++) inv += perm[i] > perm[j];
        BigZ t = a[0 * n + perm[0]];
        for (int i = 1; i < n && t.sg != 0; i++) t = bz_mul(t, a[(size_t)i * n + perm[i]]);
        det = bz_add(det, (inv & 1) ? bz_neg(t) : t);
    } while (std::next_permutation(perm, perm + n));
    return det;
}

// value = m 2^e (m = 0 for zero); false for non-finite or subnormal values
bool split_exact(double v, int64_t& m, int& e) {
    if (v == 0.0) {
        m = 0;
        e = 0;
        return true;
    }
    if (!std::isfinite(v) || std::fabs(v) < DBL_MIN) return false;
    int ex;
    const double f = std::frexp(v, &ex);  // v = f 2^ex, 0.5 <= |f| < 1
    m = (int64_t)std::ldexp(f, 53);
    e = ex - 53;
    return true;
}

// One row of the exact conflict determinant: box plane (axis, sign) or the
// bisector of the seed and point p.
struct ExRow {
    int box;  // 0: bisector; +-(axis + 1): box plane x_axis >= lo (+) / <= hi (-)
    double p[3];
    uint32_t key;
};

// +1: conflict, -1: no conflict, 0: undecided (a subnormal or non-finite value)
int exact_conflict(int dim, const double* s, uint32_t skey, const ExRow* rows) {
    const int nr = dim + 1;
    const double lo = -1e-14, hi = 1.0 + 1e-14;  // the initial cell's box
    std::vector<double> vals(s, s + dim);
    for (int k = 0; k < nr; k++) {
        if (rows[k].box) vals.push_back(rows[k].box > 0 ? lo : hi);
        else
            for (int d = 0; d < dim; d++) vals.push_back(rows[k].p[d]);
    }
    int emin = INT_MAX;
    for (double v : vals) {
        int64_t m;
        int e;
        if (!split_exact(v, m, e)) return 0;
        if (m != 0) emin = std::min(emin, e);
    }
    if (emin == INT_MAX) emin = 0;
    auto I = [&](double v) {
        int64_t m = 0;
        int e = 0;
        split_exact(v, m, e);
        return m == 0 ? BigZ() : bz_from(m, e - emin);
    };
    BigZ S[3];
    BigZ S2;
    for (int d = 0; d < dim; d++) {
        S[d] = I(s[d]);
        S2 = bz_add(S2, bz_mul(S[d], S[d]));
    }
    // matrix rows (normal..., w)
    std::vector<BigZ> M((size_t)nr * nr);
    for (int k = 0; k < nr; k++) {
        BigZ* r = &M[(size_t)k * nr];
        if (rows[k].box) {
            const int ax = std::abs(rows[k].box) - 1;
            const int sg = rows[k].box > 0 ? 1 : -1;
            for (int d = 0; d < dim; d++) r[d] = d == ax ? bz_from(sg, 0) : BigZ();
            const BigZ c = I(rows[k].box > 0 ? lo : hi);
            r[dim] = sg > 0 ? bz_neg(c) : c;
        } else {
            BigZ B2;
            for (int d = 0; d < dim; d++) {
                const BigZ b = I(rows[k].p[d]);
                const BigZ df = bz_add(S[d], bz_neg(b));
                r[d] = bz_add(df, df);
                B2 = bz_add(B2, bz_mul(b, b));
            }
            r[dim] = bz_add(B2, bz_neg(S2));
        }
    }
    const BigZ det = bz_det(M, nr);
    if (det.sg != 0) return det.sg;
    // cofactors of the w column
    std::vector<BigZ> C(nr);
    for (int k = 0; k < nr; k++) {
        std::vector<BigZ> sub;
        for (int i = 0; i < nr; i++)
            if (i != k)
                for (int j = 0; j < dim; j++) sub.push_back(M[(size_t)i * nr + j]);
        const BigZ m = bz_det(sub, dim);
        C[k] = ((k + dim) & 1) ? bz_neg(m) : m;
    }
    // candidates (key, coefficient sign) in key order
    std::vector<std::pair<uint32_t, int>> cand;
    BigZ cs;
    for (int k = 0; k < nr; k++)
        if (!rows[k].box) {
            cs = bz_add(cs, C[k]);
            cand.push_back({rows[k].key, -C[k].sg});
        }
    cand.push_back({skey, cs.sg});
    std::sort(cand.begin(), cand.end());
    for (auto& c : cand)
        if (c.second != 0) return c.second;
    return 0;
}

// ------------------------------------------------------------------ cell
// cell.cpp:42-411: the convex cell as planes + vertices (plane triples).
struct Cell {
    int dim, max_p, max_t;
    const double* pts;  // sorted points [n][dim]
    int seed_id;
    P4 seed;
    int status;
    int nb_v, nb_t, nb_r;
    uint8_t first_b;
    std::vector<P4> plane;
    std::vector<int> pvid;
    std::vector<uint8_t> tri;  // 3 bytes per vertex (third unused in 2D)
    std::vector<uint8_t> bnext;
    // exact-predicate fallback.  xmode: a test is trusted in floating point
    // only beyond XTOL times the product of its rows' magnitudes (pmw: the
    // magnitude the rounding of w is relative to); below, it is decided by
    // exact_conflict with the tie-break keys per sorted id (keys set) or it
    // flags the cell (keys nullptr).
    const uint32_t* keys = nullptr;
    bool xmode = false;
    std::vector<double> pmw;

    Cell(int dim_, int maxp, int maxt, const double* p, int sid, const uint32_t* keys_ = nullptr,
         bool xmode_ = false)
        : dim(dim_), max_p(maxp), max_t(maxt), pts(p), seed_id(sid), keys(keys_), xmode(xmode_ || keys_) {
        plane.assign(max_p, P4{0, 0, 0, 0});
        pmw.assign(max_p, 0.0);
        pvid.assign(max_p, -1);
        tri.assign(3 * (size_t)max_t, 0);
        bnext.assign(max_p, END);
        first_b = END;
        status = success;
        seed = point(sid);
        const double eps = 1e-14;
        const double lo = -eps, hi = 1.0 + eps;
        plane[0] = {1.0, 0.0, 0.0, -lo};
        plane[1] = {-1.0, 0.0, 0.0, hi};
        plane[2] = {0.0, 1.0, 0.0, -lo};
        plane[3] = {0.0, -1.0, 0.0, hi};
        if (dim == 2) {
            for (int b = 0; b < 4; b++) pmw[b] = std::fabs(plane[b].w);
            const uint8_t t0[4][2] = {{2, 0}, {1, 2}, {3, 1}, {0, 3}};
            for (int t = 0; t < 4; t++) set_tri(t, t0[t][0], t0[t][1], 0);
            nb_v = 4;
            nb_t = 4;
        } else {
            plane[4] = {0.0, 0.0, 1.0, -lo};
            plane[5] = {0.0, 0.0, -1.0, hi};
            for (int b = 0; b < 6; b++) pmw[b] = std::fabs(plane[b].w);
            const uint8_t t0[8][3] = {{2, 5, 0}, {5, 3, 0}, {1, 5, 2}, {5, 1, 3},
                                      {4, 2, 0}, {4, 0, 3}, {2, 4, 1}, {4, 3, 1}};
            for (int t = 0; t < 8; t++) set_tri(t, t0[t][0], t0[t][1], t0[t][2]);
            nb_v = 6;
            nb_t = 8;
        }
        nb_r = 0;
    }

    P4 point(int id) const {
        const double* f = pts + (size_t)dim * id;
        return {f[0], f[1], dim == 3 ? f[2] : 0.0, 1.0};
    }
    uint8_t& tp(int t, int e) { return tri[3 * (size_t)t + e]; }
    void set_tri(int t, uint8_t a, uint8_t b, uint8_t c) {
        tp(t, 0) = a;
        tp(t, 1) = b;
        tp(t, 2) = c;
    }
    void swap_tri(int a, int b) {
        for (int e = 0; e < 3; e++) std::swap(tp(a, e), tp(b, e));
    }
    bool refs(int t, int p) {
        for (int d = 0; d < dim; d++)
            if (tp(t, d) == p) return true;
        return false;
    }

    // cell.cpp:158-172
    int new_halfplane(int vid) {
        if (nb_v >= max_p) {
            status = vertex_overflow;
            return -1;
        }
        P4 B = point(vid);
        P4 dir = sub4(seed, B);
        P4 ave2 = add4(seed, B);
        double dot = dot3(ave2, dir);
        plane[nb_v] = {dir.x, dir.y, dir.z, -dot / 2.0};
        pmw[nb_v] = 0.5 * (std::fabs(ave2.x * dir.x) + std::fabs(ave2.y * dir.y) + std::fabs(ave2.z * dir.z));
        pvid[nb_v] = vid;
        nb_v++;
        return nb_v - 1;
    }

    // the degenerate test of vertex t against plane cur, decided exactly
    int exact_test(int t, int cur) {
        ExRow rows[4];
        for (int r = 0; r <= dim; r++) {
            const int pid = r < dim ? tp(t, r) : cur;
            ExRow& x = rows[r];
            x.p[0] = x.p[1] = x.p[2] = 0.0;
            x.key = 0;
            x.box = 0;
            if (pvid[pid] < 0) {
                x.box = (pid & 1) ? -(pid / 2 + 1) : (pid / 2 + 1);
            } else {
                const double* f = pts + (size_t)dim * pvid[pid];
                for (int d = 0; d < dim; d++) x.p[d] = f[d];
                x.key = keys[pvid[pid]];
            }
        }
        const double* sp = pts + (size_t)dim * seed_id;
        return exact_conflict(dim, sp, keys[seed_id], rows);
    }

    // cell.cpp:175-221
    bool conflict(int t, const P4& eqn, int cur) {
        const P4& p1 = plane[tp(t, 0)];
        const P4& p2 = plane[tp(t, 1)];
        double det, eps;
        if (dim == 2) {
            det = d3(p1.x, p2.x, eqn.x, p1.y, p2.y, eqn.y, p1.w, p2.w, eqn.w);
            double mx = std::max({std::fabs(p1.x), std::fabs(p2.x), std::fabs(eqn.x)});
            double my = std::max({std::fabs(p1.y), std::fabs(p2.y), std::fabs(eqn.y)});
            double mw = std::max({std::fabs(p1.w), std::fabs(p2.w), std::fabs(eqn.w)});
            double mm = std::max({mx, my, mw});
            eps = 1e-14 * mx * my * mw;
            eps *= mm;
        } else {
            const P4& p3 = plane[tp(t, 2)];
            det = d4(p1.x, p2.x, p3.x, eqn.x, p1.y, p2.y, p3.y, eqn.y, p1.z, p2.z, p3.z, eqn.z, p1.w,
                     p2.w, p3.w, eqn.w);
            double mx = std::max({std::fabs(p1.x), std::fabs(p2.x), std::fabs(p3.x), std::fabs(eqn.x)});
            double my = std::max({std::fabs(p1.y), std::fabs(p2.y), std::fabs(p3.y), std::fabs(eqn.y)});
            double mz = std::max({std::fabs(p1.z), std::fabs(p2.z), std::fabs(p3.z), std::fabs(eqn.z)});
            eps = 1e-14 * mx * my * mz;
            double mm = std::max(std::max(mx, my), mz);
            eps *= (mm * mm);
        }
        double xb = 0.0;
        if (xmode) {
            xb = 1e-12;  // XTOL of the device code (exact.h)
            for (int r = 0; r <= dim; r++) {
                const int pid = r < dim ? tp(t, r) : cur;
                const P4& q = plane[pid];
                xb *= std::fabs(q.x) + std::fabs(q.y) + std::fabs(q.z) + pmw[pid];
            }
        }
        if (std::fabs(det) < eps || std::fabs(det) < xb) {
            if (keys) {
                const int r = exact_test(t, cur);
                if (r != 0) return r > 0;
            }
            status = needs_exact_predicates;
        }
        return det > 0.0;
    }

    // cell.cpp:224-343
    void compute_boundary() {
        for (int i = 0; i < max_p; i++) bnext[i] = END;
        first_b = END;
        if (dim == 2) {
            std::vector<int> cnt(max_p, 0);
            for (int r = 0; r < nb_r; r++) {
                cnt[tp(nb_t + r, 0)]++;
                cnt[tp(nb_t + r, 1)]++;
            }
            uint8_t bl[2] = {0, 0};
            int nbl = 0;
            for (int p = 0; p < nb_v; p++)
                if (cnt[p] == 1 && nbl < 2) bl[nbl++] = (uint8_t)p;
            if (nbl != 2) {
                status = inconsistent_boundary;
                return;
            }
            first_b = bl[0];
            bnext[bl[0]] = bl[1];
            bnext[bl[1]] = bl[0];
            return;
        }
        int nb_iter = 0;
        int t = nb_t;
        while (nb_r > 0) {
            if (nb_iter++ > 100) {
                status = inconsistent_boundary;
                return;
            }
            bool in_border[3], opp[3];
            for (int e = 0; e < 3; e++) in_border[e] = (bnext[tp(t, e)] != END);
            for (int e = 0; e < 3; e++) opp[e] = (bnext[tp(t, (e + 1) % 3)] == tp(t, e));
            bool simple = true;
            for (int e = 0; e < 3; e++)
                if (!opp[e] && !opp[(e + 1) % 3] && in_border[(e + 1) % 3]) simple = false;
            if (!opp[0] && !opp[1] && !opp[2]) {
                if (first_b == END) {
                    for (int e = 0; e < 3; e++) bnext[tp(t, e)] = tp(t, (e + 1) % 3);
                    first_b = tp(t, 0);
                } else {
                    simple = false;
                }
            }
            if (!simple) {
                t++;
                if (t == nb_t + nb_r) t = nb_t;
                continue;
            }
            for (int e = 0; e < 3; e++)
                if (!opp[e]) bnext[tp(t, e)] = tp(t, (e + 1) % 3);
            for (int e = 0; e < 3; e++) {
                if (opp[e] && opp[(e + 1) % 3]) {
                    uint8_t q = tp(t, (e + 1) % 3);
                    if (first_b == q) first_b = bnext[q];
                    bnext[q] = END;
                }
            }
            swap_tri(t, nb_t + nb_r - 1);
            t = nb_t;
            nb_r--;
        }
    }

    // cell.cpp:346-364
    void new_vertex(uint8_t i, uint8_t j, uint8_t k) {
        if (nb_t + 1 >= max_t) {
            status = triangle_overflow;
            return;
        }
        if (dim == 2) {
            double rw = d2(plane[i].x, plane[i].y, plane[j].x, plane[j].y);
            if (rw > 0)
                set_tri(nb_t, j, i, 0);
            else
                set_tri(nb_t, i, j, 0);
        } else {
            set_tri(nb_t, i, j, k);
        }
        nb_t++;
    }

    // cell.cpp:110-154
    void clip(int vid) {
        int cur = new_halfplane(vid);
        if (status == vertex_overflow) return;
        P4 eqn = plane[cur];
        nb_r = 0;
        int i = 0;
        while (i < nb_t) {
            if (conflict(i, eqn, cur)) {
                nb_t--;
                swap_tri(i, nb_t);
                nb_r++;
            } else {
                i++;
            }
        }
        if (status == needs_exact_predicates) return;
        if (nb_r == 0) {
            nb_v--;
            return;
        }
        compute_boundary();
        if (status != success) return;
        if (first_b == END) return;
        uint8_t cir = first_b;
        do {
            new_vertex((uint8_t)cur, cir, dim == 3 ? bnext[cir] : 0);
            if (status != success) return;
            cir = bnext[cir];
        } while (cir != first_b);
    }

    // cell.cpp:388-411
    P4 vertex(int t, bool divide = true) {
        const P4& a = plane[tp(t, 0)];
        const P4& b = plane[tp(t, 1)];
        P4 r;
        if (dim == 2) {
            r.x = -d2(a.w, a.y, b.w, b.y);
            r.y = -d2(a.x, a.w, b.x, b.w);
            r.z = 0;
            r.w = d2(a.x, a.y, b.x, b.y);
            if (divide) return {r.x / r.w, r.y / r.w, 0, 1};
        } else {
            const P4& c = plane[tp(t, 2)];
            r.x = -d3(a.w, a.y, a.z, b.w, b.y, b.z, c.w, c.y, c.z);
            r.y = -d3(a.x, a.w, a.z, b.x, b.w, b.z, c.x, c.w, c.z);
            r.z = -d3(a.x, a.y, a.w, b.x, b.y, b.w, c.x, c.y, c.w);
            r.w = d3(a.x, a.y, a.z, b.x, b.y, b.z, c.x, c.y, c.z);
            if (divide) return {r.x / r.w, r.y / r.w, r.z / r.w, 1};
        }
        return r;
    }

    // cell.cpp:369-384; with the exact fallback on, only neighbours clearly
    // beyond twice the largest vertex distance certify (relative margin
    // 1e-9): a tied neighbour can still cut a zero-area face under the
    // symbolic perturbation
    bool radius_reached(const P4& last, double margin = -1.0) {
        double vd = 0;
        for (int i = 0; i < nb_t; i++) {
            P4 pc = vertex(i);
            P4 df = sub4(pc, seed);
            vd = dmax(dot3(df, df), vd);
        }
        P4 df = sub4(last, seed);
        if (margin < 0) margin = keys ? 1.0 + 1e-9 : 1.0;
        return dot3(df, df) > 4 * vd * margin;
    }
    // exact mode without the exact test: a certificate within the margin of
    // a tie flags the cell for the exact pass
    bool near_tie(const P4& last) { return radius_reached(last, 1.0) && !radius_reached(last, 1.0 + 1e-9); }
};

// Extracted geometry of one cell, faces in plane order (cell.cpp:416-470).
struct CellOut {
    double volume = 0;
    double centroid[3] = {0, 0, 0};
    std::vector<int> nbr;
    std::vector<double> normal, area, fcen;
};

// cell.cpp:475-498 — 2D area from atan2-sorted vertices.
double area2d(const std::vector<P4>& v, int nt) {
    double cx = 0, cy = 0;
    for (int i = 0; i < nt; i++) {
        cx += v[i].x;
        cy += v[i].y;
    }
    cx /= nt;
    cy /= nt;
    std::vector<int> order(nt);
    for (int i = 0; i < nt; i++) order[i] = i;
    std::sort(order.begin(), order.end(), [&](int a, int b) {
        return std::atan2(v[a].y - cy, v[a].x - cx) < std::atan2(v[b].y - cy, v[b].x - cx);
    });
    double a = 0;
    for (int i = 0; i < nt; i++) {
        int j = (i + 1) % nt;
        a += v[order[i]].x * v[order[j]].y;
        a -= v[order[j]].x * v[order[i]].y;
    }
    return std::fabs(a) / 2.0;
}

// cell.cpp:416-669 plus centroids (no reference counterpart: the cell
// centroid is the volume-weighted mean of the cone/triangle centroids of the
// fan, the face centroid the area-weighted mean of the fan triangles).
void extract(Cell& c, CellOut& out) {
    const int dim = c.dim;
    std::vector<P4> vtx(c.nb_t);
    for (int i = 0; i < c.nb_t; i++) vtx[i] = c.vertex(i, true);
    double vol = 0;
    double cx = 0, cy = 0, cz = 0;
    if (dim == 2) out.volume = area2d(vtx, c.nb_t);
    for (int p = 0; p < c.nb_v; p++) {
        std::vector<int> idx;
        for (int i = 0; i < c.nb_t; i++)
            if (c.refs(i, p)) idx.push_back(i);
        if ((int)idx.size() < dim) continue;
        std::vector<P4> fv;
        if (dim == 2) {
            for (int i : idx) fv.push_back(vtx[i]);
        } else {
            // cell.cpp:515-572 — order by adjacency
            std::vector<int> ord{idx[0]};
            std::vector<bool> used(idx.size(), false);
            used[0] = true;
            for (size_t step = 1; step < idx.size(); step++) {
                int last = ord.back();
                uint8_t oth[2];
                int cnt = 0;
                for (int d = 0; d < 3; d++) {
                    uint8_t pl = c.tp(last, d);
                    if (pl != (uint8_t)p) oth[cnt++] = pl;
                }
                bool found = false;
                for (size_t j = 0; j < idx.size() && !found; j++) {
                    if (used[j]) continue;
                    for (int o = 0; o < 2; o++) {
                        if (c.refs(idx[j], oth[o])) {
                            ord.push_back(idx[j]);
                            used[j] = true;
                            found = true;
                            break;
                        }
                    }
                }
                if (!found) break;
            }
            if ((int)ord.size() < 3) continue;
            for (int i : ord) fv.push_back(vtx[i]);
        }
        const P4& pe = c.plane[p];
        double nl = std::sqrt(dot3(pe, pe));
        double n3[3] = {pe.x / nl, pe.y / nl, pe.z / nl};
        double meas = 0;
        double fc[3] = {0, 0, 0};
        if (dim == 2) {
            double dx = fv[1].x - fv[0].x, dy = fv[1].y - fv[0].y;
            meas = std::sqrt(dx * dx + dy * dy);
            fc[0] = 0.5 * (fv[0].x + fv[1].x);
            fc[1] = 0.5 * (fv[0].y + fv[1].y);
            // triangle (seed, v0, v1): area-weighted centroid contribution
            double ta = 0.5 * std::fabs((fv[0].x - c.seed.x) * (fv[1].y - c.seed.y) -
                                        (fv[1].x - c.seed.x) * (fv[0].y - c.seed.y));
            cx += ta * (c.seed.x + fv[0].x + fv[1].x) / 3.0;
            cy += ta * (c.seed.y + fv[0].y + fv[1].y) / 3.0;
        } else {
            {
                // orientation from the sum of the fan's cross products (the
                // reference takes the first triangle's, cell.cpp:515-572: the
                // same decision for a convex face, but also defined when the
                // first vertices coincide, as on the zero-area faces of the
                // exact-predicate fallback)
                P4 cr{0, 0, 0, 0};
                for (size_t i = 1; i + 1 < fv.size(); i++) {
                    const P4 t = cross3(sub4(fv[i], fv[0]), sub4(fv[i + 1], fv[0]));
                    cr.x += t.x;
                    cr.y += t.y;
                    cr.z += t.z;
                }
                P4 cen{0, 0, 0, 0};
                for (const P4& q : fv) {
                    cen.x += q.x;
                    cen.y += q.y;
                    cen.z += q.z;
                }
                cen.x /= fv.size();
                cen.y /= fv.size();
                cen.z /= fv.size();
                P4 outw = sub4(cen, c.seed);
                if (dot3(cr, outw) < 0) std::reverse(fv.begin(), fv.end());
            }
            P4 v0 = fv[0];
            for (size_t i = 1; i + 1 < fv.size(); i++) {
                P4 cr = cross3(sub4(fv[i], v0), sub4(fv[i + 1], v0));
                double ta = 0.5 * std::sqrt(cr.x * cr.x + cr.y * cr.y + cr.z * cr.z);
                meas += ta;
                fc[0] += ta * (v0.x + fv[i].x + fv[i + 1].x) / 3.0;
                fc[1] += ta * (v0.y + fv[i].y + fv[i + 1].y) / 3.0;
                fc[2] += ta * (v0.z + fv[i].z + fv[i + 1].z) / 3.0;
            }
            for (size_t i = 1; i + 1 < fv.size(); i++) {
                P4 a = sub4(fv[0], c.seed);
                P4 b = sub4(fv[i], c.seed);
                P4 cc = sub4(fv[i + 1], c.seed);
                double tv = dot3(a, cross3(b, cc)) / 6.0;
                vol += tv;
                // tetra (seed, v0, vi, vi+1) centroid
                cx += tv * (c.seed.x + fv[0].x + fv[i].x + fv[i + 1].x) / 4.0;
                cy += tv * (c.seed.y + fv[0].y + fv[i].y + fv[i + 1].y) / 4.0;
                cz += tv * (c.seed.z + fv[0].z + fv[i].z + fv[i + 1].z) / 4.0;
            }
            if (meas > 0) {
                fc[0] /= meas;
                fc[1] /= meas;
                fc[2] /= meas;
            }
        }
        out.nbr.push_back(c.pvid[p]);
        out.normal.insert(out.normal.end(), n3, n3 + 3);
        out.area.push_back(meas);
        out.fcen.insert(out.fcen.end(), fc, fc + 3);
    }
    if (dim == 3) out.volume = std::fabs(vol);
    double wv = (dim == 3) ? vol : out.volume;
    if (wv != 0) {
        out.centroid[0] = cx / wv;
        out.centroid[1] = cy / wv;
        out.centroid[2] = (dim == 3) ? cz / wv : 0.0;
    }
}

// ------------------------------------------------------------------ mesh
struct Mesh {
    int dim = 3, k = 0, N_grid = 0;
    int boundary = 0;
    double scale = 1.0;  // unscale factor of the sorted points (periodic: 1 + 2 buff)
    int64_t n_seeds = 0, n_hydro = 0;
    std::vector<double> sorted;
    std::vector<uint32_t> perm, knn;
    std::vector<int64_t> o2s;  // original id -> sorted id
    double times[4] = {0, 0, 0, 0};  // seconds: ghosts + grid, KNN, first clip pass, restarts
    // reference VMesh layout, original order
    std::vector<double> seeds, volumes, normal, area, ccen, fcen;
    std::vector<int64_t> fcount, fptr, ghost_ids;
    std::vector<int32_t> nbr, status, kcert;
    int64_t n_failed = 0;
    int64_t n_ball = 0;  // cells finished by the vertex-ball completion
};

// knn.cpp:16-117 + 119-204
void grid_sort(int dim, const double* pts, int64_t n, int N, double* sorted, uint32_t* perm,
               int32_t* counters, int32_t* ptrs) {
    int64_t ncell = (dim == 2) ? (int64_t)N * N : (int64_t)N * N * N;
    std::vector<int32_t> cnt(ncell, 0), start(ncell, 0);
    std::vector<int> cell(n);
    for (int64_t i = 0; i < n; i++) {
        cell[i] = cell_of(dim, N, pts + dim * i);
        cnt[cell[i]]++;
    }
    int32_t run = 0;
    for (int64_t c = 0; c < ncell; c++) {
        start[c] = cnt[c] > 0 ? run : 0;  // knn.cpp:176-180 leaves empty cells at 0
        run += cnt[c];
    }
    std::vector<int32_t> fill(ncell, 0);
    for (int64_t i = 0; i < n; i++) {
        int c = cell[i];
        int32_t pos = start[c] + fill[c]++;
        std::memcpy(sorted + dim * (int64_t)pos, pts + dim * i, sizeof(double) * dim);
        perm[pos] = (uint32_t)i;
    }
    if (counters) std::memcpy(counters, cnt.data(), sizeof(int32_t) * ncell);
    if (ptrs) std::memcpy(ptrs, start.data(), sizeof(int32_t) * ncell);
}

// knn.cpp:237-347: ring search, heap keyed by (dist^2, original index).
// Rows of the queries qlist[0..nq) (every point when qlist is null), row r
// written at out + r * k.
void knn_search_list(int dim, const double* sorted, int64_t n, int N, const int32_t* cnt,
                     const int32_t* start, const uint32_t* perm, int k, const int64_t* qlist, int64_t nq,
                     uint32_t* out) {
    (void)n;  // the grid tables bound every access
#pragma omp parallel
    {
        std::vector<double> hd(k);
        std::vector<uint32_t> hi(k), ho(k);
#pragma omp for schedule(dynamic, 256)
        for (int64_t r = 0; r < nq; r++) {
            const int64_t q = qlist ? qlist[r] : r;
            const double* p = sorted + dim * q;
            int c = cell_of(dim, N, p);
            int ci = c % N, cj = (c / N) % N, ck = (dim == 3) ? c / (N * N) : 0;
            int size = 0;
            auto less = [&](double da, uint32_t oa, double db, uint32_t ob) {
                return da < db || (da == db && oa < ob);
            };
            auto sift = [&](int j) {
                for (;;) {
                    int l = 2 * j + 1, r = l + 1, m = j;
                    if (l < size && less(hd[m], ho[m], hd[l], ho[l])) m = l;
                    if (r < size && less(hd[m], ho[m], hd[r], ho[r])) m = r;
                    if (m == j) return;
                    std::swap(hd[j], hd[m]);
                    std::swap(hi[j], hi[m]);
                    std::swap(ho[j], ho[m]);
                    j = m;
                }
            };
            auto push = [&](double d, uint32_t idx) {
                uint32_t o = perm[idx];
                if (size < k) {
                    // sift-up insert
                    int j = size++;
                    hd[j] = d;
                    hi[j] = idx;
                    ho[j] = o;
                    while (j > 0) {
                        int par = (j - 1) / 2;
                        if (!less(hd[par], ho[par], hd[j], ho[j])) break;
                        std::swap(hd[j], hd[par]);
                        std::swap(hi[j], hi[par]);
                        std::swap(ho[j], ho[par]);
                        j = par;
                    }
                } else if (less(d, o, hd[0], ho[0])) {
                    hd[0] = d;
                    hi[0] = idx;
                    ho[0] = o;
                    sift(0);
                }
            };
            int maxring = N;  // all rings: the grid is covered once ring >= N
            for (int ring = 0; ring <= maxring; ring++) {
                if (ring >= 1 && size == k) {
                    double md = (double)(ring - 1) / (double)N;
                    if (hd[0] < md * md) break;
                }
                int kl = (dim == 3) ? -ring : 0, kh = (dim == 3) ? ring : 0;
                for (int dk = kl; dk <= kh; dk++)
                    for (int dj = -ring; dj <= ring; dj++)
                        for (int di = -ring; di <= ring; di++) {
                            int cheb = std::max(std::abs(di), std::max(std::abs(dj), std::abs(dk)));
                            if (cheb != ring) continue;
                            int ii = ci + di, jj = cj + dj, kk = ck + dk;
                            if (ii < 0 || jj < 0 || kk < 0 || ii >= N || jj >= N || (dim == 3 && kk >= N))
                                continue;
                            int cc = ii + jj * N + kk * N * N;
                            for (int32_t s = start[cc]; s < start[cc] + cnt[cc]; s++) {
                                if (s == q) continue;
                                push(dist2(dim, p, sorted + dim * (int64_t)s), (uint32_t)s);
                            }
                        }
            }
            // ascending order
            std::vector<int> ord(size);
            for (int i = 0; i < size; i++) ord[i] = i;
            std::sort(ord.begin(), ord.end(),
                      [&](int a, int b) { return less(hd[a], ho[a], hd[b], ho[b]); });
            for (int i = 0; i < k; i++) out[r * k + i] = i < size ? hi[ord[i]] : UINT32_MAX;
        }
    }
}

void knn_search(int dim, const double* sorted, int64_t n, int N, const int32_t* cnt,
                const int32_t* start, const uint32_t* perm, int k, uint32_t* out) {
    knn_search_list(dim, sorted, n, N, cnt, start, perm, k, nullptr, n, out);
}

int ngrid(int dim, int64_t n) {
    return std::max(1, (int)std::round(std::pow((float)n / 3.1f, 1.0f / (float)dim)));
}

// periodic_mesh.cpp:63-126.  Region order of the reference: 'A' = (0,1),
// 'L' = (0,buff) shifted by +1, 'H' = (1-buff,1) shifted by -1.
const char* REG2[8] = {"LA", "AH", "HA", "AL", "LL", "LH", "HH", "HL"};
const char* REG3[26] = {"AAL", "LAA", "AHA", "HAA", "ALA", "AAH", "ALL", "LAL", "AHL",
                        "HAL", "LLA", "LHA", "HHA", "HLA", "LAH", "AHH", "HAH", "ALH",
                        "LLL", "LHL", "HHL", "HLL", "LLH", "LHH", "HHH", "HLH"};

int64_t periodic_ghosts(int dim, const double* pts, int64_t n, double buff, double* out,
                        int64_t* gid) {
    const char** reg = dim == 2 ? REG2 : REG3;
    int nreg = dim == 2 ? 8 : 26;
    int64_t ng = 0;
    std::vector<double> tmp;  // ghosts, unscaled
    for (int64_t i = 0; i < n; i++) {
        const double* p = pts + dim * i;
        for (int r = 0; r < nreg; r++) {
            bool in = true;
            double sh[3] = {0, 0, 0};
            for (int d = 0; d < dim && in; d++) {
                char c = reg[r][d];
                double lo = c == 'H' ? 1. - buff : 0.;
                double hi = c == 'L' ? buff : 1.;
                in = (p[d] > lo && p[d] < hi);
                sh[d] = c == 'L' ? 1. : (c == 'H' ? -1. : 0.);
            }
            if (!in) continue;
            for (int d = 0; d < dim; d++) tmp.push_back(p[d] + sh[d]);
            gid[ng++] = i;
        }
    }
    double scale = 1. / (1. + (2 * buff));
    for (int64_t i = 0; i < n; i++)
        for (int d = 0; d < dim; d++) out[dim * i + d] = scale * (pts[dim * i + d] - 0.5) + 0.5;
    for (int64_t g = 0; g < ng; g++)
        for (int d = 0; d < dim; d++) out[dim * (n + g) + d] = scale * (tmp[dim * g + d] - 0.5) + 0.5;
    return n + ng;
}

// voronoi.cpp:115-156 for one seed with its neighbour row; returns kcert.
// A row that holds every other seed (k == n - 1) certifies the cell by
// exhaustion; the reference would report security_radius_not_reached there.
int run_cell(int dim, int maxp, int maxt, const double* sorted, const uint32_t* row, int k, int sid,
             CellOut* out, int* status, bool exhaustive, const uint32_t* keys, bool xmode) {
    Cell c(dim, maxp, maxt, sorted, sid, keys, xmode);
    const bool flag_ties = xmode && !keys;
    int kc = 0;
    for (int v = 0; v < k; v++) {
        uint32_t z = row[v];
        c.clip((int)z);
        if (c.radius_reached(c.point((int)z))) {
            if (flag_ties && c.status == success && c.near_tie(c.point((int)z)))
                c.status = needs_exact_predicates;
            else
                kc = v + 1;
            break;
        }
        if (c.status != success) break;
    }
    if (flag_ties && c.status == success && c.near_tie(c.point((int)row[k - 1])))
        c.status = needs_exact_predicates;
    if (!c.radius_reached(c.point((int)row[k - 1]))) {
        if (exhaustive && c.status == success)
            kc = k;
        else
            c.status = security_radius_not_reached;
    }
    if (c.status != success) kc = 0;
    *status = c.status;
    if (c.status == success && out) extract(c, *out);
    return kc;
}

// Vertex-ball completion (no reference counterpart).  A cell whose
// certificate no neighbour row of k_max entries reaches (clustered input: a
// sparse cell beside a dense cluster fills its whole row with cluster points
// while its far vertices keep the security radius large) is finished by the
// empty-sphere test instead: after the first kb neighbours of its row, while
// the ball of some vertex v (centre v, radius |v - s|) strictly contains a
// point that is not one of the cell's planes, the cell is clipped by the
// nearest such point (smallest input index on ties) of the first vertex
// whose ball holds one; when no ball does, every vertex is a vertex of the
// Voronoi diagram and the cell is certified.  A ball found empty stays empty
// while its vertex lives (the points do not change), so each vertex is
// searched once.  Faces after the row follow their neighbours' (distance,
// index) order, so a cell's faces keep the neighbour-row order.
struct GridView {
    int dim, N;
    const double* sorted;
    const int32_t* cnt;
    const int32_t* start;
    const uint32_t* perm;
};

// grid rings a ball search may visit (beyond: the completion gives up)
const int BALL_RINGS = 32;

// Lower bound of the squared distance from v to any point of grid cell ix:
// boundary cells are open outwards (points beyond the unit box are clamped
// into them); a slack of 1e-9 cell widths per axis and a factor 1 - 1e-12
// keep it below the rounded distance of every point of the cell, so
// skipping a cell whose bound is at least the best distance so far never
// skips a point that would be picked.
double ball_cell_lb2(int dim, int N, const double* vv, const int* ix) {
    const double h = 1.0 / N;
    double s = 0.0;
    for (int d = 0; d < dim; d++) {
        const double lo = ix[d] == 0 ? -1e300 : ix[d] * h;
        const double hi = ix[d] == N - 1 ? 1e300 : (ix[d] + 1) * h;
        const double e = std::max(std::max(lo - vv[d], vv[d] - hi) - 1e-9 * h, 0.0);
        s += e * e;
    }
    return s * (1.0 - 1e-12);
}

// nearest point strictly inside the ball (v, r2), (distance, input index)
// order; -1: none; -2: the search would need more than BALL_RINGS rings.
// With the bounds above, the point found is the argmin over all points, and
// the ring at which the walk stops depends only on the points of the rings
// before it: any visiting order within a ring gives the same answer.
int64_t ball_point(const GridView& g, const Cell& c, const P4& v, double r2) {
    const int N = g.N, dim = g.dim;
    const double vv[3] = {v.x, v.y, v.z};
    int cc[3] = {0, 0, 0};
    for (int d = 0; d < dim; d++) cc[d] = std::max(0, std::min(N - 1, (int)std::floor(vv[d] * N)));
    double best = r2;
    int64_t pick = -1;
    const double h = 1.0 / N;
    for (int ring = 0;; ring++) {
        if (ring > 1) {
            // a point of ring r lies more than r - 1 cell widths away
            const double lb = ((ring - 1) - 1e-9) * h;
            if (lb * lb * (1.0 - 1e-12) >= best) break;
        }
        if (ring > N) break;
        if (ring > BALL_RINGS) return -2;
        const int kz0 = dim == 3 ? -ring : 0, kz1 = dim == 3 ? ring : 0;
        for (int dk = kz0; dk <= kz1; dk++)
            for (int dj = -ring; dj <= ring; dj++)
                for (int di = -ring; di <= ring; di++) {
                    if (std::max(std::abs(di), std::max(std::abs(dj), std::abs(dk))) != ring) continue;
                    const int ii = cc[0] + di, jj = cc[1] + dj, kk = cc[2] + dk;
                    if (ii < 0 || jj < 0 || kk < 0 || ii >= N || jj >= N || (dim == 3 && kk >= N)) continue;
                    const int ix[3] = {ii, jj, kk};
                    if (ball_cell_lb2(dim, N, vv, ix) >= best) continue;
                    const int64_t cell = ii + (int64_t)jj * N + (int64_t)kk * N * N;
                    for (int t = 0; t < g.cnt[cell]; t++) {
                        const int64_t q = g.start[cell] + t;
                        if (q == c.seed_id) continue;
                        const double* pq = g.sorted + (size_t)dim * q;
                        double d2 = 0;
                        for (int d = 0; d < dim; d++) d2 += (pq[d] - vv[d]) * (pq[d] - vv[d]);
                        // strictly inside; on ties the smaller input index
                        if (d2 > best || (d2 == best && (pick < 0 || g.perm[q] > g.perm[pick]))) continue;
                        bool plane = false;
                        for (int pl = 0; pl < c.nb_v && !plane; pl++) plane = c.pvid[pl] == (int)q;
                        if (plane) continue;
                        best = d2;
                        pick = q;
                    }
                }
    }
    return pick;
}

// The row's first kb neighbours, then vertex-ball completion; kcert on
// success: kmark.
int run_cell_ball(const GridView& g, int maxp, int maxt, const uint32_t* row, int kb, int sid, CellOut* out,
                  int* status, int kmark) {
    Cell c(g.dim, maxp, maxt, g.sorted, sid);
    for (int v = 0; v < kb && c.status == success; v++) c.clip((int)row[v]);
    const int first_ball = c.nb_v;
    // per vertex (sorted plane triple): the nearest point inside its ball,
    // or -1 when the ball is empty; valid while the vertex lives (a clip by p
    // removes every vertex whose ball holds p)
    std::vector<std::pair<std::array<uint8_t, 3>, int64_t>> known;
    auto key3 = [&](int t) {
        std::array<uint8_t, 3> a = {c.tp(t, 0), c.tp(t, 1), (uint8_t)(g.dim == 3 ? c.tp(t, 2) : 0)};
        std::sort(a.begin(), a.end());
        return a;
    };
    const double* sp = g.sorted + (size_t)g.dim * sid;
    int iter = 0;
    while (c.status == success) {
        // the candidate nearest the seed ((distance, index) order) among the
        // vertices' candidates: the completion follows the row's order
        int64_t pick = -1;
        double pd = 0;
        bool far = false;
        for (int t = 0; t < c.nb_t; t++) {
            const auto kt = key3(t);
            int64_t cand = -2;
            for (auto& e : known)
                if (e.first == kt) cand = e.second;
            if (cand == -2) {
                const P4 v = c.vertex(t);
                const P4 df = sub4(v, c.seed);
                cand = ball_point(g, c, v, dot3(df, df));
                known.push_back({kt, cand});
            }
            if (cand == -2) far = true;
            if (cand < 0) continue;
            const double d2 = dist2(g.dim, g.sorted + (size_t)g.dim * cand, sp);
            if (pick < 0 || d2 < pd || (d2 == pd && g.perm[cand] < g.perm[pick])) {
                pick = cand;
                pd = d2;
            }
        }
        if (far) {
            c.status = security_radius_not_reached;
            break;
        }
        if (pick < 0) break;
        if (++iter > 4 * maxp) {
            c.status = security_radius_not_reached;
            break;
        }
        c.clip((int)pick);
    }
    *status = c.status;
    if (c.status != success) return 0;
    if (out) {
        extract(c, *out);
        // faces of the planes added by the completion in (distance, index)
        // order, after the row's
        std::vector<int> idx(out->nbr.size());
        std::iota(idx.begin(), idx.end(), 0);
        auto dkey = [&](int f) {
            const int q = out->nbr[f];
            return std::make_pair(dist2(g.dim, g.sorted + (size_t)g.dim * q, sp), g.perm[q]);
        };
        // planes are stored in face order: faces from planes >= first_ball
        // come after the others
        int nrow = 0;
        for (size_t f = 0; f < out->nbr.size(); f++) {
            int pl = -1;
            for (int q = 0; q < c.nb_v; q++)
                if (c.pvid[q] == out->nbr[f]) pl = q;
            if (pl < first_ball) nrow++;
        }
        std::sort(idx.begin() + nrow, idx.end(), [&](int a, int b) { return dkey(a) < dkey(b); });
        CellOut o2 = *out;
        for (size_t f = 0; f < idx.size(); f++) {
            o2.nbr[f] = out->nbr[idx[f]];
            o2.area[f] = out->area[idx[f]];
            for (int d = 0; d < 3; d++) {
                o2.normal[3 * f + d] = out->normal[3 * idx[f] + d];
                o2.fcen[3 * f + d] = out->fcen[3 * idx[f] + d];
            }
        }
        *out = o2;
    }
    return kmark;
}

double now() { return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count(); }

// exact != 0: degenerate conflict tests are decided by exact_conflict, with
// the input index (ghosts: n + ghost index) as tie-break key.
Mesh* build(int dim, const double* in_pts, int64_t n, int boundary, double buff, int k, int maxp,
            int maxt, int restart, int k_max = 1024, int exact = 0) {
    Mesh* m = new Mesh();
    double t0 = now();
    m->dim = dim;
    std::vector<double> pts;
    int64_t ntot = n;
    if (boundary == 0) {
        int64_t cap = n;
        for (int d = 0; d < dim; d++) cap *= 3;
        pts.resize(cap * dim);
        m->ghost_ids.resize(cap);
        ntot = periodic_ghosts(dim, in_pts, n, buff, pts.data(), m->ghost_ids.data());
        pts.resize(ntot * dim);
        m->ghost_ids.resize(ntot - n);
        m->n_hydro = n;
    } else {
        pts.assign(in_pts, in_pts + n * dim);
        m->n_hydro = (boundary == 1) ? n : 0;
    }
    m->n_seeds = ntot;
    int N = ngrid(dim, ntot);
    m->N_grid = N;
    m->sorted.resize(ntot * dim);
    m->perm.resize(ntot);
    int64_t ncell = (dim == 2) ? (int64_t)N * N : (int64_t)N * N * N;
    std::vector<int32_t> cnt(ncell), start(ncell);
    grid_sort(dim, pts.data(), ntot, N, m->sorted.data(), m->perm.data(), cnt.data(), start.data());
    int kmax = k;
    if (restart) {
        kmax = k_max;
        if (kmax > ntot - 1) kmax = (int)(ntot - 1);
        if (kmax < k) kmax = k;
    }
    m->k = k;
    m->knn.resize((size_t)ntot * k);
    double t1 = now();
    knn_search(dim, m->sorted.data(), ntot, N, cnt.data(), start.data(), m->perm.data(), k, m->knn.data());
    double t2 = now();
    std::vector<CellOut> cells(ntot);
    std::vector<int> st(ntot), kc(ntot);
    // First pass: every seed with its row of length k (voronoi.cpp:115-156).
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t s = 0; s < ntot; s++)
        kc[s] = run_cell(dim, maxp, maxt, m->sorted.data(), m->knn.data() + (size_t)s * k, k, (int)s, &cells[s],
                         &st[s], k >= ntot - 1, nullptr, exact != 0);
    double t3 = now();
    // Restarts: a seed the first pass leaves uncertified is clipped again from
    // the first neighbour with a row twice as long, up to kmax.  Rows are
    // ordered by (dist^2, original id), so a longer row extends the shorter
    // one and the restarted cell is the cell a fixed k of that length gives.
    // exact != 0: afterwards the cells left with needs_exact_predicates are
    // clipped again in exact mode from the first neighbour at k, and
    // restarted the same way while uncertified.
    const uint32_t* keys = nullptr;
    for (int pass = 0; pass < (exact ? 2 : 1); pass++) {
        std::vector<int64_t> todo;
        std::vector<uint32_t> rows;
        int kk = k;
        if (pass == 1) {
            keys = m->perm.data();
            for (int64_t s = 0; s < ntot; s++)
                if (st[s] == needs_exact_predicates) todo.push_back(s);
            const int64_t nq = (int64_t)todo.size();
            if (nq == 0) break;
#pragma omp parallel for schedule(dynamic, 16)
            for (int64_t r = 0; r < nq; r++) {
                const int64_t s = todo[r];
                cells[s] = CellOut();
                kc[s] = run_cell(dim, maxp, maxt, m->sorted.data(), m->knn.data() + (size_t)s * k, k, (int)s,
                                 &cells[s], &st[s], k >= ntot - 1, keys, true);
            }
            std::vector<int64_t> left;
            for (int64_t s : todo)
                if (st[s] != success) left.push_back(s);
            todo.swap(left);
        } else {
            for (int64_t s = 0; s < ntot; s++)
                if (st[s] != success) todo.push_back(s);
        }
        if (!restart) continue;
        while (!todo.empty() && kk < kmax) {
            kk = std::min(kmax, 2 * kk);
            const int64_t nq = (int64_t)todo.size();
            rows.resize((size_t)nq * kk);
            knn_search_list(dim, m->sorted.data(), ntot, N, cnt.data(), start.data(), m->perm.data(), kk,
                            todo.data(), nq, rows.data());
#pragma omp parallel for schedule(dynamic, 16)
            for (int64_t r = 0; r < nq; r++) {
                const int64_t s = todo[r];
                cells[s] = CellOut();
                kc[s] = run_cell(dim, maxp, maxt, m->sorted.data(), rows.data() + (size_t)r * kk, kk, (int)s,
                                 &cells[s], &st[s], kk >= ntot - 1, keys, exact != 0);
            }
            std::vector<int64_t> left;
            for (int64_t s : todo)
                if (st[s] != success) left.push_back(s);
            todo.swap(left);
        }
        // seeds still uncertified at the largest row: vertex-ball completion
        if (pass == 0) {
            std::vector<int64_t> far;
            for (int64_t s = 0; s < ntot; s++)
                if (st[s] == security_radius_not_reached) far.push_back(s);
            const int64_t nf = (int64_t)far.size();
            if (nf > 0) {
                const int kb = (int)std::min<int64_t>(std::min(kmax, 64), ntot - 1);
                rows.resize((size_t)nf * kb);
                knn_search_list(dim, m->sorted.data(), ntot, N, cnt.data(), start.data(), m->perm.data(), kb,
                                far.data(), nf, rows.data());
                const GridView gv{dim, N, m->sorted.data(), cnt.data(), start.data(), m->perm.data()};
#pragma omp parallel for schedule(dynamic, 4)
                for (int64_t r = 0; r < nf; r++) {
                    const int64_t s = far[r];
                    cells[s] = CellOut();
                    kc[s] = run_cell_ball(gv, maxp, maxt, rows.data() + (size_t)r * kb, kb, (int)s, &cells[s],
                                          &st[s], kmax + 1);
                }
                m->n_ball = nf;
            }
        }
    }
    double t4 = now();
    m->times[0] = t1 - t0;
    m->times[1] = t2 - t1;
    m->times[2] = t3 - t2;
    m->times[3] = t4 - t3;
    // unpermute (voronoi.cpp:212-333) and, for periodic, unscale (periodic_mesh.cpp:146-166)
    m->boundary = boundary;
    m->scale = boundary == 0 ? 1. + (2 * buff) : 1.0;
    std::vector<int64_t>& o2s = m->o2s;
    o2s.resize(ntot);
    for (int64_t s = 0; s < ntot; s++) o2s[m->perm[s]] = s;
    m->seeds.assign(ntot * 3, 0.0);
    m->volumes.assign(ntot, 0.0);
    m->ccen.assign(ntot * 3, 0.0);
    m->fcount.assign(ntot, 0);
    m->fptr.assign(ntot, 0);
    m->status.assign(ntot, 0);
    m->kcert.assign(ntot, 0);
    int64_t cursor = 0;
    for (int64_t o = 0; o < ntot; o++) {
        int64_t s = o2s[o];
        const CellOut& c = cells[s];
        m->status[o] = st[s];
        m->kcert[o] = kc[s];
        if (st[s] != success) m->n_failed++;
        m->fptr[o] = cursor;
        m->fcount[o] = (int64_t)c.nbr.size();
        m->volumes[o] = c.volume;
        if (st[s] == success)
            for (int d = 0; d < dim; d++) m->seeds[3 * o + d] = m->sorted[dim * s + d];
        for (int d = 0; d < 3; d++) m->ccen[3 * o + d] = c.centroid[d];
        for (size_t f = 0; f < c.nbr.size(); f++) {
            int nb = c.nbr[f];
            m->nbr.push_back(nb >= 0 ? (int32_t)m->perm[nb] : nb);
            for (int d = 0; d < 3; d++) m->normal.push_back(c.normal[3 * f + d]);
            m->area.push_back(c.area[f]);
            for (int d = 0; d < 3; d++) m->fcen.push_back(c.fcen[3 * f + d]);
        }
        cursor += (int64_t)c.nbr.size();
    }
    if (boundary == 0) {
        double scale = 1. + (2 * buff);
        double vs = dim == 2 ? scale * scale : scale * scale * scale;
        double as = dim == 2 ? scale : scale * scale;
        for (int64_t i = 0; i < ntot; i++) {
            for (int d = 0; d < dim; d++) {
                m->seeds[3 * i + d] = (m->seeds[3 * i + d] - 0.5) * scale + 0.5;
                m->ccen[3 * i + d] = (m->ccen[3 * i + d] - 0.5) * scale + 0.5;
            }
            m->volumes[i] = vs * m->volumes[i];
        }
        for (size_t f = 0; f < m->area.size(); f++) {
            m->area[f] = as * m->area[f];
            for (int d = 0; d < dim; d++) m->fcen[3 * f + d] = (m->fcen[3 * f + d] - 0.5) * scale + 0.5;
        }
    }
    return m;
}

// ------------------------------------------------------------------ hydro
struct Prim {
    double rho = 0;
    double v[3] = {0, 0, 0};
    double E = 0;
};
struct Geom {
    double n[3], m[3], p[3];
};

// riemann.cpp:75-97
Geom geom_of(const double* nr) {
    Geom g;
    double nn = std::sqrt(nr[0] * nr[0] + nr[1] * nr[1] + nr[2] * nr[2]);
    g.n[0] = nr[0] / nn;
    g.n[1] = nr[1] / nn;
    g.n[2] = nr[2] / nn;
    if (g.n[0] != 0.0 || g.n[1] != 0.0) {
        g.m[0] = -g.n[1];
        g.m[1] = g.n[0];
        g.m[2] = 0.0;
    } else {
        g.m[0] = 1.0;
        g.m[1] = 0.0;
        g.m[2] = 0.0;
    }
    double mm = std::sqrt(g.m[0] * g.m[0] + g.m[1] * g.m[1] + g.m[2] * g.m[2]);
    g.m[0] /= mm;
    g.m[1] /= mm;
    g.m[2] /= mm;
    g.p[0] = g.n[1] * g.m[2] - g.n[2] * g.m[1];
    g.p[1] = g.n[2] * g.m[0] - g.n[0] * g.m[2];
    g.p[2] = g.n[0] * g.m[1] - g.n[1] * g.m[0];
    return g;
}

// riemann.cpp:47-73
void to_face(int dim, Prim& s, const Geom& g) {
    double vx = s.v[0], vy = s.v[1];
    if (dim == 2) {
        s.v[0] = vx * g.n[0] + vy * g.n[1];
        s.v[1] = vx * g.m[0] + vy * g.m[1];
    } else {
        double vz = s.v[2];
        s.v[0] = vx * g.n[0] + vy * g.n[1] + vz * g.n[2];
        s.v[1] = vx * g.m[0] + vy * g.m[1] + vz * g.m[2];
        s.v[2] = vx * g.p[0] + vy * g.p[1] + vz * g.p[2];
    }
}
void from_face(int dim, Prim& s, const Geom& g) {
    double vx = s.v[0], vy = s.v[1];
    if (dim == 2) {
        s.v[0] = vx * g.n[0] + vy * g.m[0];
        s.v[1] = vx * g.n[1] + vy * g.m[1];
    } else {
        double vz = s.v[2];
        s.v[0] = vx * g.n[0] + vy * g.m[0] + vz * g.p[0];
        s.v[1] = vx * g.n[1] + vy * g.m[1] + vz * g.p[1];
        s.v[2] = vx * g.n[2] + vy * g.m[2] + vz * g.p[2];
    }
}
// riemann.cpp:118-124
double pressure(int dim, double gam, const Prim& s) {
    if (dim == 2) return (gam - 1) * (s.E - (0.5 * s.rho * (s.v[0] * s.v[0] + s.v[1] * s.v[1])));
    return (gam - 1) * (s.E - (0.5 * s.rho * (s.v[0] * s.v[0] + s.v[1] * s.v[1] + s.v[2] * s.v[2])));
}
// riemann.cpp:99-116
Prim euler_flux(int dim, double gam, const Prim& s) {
    double P = pressure(dim, gam, s);
    Prim f;
    f.rho = s.rho * s.v[0];
    f.v[0] = s.rho * s.v[0] * s.v[0] + P;
    f.v[1] = s.rho * s.v[0] * s.v[1];
    if (dim == 3) f.v[2] = s.rho * s.v[0] * s.v[2];
    f.E = (s.E + P) * s.v[0];
    return f;
}

// riemann.cpp:7-45.  wall: the right state is the mirror of the left one in
// the face frame (normal velocity negated), so SL == -SR bit for bit.
Prim hll(int dim, double gam, Prim L, Prim R, const double* nrm, bool wall) {
    Geom g = geom_of(nrm);
    to_face(dim, L, g);
    if (wall) {
        R = L;
        R.v[0] = -L.v[0];
    } else {
        to_face(dim, R, g);
    }
    Prim fl = euler_flux(dim, gam, L), fr = euler_flux(dim, gam, R);
    double cl = std::sqrt((gam * pressure(dim, gam, L)) / L.rho);
    double cr = std::sqrt((gam * pressure(dim, gam, R)) / R.rho);
    double SL = dmin(L.v[0] - cl, R.v[0] - cr);
    double SR = dmax(L.v[0] + cl, R.v[0] + cr);
    Prim F;
    if (SL >= 0) {
        F = fl;
    } else if (SL < 0 && SR > 0) {
        F.rho = (SR * fl.rho - SL * fr.rho + SL * SR * (R.rho - L.rho)) / (SR - SL);
        for (int d = 0; d < dim; d++)
            F.v[d] = (SR * fl.v[d] - SL * fr.v[d] + SL * SR * (R.rho * R.v[d] - L.rho * L.v[d])) / (SR - SL);
        F.E = (SR * fl.E - SL * fr.E + SL * SR * (R.E - L.E)) / (SR - SL);
    } else if (SR <= 0) {
        F = fr;
    }
    from_face(dim, F, g);
    return F;
}

// ------------------------------------------------------------------ moving mesh
// No reference counterpart: the reference's mesh is built once and never
// moves (main.cpp:41-63).  Restated from the definitions of the moving-mesh
// step (Springel 2010, MNRAS 401, 791, secs. 3.3 and 4.2).  Pinned by: the
// static limit (move = 0 gives the face sums of the golden-pinned static
// step bit for bit), Galilean invariance of the face-frame flux, a
// face-velocity known answer (a rotating pair of generators), uniform flow
// staying uniform on a moving mesh, and conservation.

inline double unscale(double x, double scale) { return (x - 0.5) * scale + 0.5; }

// Velocity of the face between generators r_i and r_j (velocities w_i, w_j)
// whose centroid is fc: the mean generator velocity plus the velocity the
// bisector plane gets at fc from the pair's relative motion (its rotation
// about the midpoint of the pair).
void face_velocity(int dim, const double* ri, const double* rj, const double* wi, const double* wj,
                   const double* fc, double* w) {
    double dr[3] = {0, 0, 0}, dd = 0.0, proj = 0.0;
    for (int d = 0; d < dim; d++) {
        dr[d] = rj[d] - ri[d];
        const double mid = 0.5 * (ri[d] + rj[d]);
        dd += dr[d] * dr[d];
        proj += (wi[d] - wj[d]) * (fc[d] - mid);
    }
    for (int d = 0; d < 3; d++) w[d] = d < dim ? 0.5 * (wi[d] + wj[d]) + proj * dr[d] / dd : 0.0;
}

// Hydro index of face neighbour nb (-1: box plane, ghosts: their source).
inline int64_t hydro_of(const Mesh& m, int64_t nb) {
    if (nb < 0) return -1;
    return nb < m.n_hydro ? nb : m.ghost_ids[nb - m.n_hydro];
}

// Position of original id o in box coordinates, from the build's (scaled)
// sorted points.
inline double gen_pos(const Mesh& m, int64_t o, int d) {
    return unscale(m.sorted[(size_t)m.dim * m.o2s[o] + d], m.scale);
}

// Sum over the faces of hydro cell i of the area-integrated fluxes (mass,
// momentum[3], energy: out[5]), finite_volume_solver.cpp:53-106 for the face
// loop.  move = 0: the static flux of or_hydro_step, summed in the same
// order.  move = 1: every face moves with face_velocity() of its generators'
// velocities (wgen, or the fluid velocity when wgen is null; walls stay at
// rest); both states are boosted into the face frame, where the total energy
// per volume changes by rho (|v - w|^2 - |v|^2) / 2, the HLL problem is
// solved there and its flux transformed back: mass unchanged, momentum
// + F_m w, energy + F_p.w + |w|^2 F_m / 2.  Returns the number of box faces
// met on a periodic mesh (a ghost layer thinner than the cell).
int cell_flux(const Mesh& m, int64_t i, double gam, int flux_mode, int move, const double* rho, const double* vel,
              const double* E, const double* wgen, double* out) {
    const int dim = m.dim;
    auto gvel = [&](int64_t h, int d) { return wgen ? wgen[dim * h + d] : vel[dim * h + d]; };
    Prim si;
    si.rho = rho[i];
    si.E = E[i];
    for (int d = 0; d < dim; d++) si.v[d] = vel[dim * i + d];
    double ri[3] = {0, 0, 0}, wi[3] = {0, 0, 0};
    for (int d = 0; d < dim; d++) {
        ri[d] = gen_pos(m, i, d);
        wi[d] = gvel(i, d);
    }
    double tot[5] = {0, 0, 0, 0, 0};
    int bad = 0;
    for (int64_t q = 0; q < m.fcount[i]; q++) {
        const int64_t f = m.fptr[i] + q;
        const int64_t nb = m.nbr[f];
        const int64_t j = hydro_of(m, nb);
        const bool wall = j < 0;
        if (wall && m.boundary == 0) bad++;
        Prim sj = si;
        double w[3] = {0, 0, 0};
        if (!wall) {
            sj.rho = rho[j];
            sj.E = E[j];
            for (int d = 0; d < dim; d++) sj.v[d] = vel[dim * j + d];
            if (move) {
                double rj[3] = {0, 0, 0}, wj[3] = {0, 0, 0}, fc[3] = {0, 0, 0};
                for (int d = 0; d < dim; d++) {
                    rj[d] = gen_pos(m, nb, d);
                    wj[d] = gvel(j, d);
                    // face centroids: periodic meshes store them unscaled already
                    fc[d] = m.boundary == 0 ? m.fcen[3 * f + d] : unscale(m.fcen[3 * f + d], 1.0);
                }
                face_velocity(dim, ri, rj, wi, wj, fc, w);
            }
        }
        Prim L = si, R = sj;
        if (move && !wall) {
            double kl = 0, kr = 0, kl2 = 0, kr2 = 0;
            for (int d = 0; d < dim; d++) {
                kl += L.v[d] * L.v[d];
                kr += R.v[d] * R.v[d];
                L.v[d] -= w[d];
                R.v[d] -= w[d];
                kl2 += L.v[d] * L.v[d];
                kr2 += R.v[d] * R.v[d];
            }
            L.E += 0.5 * L.rho * (kl2 - kl);
            R.E += 0.5 * R.rho * (kr2 - kr);
        }
        double n3[3] = {m.normal[3 * f], m.normal[3 * f + 1], m.normal[3 * f + 2]};
        if (flux_mode == 1) {
            n3[0] = -n3[0];
            n3[1] = -n3[1];
            n3[2] = -n3[2];
        }
        const Prim F = hll(dim, gam, L, R, n3, wall);
        double Fm = F.rho, Fp[3] = {F.v[0], F.v[1], F.v[2]}, Fe = F.E;
        if (move && !wall) {
            double wp = 0.0, ww = 0.0;
            for (int d = 0; d < dim; d++) {
                wp += Fp[d] * w[d];
                ww += w[d] * w[d];
            }
            Fe += wp + 0.5 * ww * Fm;
            for (int d = 0; d < dim; d++) Fp[d] += Fm * w[d];
        }
        const double a = m.area[f];
        tot[0] += Fm * a;
        for (int d = 0; d < dim; d++) tot[1 + d] += Fp[d] * a;
        tot[4] += Fe * a;
    }
    for (int q = 0; q < 5; q++) out[q] = tot[q];
    return bad;
}

// Steered generator velocity of hydro cell i (regularised motion, Springel
// 2010 eq. 63 with eta = 0.25, chi = 1): w = v + chi f (c - r) / |c - r|
// with c the cell centroid, r the generator, R the radius of the cell's
// volume, f = 0 below 0.9 eta R, the sound speed above 1.1 eta R, linear in
// between.
void steered_velocity(const Mesh& m, int64_t i, double gam, const double* x, const double* rho, const double* vel,
                      const double* E, double* w) {
    const int dim = m.dim;
    const double V = m.volumes[i];
    const double R = dim == 2 ? std::sqrt(V / M_PI) : std::cbrt(3.0 * V / (4.0 * M_PI));
    Prim st;
    st.rho = rho[i];
    st.E = E[i];
    for (int d = 0; d < dim; d++) st.v[d] = vel[dim * i + d];
    const double P = pressure(dim, gam, st);
    const double cs = (P > 0.0 && st.rho > 0.0) ? std::sqrt(gam * P / st.rho) : 0.0;
    double dc[3] = {0, 0, 0}, dist = 0.0;
    for (int d = 0; d < dim; d++) {
        const double c = m.boundary == 0 ? m.ccen[3 * i + d] : unscale(m.ccen[3 * i + d], 1.0);
        dc[d] = c - x[dim * i + d];
        dist += dc[d] * dc[d];
    }
    dist = std::sqrt(dist);
    const double eta = 0.25, chi = 1.0;
    const double lo = 0.9 * eta * R, hi = 1.1 * eta * R;
    double f = 0.0;
    if (dist > hi)
        f = cs;
    else if (dist > lo)
        f = cs * (dist - lo) / (hi - lo);
    for (int d = 0; d < dim; d++) w[dim * i + d] = st.v[d] + (dist > 0.0 ? chi * f * dc[d] / dist : 0.0);
}

// CFL step of hydro cell i (finite_volume_solver.cpp:155-183): cfl R / (c +
// |v|).  With steered generators (wgen) the signal speed is the larger of
// |v| and |v - w|, and R is capped by the distance to the nearest face
// neighbour, which bounds the width of slivers.
double cell_dt(const Mesh& m, int64_t i, double gam, double cfl, const double* rho, const double* vel,
               const double* E, const double* wgen) {
    const int dim = m.dim;
    Prim s;
    s.rho = rho[i];
    s.E = E[i];
    for (int d = 0; d < dim; d++) s.v[d] = vel[dim * i + d];
    const double P = pressure(dim, gam, s);
    const double c = std::sqrt(gam * P / s.rho);
    const double V = m.volumes[i];
    double R = dim == 2 ? std::sqrt(V / M_PI) : std::cbrt(3.0 * V / (4.0 * M_PI));
    double va = std::sqrt(s.v[0] * s.v[0] + s.v[1] * s.v[1] + s.v[2] * s.v[2]);
    if (wgen) {
        const int64_t si = m.o2s[i];
        double dmin = 1e300;
        for (int64_t q = 0; q < m.fcount[i]; q++) {
            const int64_t nb = m.nbr[m.fptr[i] + q];
            if (nb < 0) continue;
            const int64_t sj = m.o2s[nb];
            double dd = 0.0;
            for (int d = 0; d < dim; d++) {
                const double t = m.sorted[(size_t)dim * sj + d] - m.sorted[(size_t)dim * si + d];
                dd += t * t;
            }
            dmin = std::min(dmin, dd);
        }
        R = std::min(R, std::sqrt(dmin) * m.scale);
        double rs = 0.0;
        for (int d = 0; d < dim; d++) rs += (s.v[d] - wgen[dim * i + d]) * (s.v[d] - wgen[dim * i + d]);
        va = std::max(va, std::sqrt(rs));
    }
    return cfl * (R / (c + va));
}

// A moving-mesh simulation: generators x, extensive FP64 state (mass,
// momentum, total energy) per hydro cell, and the current mesh.
struct Sim {
    int dim = 3, boundary = 1, k = 32, flux_mode = 1, move = 1, exact = 0;
    double buff = 0.04, gam = 1.4, cfl = 0.3;
    int64_t n = 0;
    std::vector<double> x, mass, mom, en, rho, vel, E, wgen;
    Mesh* mesh = nullptr;
    ~Sim() { delete mesh; }
    bool rebuild() {
        delete mesh;
        mesh = build(dim, x.data(), n, boundary, buff, k, 64, 96, 1, 1024, exact);
        return mesh->n_failed == 0;
    }
    // primitives from the conserved state and the current volumes
    void prim() {
        for (int64_t i = 0; i < n; i++) {
            const double V = mesh->volumes[i];
            rho[i] = mass[i] / V;
            for (int d = 0; d < dim; d++) vel[dim * i + d] = mom[dim * i + d] / mass[i];
            E[i] = en[i] / V;
        }
    }
};

// One step: primitives, (steered generator velocities), CFL dt (capped at
// dt_max > 0; a negative step from a bad state gives 0), the flux update of
// the conserved state, the drift x += w dt (wrapped into the periodic box,
// mirrored at reflective walls) and a full rebuild.  Returns the step, or
// -1 when the rebuild leaves cells uncertified.
double sim_step(Sim& S, double dt_max) {
    const int dim = S.dim;
    const Mesh& m = *S.mesh;
    S.prim();
    const bool moving = S.move != 0;
    const double* w = nullptr;
    if (moving && S.move == 2) {
        S.wgen.assign((size_t)S.n * dim, 0.0);
        for (int64_t i = 0; i < S.n; i++)
            steered_velocity(m, i, S.gam, S.x.data(), S.rho.data(), S.vel.data(), S.E.data(), S.wgen.data());
        w = S.wgen.data();
    }
    double dt = 1e100;
    for (int64_t i = 0; i < S.n; i++)
        dt = std::min(dt, cell_dt(m, i, S.gam, S.cfl, S.rho.data(), S.vel.data(), S.E.data(), w));
    if (dt < 0.0) dt = 0.0;
    if (dt_max > 0 && dt > dt_max) dt = dt_max;
    std::vector<double> fl((size_t)S.n * 5);
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < S.n; i++)
        cell_flux(m, i, S.gam, S.flux_mode, moving ? 1 : 0, S.rho.data(), S.vel.data(), S.E.data(), w,
                  fl.data() + 5 * i);
    for (int64_t i = 0; i < S.n; i++) {
        S.mass[i] -= dt * fl[5 * i];
        for (int d = 0; d < dim; d++) S.mom[dim * i + d] -= dt * fl[5 * i + 1 + d];
        S.en[i] -= dt * fl[5 * i + 4];
    }
    if (moving) {
        const double* v = w ? w : S.vel.data();
        for (int64_t i = 0; i < S.n; i++)
            for (int d = 0; d < dim; d++) {
                double p = S.x[dim * i + d] + v[dim * i + d] * dt;
                if (S.boundary == 0) {
                    p -= std::floor(p);
                    if (p >= 1.0) p = 0.0;
                } else {
                    if (p < 0.0) p = -p;
                    if (p >= 1.0) p = 2.0 - p;
                    p = std::min(std::max(p, 0.0), 0.9999999999999999);
                }
                S.x[dim * i + d] = p;
            }
    }
    return S.rebuild() ? dt : -1.0;
}

// ------------------------------------------------------------------ learned k
// Features of one seed for the k classifier (no reference counterpart: the
// reference has a compile-time _K_).  Sorted-space seed s, grid N, mean
// occupancy lam = n / N^dim.  The 3^dim cube around the seed's cell is split
// by the number m of nonzero offsets (own cell, faces, edges, corners); the
// 8 smallest FP64 squared distances to points in the cube are kept.
//   f0..f3  mean count per in-grid cell of class m=0..3, / lam (0 if none)
//   f4      in-grid cells / 3^dim
//   f5..f7  |2u-1| per axis, u = position inside the own cell (0 in 2D z)
//   f8,f9   sqrt(d2 of the 4th, 8th nearest in the cube) * N (2 if absent)
//   f10     sqrt(d2 of the nearest in the cube) * N (2 if absent)
//   f11     log2(1 + cube total / (in-grid cells * lam))
//   f12     min(distance to the unit box's faces * N, 3) / 3
//   f13     (max - min of the in-grid neighbour counts) / lam
//   f14     |sum of count * offset| / (3^dim * lam)
//   f15     f9 / f8
void kfeatures(int dim, const double* sorted, int64_t n, int N, const int32_t* cnt,
               const int32_t* start, int64_t s, float* f) {
    const double* p = sorted + (size_t)dim * s;
    const int c = cell_of(dim, N, p);
    const int ci = c % N, cj = (c / N) % N, ck = dim == 3 ? c / (N * N) : 0;
    const double ncell = (double)N * N * (dim == 3 ? N : 1);
    const double lam = (double)n / ncell;
    double best[8];
    int nb = 0;
    double csum[4] = {0, 0, 0, 0};
    int cin[4] = {0, 0, 0, 0};
    int inside = 0;
    double total = 0, gx = 0, gy = 0, gz = 0;
    int nmin = INT_MAX, nmax = -1;
    const int kr = dim == 3 ? 1 : 0;
    for (int dk = -kr; dk <= kr; dk++)
        for (int dj = -1; dj <= 1; dj++)
            for (int di = -1; di <= 1; di++) {
                const int ii = ci + di, jj = cj + dj, kk = ck + dk;
                if (ii < 0 || jj < 0 || kk < 0 || ii >= N || jj >= N || kk >= N) continue;
                const int cc = ii + jj * N + kk * N * N;
                const int m = (di != 0) + (dj != 0) + (dk != 0);
                const int ct = cnt[cc];
                inside++;
                csum[m] += ct;
                cin[m]++;
                total += ct;
                gx += (double)ct * di;
                gy += (double)ct * dj;
                gz += (double)ct * dk;
                if (m > 0) {
                    nmin = std::min(nmin, ct);
                    nmax = std::max(nmax, ct);
                }
                for (int32_t q = start[cc]; q < start[cc] + ct; q++) {
                    if (q == s) continue;
                    const double d = dist2(dim, p, sorted + (size_t)dim * q);
                    if (nb < 8) {
                        int j = nb++;
                        while (j > 0 && best[j - 1] > d) {
                            best[j] = best[j - 1];
                            j--;
                        }
                        best[j] = d;
                    } else if (d < best[7]) {
                        int j = 7;
                        while (j > 0 && best[j - 1] > d) {
                            best[j] = best[j - 1];
                            j--;
                        }
                        best[j] = d;
                    }
                }
            }
    const int cube = dim == 3 ? 27 : 9;
    for (int m = 0; m < 4; m++) f[m] = cin[m] ? (float)std::log2(1.0 + csum[m] / cin[m] / lam) : 0.0f;
    f[4] = (float)((double)inside / cube);
    const int cidx[3] = {ci, cj, ck};
    for (int d = 0; d < 3; d++) {
        if (d < dim) {
            const double u = p[d] * (double)N - (double)cidx[d];
            f[5 + d] = (float)std::fabs(2.0 * u - 1.0);
        } else {
            f[5 + d] = 0.0f;
        }
    }
    const double r4 = nb >= 4 ? std::sqrt(best[3]) * N : 2.0;
    const double r8 = nb >= 8 ? std::sqrt(best[7]) * N : 2.0;
    const double r1 = nb >= 1 ? std::sqrt(best[0]) * N : 2.0;
    f[8] = (float)r4;
    f[9] = (float)r8;
    f[10] = (float)r1;
    f[11] = (float)std::log2(1.0 + total / (inside * lam));
    double wall = 1e300;
    for (int d = 0; d < dim; d++) wall = std::min(wall, std::min(p[d], 1.0 - p[d]));
    f[12] = (float)(std::min(std::max(wall, 0.0) * N, 3.0) / 3.0);
    f[13] = nmax >= 0 ? (float)std::log2(1.0 + (double)(nmax - nmin) / lam) : 0.0f;
    f[14] = (float)std::log2(1.0 + std::sqrt(gx * gx + gy * gy + gz * gz) / (cube * lam));
    f[15] = (float)(r8 / r4);
}

}  // namespace

// ======================================================================= C-ABI
extern "C" {

int or_ngrid(int dim, int64_t n) { return ngrid(dim, n); }

int64_t or_periodic_ghosts(int dim, const double* pts, int64_t n, double buff, double* out_pts,
                           int64_t* ghost_ids) {
    return periodic_ghosts(dim, pts, n, buff, out_pts, ghost_ids);
}

int or_grid_sort(int dim, const double* pts, int64_t n, int N_grid, double* sorted, uint32_t* perm,
                 int32_t* counters, int32_t* ptrs) {
    grid_sort(dim, pts, n, N_grid, sorted, perm, counters, ptrs);
    return 0;
}

int or_kfeatures(int dim, const double* sorted, int64_t n, int N_grid, const int32_t* counters,
                 const int32_t* ptrs, float* out) {
#pragma omp parallel for schedule(static)
    for (int64_t s = 0; s < n; s++) kfeatures(dim, sorted, n, N_grid, counters, ptrs, s, out + 16 * s);
    return 0;
}

int or_knn(int dim, const double* sorted, int64_t n, int N_grid, const int32_t* counters,
           const int32_t* ptrs, const uint32_t* perm, int k, uint32_t* out) {
    knn_search(dim, sorted, n, N_grid, counters, ptrs, perm, k, out);
    return 0;
}

void* or_mesh_build(int dim, const double* pts, int64_t n, int boundary, double buff, int k,
                    int max_p, int max_t, int restart, int k_max, int exact) {
    if ((dim != 2 && dim != 3) || n < 2 || k < 1 || k > n - 1 || max_p > 255 || max_t > 255)
        return nullptr;
    return build(dim, pts, n, boundary, buff, k, max_p, max_t, restart, k_max, exact);
}

int or_exact_conflict(int dim, const double* seed, uint32_t skey, const int* box, const double* pts,
                      const uint32_t* keys) {
    if (dim != 2 && dim != 3) return -2;
    ExRow rows[4];
    for (int r = 0; r <= dim; r++) {
        rows[r].box = box[r];
        rows[r].key = keys[r];
        for (int d = 0; d < 3; d++) rows[r].p[d] = d < dim ? pts[3 * r + d] : 0.0;
    }
    return exact_conflict(dim, seed, skey, rows);
}

void or_mesh_free(void* h) { delete (Mesh*)h; }

void or_mesh_times(void* h, double* t4) {
    for (int i = 0; i < 4; i++) t4[i] = ((Mesh*)h)->times[i];
}

void or_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }

int or_get_threads(void) { return omp_get_max_threads(); }

void or_mesh_sizes(void* h, int64_t* s) {
    Mesh* m = (Mesh*)h;
    s[0] = m->n_seeds;
    s[1] = m->n_hydro;
    s[2] = (int64_t)m->nbr.size();
    s[3] = m->n_failed;
    s[4] = m->N_grid;
    s[5] = m->k;
}

int or_mesh_get(void* h, const char* name, void* dst) {
    Mesh* m = (Mesh*)h;
    auto cp = [&](const auto& v) {
        if (!v.empty()) std::memcpy(dst, v.data(), v.size() * sizeof(v[0]));
        return 0;
    };
    std::string s(name);
    if (s == "seeds") return cp(m->seeds);
    if (s == "volumes") return cp(m->volumes);
    if (s == "face_counts") return cp(m->fcount);
    if (s == "face_ptr") return cp(m->fptr);
    if (s == "neighbor_cell") return cp(m->nbr);
    if (s == "face_normal") return cp(m->normal);
    if (s == "face_area") return cp(m->area);
    if (s == "ghost_ids") return cp(m->ghost_ids);
    if (s == "status") return cp(m->status);
    if (s == "kcert") return cp(m->kcert);
    if (s == "cell_centroid") return cp(m->ccen);
    if (s == "face_centroid") return cp(m->fcen);
    if (s == "sorted_points") return cp(m->sorted);
    if (s == "permutation") return cp(m->perm);
    if (s == "knearest") return cp(m->knn);
    return -1;
}

double or_dt_cfl(void* h, double gam, double cfl, const double* rho, const double* vel, const double* E) {
    Mesh* m = (Mesh*)h;
    const int dim = m->dim;
    double mn = 1e100;
    for (int64_t i = 0; i < m->n_hydro; i++) {
        Prim s;
        s.rho = rho[i];
        s.E = E[i];
        for (int d = 0; d < dim; d++) s.v[d] = vel[dim * i + d];
        double P = pressure(dim, gam, s);
        double c = std::sqrt(gam * P / s.rho);
        double R, va;
        if (dim == 2) {
            R = std::sqrt(m->volumes[i] / M_PI);
            va = std::sqrt(s.v[0] * s.v[0] + s.v[1] * s.v[1]);
        } else {
            R = std::cbrt(3.0 * m->volumes[i] / (4.0 * M_PI));
            va = std::sqrt(s.v[0] * s.v[0] + s.v[1] * s.v[1] + s.v[2] * s.v[2]);
        }
        double dti = cfl * (R / (c + va));
        if (dti < mn) mn = dti;
    }
    return mn;
}

int or_hydro_step(void* h, double gam, double dt, int flux_mode, double* rho, double* vel, double* E) {
    Mesh* m = (Mesh*)h;
    const int dim = m->dim;
    const int64_t nh = m->n_hydro;
    std::vector<double> nr(nh), nv(nh * dim), nE(nh);
    int bad = 0;
#pragma omp parallel for reduction(+ : bad)
    for (int64_t i = 0; i < nh; i++) {
        Prim si;
        si.rho = rho[i];
        si.E = E[i];
        for (int d = 0; d < dim; d++) si.v[d] = vel[dim * i + d];
        Prim tot;
        for (int64_t j = 0; j < m->fcount[i]; j++) {
            int64_t f = m->fptr[i] + j;
            int nb = m->nbr[f];
            bool wall = nb < 0;
            Prim sj;
            if (!wall) {
                int64_t idx = nb;
                if (idx >= nh) idx = m->ghost_ids[idx - nh];
                sj.rho = rho[idx];
                sj.E = E[idx];
                for (int d = 0; d < dim; d++) sj.v[d] = vel[dim * idx + d];
            }
            double n3[3] = {m->normal[3 * f], m->normal[3 * f + 1], m->normal[3 * f + 2]};
            if (flux_mode == 1) {
                n3[0] = -n3[0];
                n3[1] = -n3[1];
                n3[2] = -n3[2];
            }
            Prim F = hll(dim, gam, si, sj, n3, wall);
            double a = m->area[f];
            tot.rho += F.rho * a;
            for (int d = 0; d < dim; d++) tot.v[d] += F.v[d] * a;
            tot.E += F.E * a;
        }
        if (m->fcount[i] == 0) bad++;
        double frac = dt / m->volumes[i];
        nr[i] = si.rho - frac * tot.rho;
        double inv = 1.0 / nr[i];
        for (int d = 0; d < dim; d++) nv[dim * i + d] = (si.rho * si.v[d] - frac * tot.v[d]) * inv;
        nE[i] = si.E - frac * tot.E;
    }
    std::memcpy(rho, nr.data(), sizeof(double) * nh);
    std::memcpy(vel, nv.data(), sizeof(double) * nh * dim);
    std::memcpy(E, nE.data(), sizeof(double) * nh);
    return bad;
}

void or_hll(int dim, double gam, const double* UL, const double* UR, const double* n, int wall, double* out) {
    Prim L, R;
    L.rho = UL[0];
    R.rho = UR[0];
    for (int d = 0; d < 3; d++) {
        L.v[d] = d < dim ? UL[1 + d] : 0.0;
        R.v[d] = d < dim ? UR[1 + d] : 0.0;
    }
    L.E = UL[4];
    R.E = UR[4];
    Prim F = hll(dim, gam, L, R, n, wall != 0);
    out[0] = F.rho;
    for (int d = 0; d < 3; d++) out[1 + d] = F.v[d];
    out[4] = F.E;
}

void or_face_velocity(int dim, const double* ri, const double* rj, const double* wi, const double* wj,
                      const double* fc, double* w) {
    face_velocity(dim, ri, rj, wi, wj, fc, w);
}

int or_cell_fluxes(void* h, double gam, int flux_mode, int move, const double* rho, const double* vel,
                   const double* E, const double* wgen, double* out) {
    const Mesh* m = (const Mesh*)h;
    int bad = 0;
#pragma omp parallel for reduction(+ : bad) schedule(dynamic, 256)
    for (int64_t i = 0; i < m->n_hydro; i++) bad += cell_flux(*m, i, gam, flux_mode, move, rho, vel, E, wgen, out + 5 * i);
    return bad;
}

void* or_sim_create(int dim, const double* pts, const double* rho, const double* vel, const double* E, int64_t n,
                    int boundary, double buff, int k, double gamma, double cfl, int flux_mode, int move, int exact) {
    if ((dim != 2 && dim != 3) || n < 2 || k < 1 || k > n - 1 || (boundary != 0 && boundary != 1) ||
        move < 0 || move > 2 || (flux_mode != 0 && flux_mode != 1))
        return nullptr;
    Sim* S = new Sim();
    S->dim = dim;
    S->boundary = boundary;
    S->k = k;
    S->flux_mode = flux_mode;
    S->move = move;
    S->exact = exact;
    S->buff = buff;
    S->gam = gamma;
    S->cfl = cfl;
    S->n = n;
    S->x.assign(pts, pts + n * dim);
    S->rho.assign(rho, rho + n);
    S->vel.assign(vel, vel + n * dim);
    S->E.assign(E, E + n);
    S->mass.resize(n);
    S->mom.resize(n * dim);
    S->en.resize(n);
    if (!S->rebuild()) {
        delete S;
        return nullptr;
    }
    for (int64_t i = 0; i < n; i++) {
        const double V = S->mesh->volumes[i];
        S->mass[i] = rho[i] * V;
        for (int d = 0; d < dim; d++) S->mom[dim * i + d] = rho[i] * vel[dim * i + d] * V;
        S->en[i] = E[i] * V;
    }
    return S;
}

double or_sim_step(void* h, double dt_max) { return sim_step(*(Sim*)h, dt_max); }

void or_sim_get(void* h, double* pts, double* rho, double* vel, double* E) {
    Sim* S = (Sim*)h;
    S->prim();
    if (pts) std::memcpy(pts, S->x.data(), sizeof(double) * S->x.size());
    if (rho) std::memcpy(rho, S->rho.data(), sizeof(double) * S->n);
    if (vel) std::memcpy(vel, S->vel.data(), sizeof(double) * S->vel.size());
    if (E) std::memcpy(E, S->E.data(), sizeof(double) * S->n);
}

void or_sim_totals(void* h, double* tot) {
    const Sim* S = (const Sim*)h;
    for (int q = 0; q < 5; q++) tot[q] = 0.0;
    for (int64_t i = 0; i < S->n; i++) {
        tot[0] += S->mass[i];
        for (int d = 0; d < S->dim; d++) tot[1 + d] += S->mom[S->dim * i + d];
        tot[4] += S->en[i];
    }
}

void* or_sim_mesh(void* h) { return ((Sim*)h)->mesh; }

void or_sim_free(void* h) { delete (Sim*)h; }

}  // extern "C"
