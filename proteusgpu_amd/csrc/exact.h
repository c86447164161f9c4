// Exact conflict predicate with symbolic perturbation, for the cells whose
// filtered predicate is degenerate (the reference marks them
// needs_exact_predicates, cell.cpp:216-218, and leaves them failed,
// voronoi.cpp:104-110).
//
// The conflict test of the reference is the sign of the 4x4 (3x3 in 2D)
// determinant whose rows are the planes of the vertex and the new plane.
// Here that determinant is evaluated without rounding from the generators
// themselves:
//   * translating to the seed s (a column operation on the w column) and
//     rescaling by 2^-E (E: the smallest unit in the last place among all
//     coordinates and box bounds involved) turns every row into integers:
//       bisector of s and point p:  (-2 U, |U|^2),      U = (p - s) / 2^E
//       box plane a with sign sg:   (sg e_a, sg (s_a - c) / 2^E)
//     and changes the determinant by a positive factor only;
//   * the determinant is the sum over rows k of L_k C_k (L: last column,
//     C_k: its cofactor), in fixed-width two's complement integers.
// Exact zeros are broken by Simulation of Simplicity on the power-diagram
// weights (Edelsbrunner & Muecke 1990): point x gets weight eps^r(x), r the
// rank of its key, which adds (w_s - w_k) C_k to the determinant; the first
// non-zero coefficient in key order (the seed's is the sum of the bisector
// cofactors, a bisector point's is -C_k) decides.  The new plane's cofactor
// is the vertex's own Cramer denominator, non-zero for every vertex, so a
// decision is always reached.  Every cell evaluates the same perturbed
// diagram, so neighbouring cells agree on every face.
//
// Supported range: |coordinates| < 2^top and multiples of 2^E with
// top - E <= 175 (e.g. coordinates in [2^-120, 2), or zero); outside it the
// predicate reports "undecided" and the cell keeps needs_exact_predicates.
#pragma once

#include "common.h"

namespace pg {

// Fixed-width two's complement integer of W 64-bit limbs (little endian).
template <int W>
struct XI {
    uint64_t w[W];
};

template <int W>
PG_DEV bool xi_neg(const XI<W>& a) {
    return (a.w[W - 1] >> 63) != 0;
}

template <int W>
PG_DEV int xi_sign(const XI<W>& a) {
    if (xi_neg(a)) return -1;
    uint64_t o = 0;
#pragma unroll
    for (int i = 0; i < W; i++) o |= a.w[i];
    return o ? 1 : 0;
}

template <int W>
PG_DEV XI<W> xi_add(const XI<W>& a, const XI<W>& b) {
    XI<W> r;
    uint64_t c = 0;
#pragma unroll
    for (int i = 0; i < W; i++) {
        const uint64_t s = a.w[i] + c;
        const uint64_t c1 = s < c;
        r.w[i] = s + b.w[i];
        c = c1 + (r.w[i] < s);
    }
    return r;
}

template <int W>
PG_DEV XI<W> xi_negate(const XI<W>& a) {
    XI<W> r;
    uint64_t c = 1;
#pragma unroll
    for (int i = 0; i < W; i++) {
        r.w[i] = ~a.w[i] + c;
        c = (c && r.w[i] == 0) ? 1 : 0;
    }
    return r;
}

template <int W>
PG_DEV XI<W> xi_sub(const XI<W>& a, const XI<W>& b) {
    return xi_add(a, xi_negate(b));
}

// sign extension to a wider type
template <int W2, int W>
PG_DEV XI<W2> xi_widen(const XI<W>& a) {
    XI<W2> r;
    const uint64_t fill = xi_neg(a) ? ~0ull : 0ull;
#pragma unroll
    for (int i = 0; i < W2; i++) r.w[i] = i < W ? a.w[i] : fill;
    return r;
}

// full signed product: magnitudes multiplied limb by limb, sign applied
template <int W1, int W2>
PG_DEV XI<W1 + W2> xi_mul(const XI<W1>& a, const XI<W2>& b) {
    const bool na = xi_neg(a), nb = xi_neg(b);
    const XI<W1> ma = na ? xi_negate(a) : a;
    const XI<W2> mb = nb ? xi_negate(b) : b;
    XI<W1 + W2> r;
#pragma unroll
    for (int i = 0; i < W1 + W2; i++) r.w[i] = 0;
    for (int i = 0; i < W1; i++) {
        if (ma.w[i] == 0) continue;
        uint64_t carry = 0;
        for (int j = 0; j < W2; j++) {
            const unsigned __int128 p = (unsigned __int128)ma.w[i] * mb.w[j] + r.w[i + j] + carry;
            r.w[i + j] = (uint64_t)p;
            carry = (uint64_t)(p >> 64);
        }
        for (int k = i + W2; k < W1 + W2 && carry; k++) {
            const uint64_t s = r.w[k] + carry;
            carry = s < carry;
            r.w[k] = s;
        }
    }
    return (na != nb) ? xi_negate(r) : r;
}

// Exact decomposition of a finite double: v = m * 2^e, m a signed integer of
// at most 53 bits (m = 0 for zero).  Subnormals and non-finite values are
// outside the supported range (false).
PG_DEV bool split_double(double v, int64_t& m, int& e) {
    const uint64_t b = (uint64_t)__double_as_longlong(v);
    const int ef = (int)((b >> 52) & 0x7ff);
    const uint64_t fr = b & ((1ull << 52) - 1);
    if (ef == 0) {
        m = 0;
        e = 0;
        return fr == 0;
    }
    if (ef == 0x7ff) return false;
    m = (int64_t)(fr | (1ull << 52));
    if (b >> 63) m = -m;
    e = ef - 1075;
    return true;
}

// m * 2^sh as a W-limb integer (0 <= sh, |m| < 2^53, the result fits)
template <int W>
PG_DEV XI<W> xi_from(int64_t m, int sh) {
    XI<W> r;
#pragma unroll
    for (int i = 0; i < W; i++) r.w[i] = 0;
    const bool neg = m < 0;
    const uint64_t a = (uint64_t)(neg ? -m : m);
    const int q = sh >> 6, b = sh & 63;
    if (q < W) r.w[q] = a << b;
    if (b && q + 1 < W) r.w[q + 1] = a >> (64 - b);
    return neg ? xi_negate(r) : r;
}

// Certificate margin of exact-predicate runs (see voronoi_group).
constexpr double CERT_TIE = 1.0 + 1e-9;
// Exact mode trusts a floating-point conflict determinant only beyond
// XTOL * prod_rows (|x| + |y| + |z| + wm), wm = (|ax dx| + |ay dy| + |az dz|) / 2
// for a bisector (a = s + b, d = s - b: the terms w = -(a . d) / 2 is rounded
// from) and |w| for a box plane.  The rounding of the plane coefficients
// (relative 4u of those magnitudes) and of the 4x4 evaluation move the
// determinant by at most about 36 u times that product (u = 2^-53, the
// product bounds the permanent); XTOL = 1e-12 leaves a factor of 250.
constexpr double XTOL = 1e-12;

constexpr int XN = 3;   // normals: |N| < 2^177
constexpr int XL2 = 6;  // lifted entries, 2x2 minors: < 2^355
constexpr int XC = 9;   // 3x3 cofactors: < 2^535
constexpr int XD = 15;  // determinant: < 2^891
constexpr int XBITS = 175;  // largest supported top - E

// One row of the conflict determinant: a bisector (point p, key) or a box
// plane (axis, sign, bound c).
struct XRow {
    bool box;
    int axis;
    double sg, c;
    double p[3];
    uint32_t key;
};

// Conflict of the vertex of planes rows[0..DIM-1] with plane rows[DIM],
// for the seed at s with key skey: +1 conflict (the vertex leaves the
// cell), -1 no conflict, 0 undecided (outside the supported range).
template <int DIM>
PG_DEV int conflict_sos(const double* s, uint32_t skey, const XRow* rows) {
    constexpr int NR = DIM + 1;
    // exponent range of every value involved
    int emin = 1 << 30, top = -(1 << 30);
    bool ok = true;
    auto scan = [&](double v) {
        int64_t m = 0;
        int e = 0;
        ok = ok && split_double(v, m, e);
        if (m != 0) {
            emin = min(emin, e);
            top = max(top, e + 53);
        }
    };
    for (int d = 0; d < DIM; d++) scan(s[d]);
    for (int k = 0; k < NR; k++) {
        if (rows[k].box) scan(rows[k].c);
        else
            for (int d = 0; d < DIM; d++) scan(rows[k].p[d]);
    }
    if (!ok) return 0;
    if (top < emin) return 0;  // every value zero: no cell
    if (top - emin > XBITS) return 0;
    auto conv = [&](double v) {
        int64_t m;
        int e;
        split_double(v, m, e);
        return m == 0 ? xi_from<XN>(0, 0) : xi_from<XN>(m, e - emin);
    };
    XI<XN> S[3], N[NR][3];
    XI<XL2> L[NR];
    for (int d = 0; d < DIM; d++) S[d] = conv(s[d]);
    for (int k = 0; k < NR; k++) {
        if (rows[k].box) {
            for (int d = 0; d < DIM; d++) {
                const int64_t v = d == rows[k].axis ? (rows[k].sg > 0 ? 1 : -1) : 0;
                N[k][d] = xi_from<XN>(v, 0);
            }
            const XI<XN> w = xi_sub(S[rows[k].axis], conv(rows[k].c));
            L[k] = xi_widen<XL2>(rows[k].sg > 0 ? w : xi_negate(w));
        } else {
            XI<XL2> l = xi_from<XL2>(0, 0);
            for (int d = 0; d < DIM; d++) {
                const XI<XN> u = xi_sub(conv(rows[k].p[d]), S[d]);
                N[k][d] = xi_negate(xi_add(u, u));
                l = xi_add(l, xi_widen<XL2>(xi_mul(u, u)));
            }
            L[k] = l;
        }
    }
    // cofactors of the last column
    XI<XC> C[NR];
    for (int k = 0; k < NR; k++) {
        int o[3], q = 0;
        for (int j = 0; j < NR; j++)
            if (j != k) o[q++] = j;
        XI<XC> m;
        if (DIM == 2) {
            m = xi_widen<XC>(xi_sub(xi_mul(N[o[0]][0], N[o[1]][1]), xi_mul(N[o[0]][1], N[o[1]][0])));
        } else {
            // N_a . (N_b x N_c)
            const XI<XL2> c0 = xi_sub(xi_mul(N[o[1]][1], N[o[2]][2]), xi_mul(N[o[1]][2], N[o[2]][1]));
            const XI<XL2> c1 = xi_sub(xi_mul(N[o[1]][2], N[o[2]][0]), xi_mul(N[o[1]][0], N[o[2]][2]));
            const XI<XL2> c2 = xi_sub(xi_mul(N[o[1]][0], N[o[2]][1]), xi_mul(N[o[1]][1], N[o[2]][0]));
            m = xi_add(xi_add(xi_mul(N[o[0]][0], c0), xi_mul(N[o[0]][1], c1)), xi_mul(N[o[0]][2], c2));
        }
        // sign (-1)^(k + DIM) of the cofactor of entry (k, DIM)
        C[k] = ((k + DIM) & 1) ? xi_negate(m) : m;
    }
    XI<XD> det = xi_from<XD>(0, 0);
    for (int k = 0; k < NR; k++) det = xi_add(det, xi_widen<XD>(xi_mul(L[k], C[k])));
    const int sd = xi_sign(det);
    if (sd != 0) return sd;
    // symbolic perturbation: candidates in key order
    XI<XC> cs = xi_from<XC>(0, 0);
    for (int k = 0; k < NR; k++)
        if (!rows[k].box) cs = xi_add(cs, C[k]);
    uint32_t last = 0;
    bool first = true;
    for (int it = 0; it < NR + 1; it++) {
        // next smallest key above `last`
        int best = -2;
        uint32_t bk = 0xffffffffu;
        if ((first || skey > last) && skey <= bk) {
            best = -1;
            bk = skey;
        }
        for (int k = 0; k < NR; k++) {
            if (rows[k].box) continue;
            const uint32_t kk = rows[k].key;
            if ((first || kk > last) && kk < bk) {
                best = k;
                bk = kk;
            }
        }
        if (best == -2) break;
        first = false;
        last = bk;
        const int sg = best < 0 ? xi_sign(cs) : -xi_sign(C[best]);
        if (sg != 0) return sg;
    }
    return 0;
}

}  // namespace pg
