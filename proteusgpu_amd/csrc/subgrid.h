// Sub-cell geometry of the crowded-cell sub-grids (SubGrid in kernels.h),
// shared by the sub-grid build (grid.hip) and the searches (knn_core.h,
// voronoi_kernel.h).
#pragma once

#include "common.h"
#include "kernels.h"

namespace pg {

// Split of a cell of cnt points (> SUB_MIN): about 3 points per sub-cell,
// at most 4096 sub-cells.
template <int DIM>
PG_DEV int sub_split(int cnt) {
    int m = DIM == 3 ? (int)cbrt(cnt / 3.0) : (int)sqrt(cnt / 3.0);
    const int mmax = DIM == 3 ? 16 : 64;
    return m < 2 ? 2 : (m > mmax ? mmax : m);
}

// Sub-cell of a point p of grid cell ix: per axis floor((p N - i) m),
// clamped into [0, m) (points beyond the unit box are clamped into the
// boundary cells, and so into their outer sub-cells).
template <int DIM>
PG_DEV int sub_index(const double* p, const int (&ix)[3], int m, int N) {
    int j = 0, scale = 1;
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        int a = (int)floor((p[d] * (double)N - (double)ix[d]) * (double)m);
        a = a < 0 ? 0 : (a >= m ? m - 1 : a);
        j += a * scale;
        scale *= m;
    }
    return j;
}

// Lower bound of the squared distance from a query (P: its coordinates
// times N) to any point of sub-cell sx of grid cell ix: the sub-cell spans
// [i + a / m, i + (a + 1) / m) in grid units, open outwards at the grid's
// faces; a slack of 1e-9 sub-cell widths and a factor 1 - 1e-12 keep it
// below the rounded distance of every point the sub-cell holds (as in
// cell_lower_d2), so a sub-cell whose bound exceeds a search's pruning
// distance holds no point that search could take.
template <int DIM>
PG_DEV double sub_lower_d2(const double (&P)[3], const int (&ix)[3], const int (&sx)[3], int m, int N) {
    const double w = 1.0 / (double)m;
    double s = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        const double lo = (ix[d] == 0 && sx[d] == 0) ? -1e300 : (double)ix[d] + sx[d] * w;
        const double hi = (ix[d] == N - 1 && sx[d] == m - 1) ? 1e300 : (double)ix[d] + (sx[d] + 1) * w;
        const double g = fmax(fmax(lo - P[d], P[d] - hi) - 1e-9 * w, 0.0);
        s += g * g;
    }
    const double n = (double)N;
    return s / (n * n) * (1.0 - 1e-12);
}

// Sub-cell index range of a grid cell ix along each axis that a ball of
// squared radius r2 (unit coordinates) around P (grid units) can reach,
// one sub-cell of margin; the whole cell for r2 = +inf.
template <int DIM>
PG_DEV void sub_box(const double (&P)[3], const int (&ix)[3], int m, int N, double r2, int (&lo)[3], int (&hi)[3]) {
    const double rad = r2 < 1e300 ? sqrt(r2) * (double)N : 1e300;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        lo[d] = 0;
        hi[d] = d < DIM ? m - 1 : 0;
        if (d < DIM && rad < 1e300) {
            const double a0 = floor((P[d] - rad - (double)ix[d]) * (double)m) - 1.0;
            const double a1 = floor((P[d] + rad - (double)ix[d]) * (double)m) + 1.0;
            lo[d] = a0 < 0.0 ? 0 : (a0 > m - 1 ? m - 1 : (int)a0);
            hi[d] = a1 < 0.0 ? 0 : (a1 > m - 1 ? m - 1 : (int)a1);
        }
    }
}

}  // namespace pg
