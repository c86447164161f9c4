#include "knn.h"

#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdio>
#include <iostream>
#include <vector>

namespace knn {

knn_problem* init(POINT_TYPE* pts, int len_pts) {
    knn_problem* k = (knn_problem*)std::malloc(sizeof(knn_problem));
    k->len_pts = len_pts;
    k->k = 0;
    k->dev = nullptr;
    int rc = pg_knn_init(pgpu_host::ctx(), DIMENSION, (const double*)pts, len_pts, &k->dev);
    if (rc == PG_ERR_GRID) {  // knn.cpp:32-36
        std::cerr << "KNN: We don't support meshes with less than approx 12700 cells (3D)." << std::endl;
        std::exit(EXIT_FAILURE);
    }
    pgpu_host::check(rc, "knn::init");
    int64_t n;
    pg_knn_info(k->dev, &n, &k->N_grid, nullptr);
    return k;
}

void solve(knn_problem* k) {
    pgpu_host::check(pg_knn_solve(pgpu_host::ctx(), k->dev, _K_), "knn::solve");
    k->k = _K_;
}

void knn_free(knn_problem** k) {
    if (!k || !*k) return;
    pg_knn_free(&(*k)->dev);
    std::free(*k);
    *k = nullptr;
}

POINT_TYPE* get_points(knn_problem* k) {
    POINT_TYPE* p = (POINT_TYPE*)std::malloc(sizeof(POINT_TYPE) * k->len_pts);
    pgpu_host::check(pg_knn_get_points(pgpu_host::ctx(), k->dev, (double*)p), "knn::get_points");
    return p;
}

unsigned int* get_knearest(knn_problem* k) {
    unsigned int* o = (unsigned int*)std::malloc(sizeof(unsigned int) * (size_t)k->len_pts * k->k);
    pgpu_host::check(pg_knn_get_knearest(pgpu_host::ctx(), k->dev, o), "knn::get_knearest");
    return o;
}

unsigned int* get_permutation(knn_problem* k) {
    unsigned int* o = (unsigned int*)std::malloc(sizeof(unsigned int) * k->len_pts);
    pgpu_host::check(pg_knn_get_permutation(pgpu_host::ctx(), k->dev, o), "knn::get_permutation");
    return o;
}

// knn.cpp:350-422: brute-force check of the distances of every list (a test
// utility of the reference, run on the host over the copied results).
// pts: n points (sorted order), idx: n rows of k (sorted-index space).
bool verify_rows(const double* pts, const unsigned int* idx, int n, int kk, double tol, int max_report) {
    auto d2 = [&](int a, int b) {
        const double* p = pts + (size_t)DIMENSION * a;
        const double* q = pts + (size_t)DIMENSION * b;
        double s = 0;
        for (int d = 0; d < DIMENSION; d++) s += (p[d] - q[d]) * (p[d] - q[d]);
        return s;
    };
    int bad = 0;
    std::vector<double> best(kk);
    for (int i = 0; i < n; i++) {
        std::fill(best.begin(), best.end(), DBL_MAX);
        for (int j = 0; j < n; j++) {
            if (j == i) continue;
            double d = d2(i, j);
            if (d >= best[kk - 1]) continue;
            int pos = kk - 1;
            while (pos > 0 && d < best[pos - 1]) {
                best[pos] = best[pos - 1];
                pos--;
            }
            best[pos] = d;
        }
        bool ok = true;
        for (int m = 0; m < kk && ok; m++) {
            unsigned int id = idx[(size_t)i * kk + m];
            ok = id != UINT_MAX && id < (unsigned int)n &&
                 std::fabs(d2(i, (int)id) - best[m]) <= tol * (1.0 + best[m]);
        }
        if (!ok && ++bad <= max_report) std::cerr << "KNN VERIFY: mismatch at point " << i << std::endl;
    }
    if (bad) {
        std::cerr << "KNN VERIFY: failed " << bad << " / " << n << " points mismatch." << std::endl;
        return false;
    }
    std::cout << "KNN VERIFY: passed - all points match brute-force." << std::endl;
    return true;
}

bool verify(knn_problem* k, double tol, int max_report) {
    if (!k || k->k == 0) {
        std::cerr << "KNN VERIFY: null or unsolved knn problem." << std::endl;
        return false;
    }
    POINT_TYPE* pts = get_points(k);
    unsigned int* idx = get_knearest(k);
    const bool ok = verify_rows((const double*)pts, idx, k->len_pts, k->k, tol, max_report);
    std::free(pts);
    std::free(idx);
    return ok;
}

}  // namespace knn
