// knn:: interface of the reference (src/knn/knn.h:19-62) on libpgpu.
// The grid and the neighbour lists live in HBM; the get_* copy-outs return
// malloc'd host arrays the caller free()s, as in the reference.
#ifndef PGPU_HOST_KNN_H
#define PGPU_HOST_KNN_H

#include "../global/allvars.h"

typedef struct {
    int len_pts;  // number of input points
    int N_grid;   // grid resolution (knn.cpp:22)
    int k;        // neighbours per point of the last solve
    pg_knn* dev;  // device-side problem (grid, sorted points, lists)
} knn_problem;

namespace knn {

knn_problem* init(POINT_TYPE* pts, int len_pts);  // knn.h:36
void solve(knn_problem* knn);                      // knn.h:46, k = _K_
void knn_free(knn_problem** knn);                  // knn.h:54
POINT_TYPE* get_points(knn_problem* knn);          // knn.h:55 (sorted order)
unsigned int* get_knearest(knn_problem* knn);      // knn.h:56 (sorted-index space)
unsigned int* get_permutation(knn_problem* knn);   // knn.h:57 (sorted -> input index)
bool verify(knn_problem* knn, double tol = 1e-8, int max_report = 5);  // knn.h:60
// The check of verify on host arrays: n points (sorted order) and their rows
// of length k (sorted-index space), e.g. the KNN problem of a mesh build.
bool verify_rows(const double* pts, const unsigned int* idx, int n, int k, double tol = 1e-8, int max_report = 5);

}  // namespace knn

#endif
