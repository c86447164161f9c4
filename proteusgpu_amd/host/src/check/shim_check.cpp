// Check of the reference-shaped entry points that the driver itself does not
// call: knn::verify on a solved problem (and its failure on a corrupted
// row), voronoi::allocate_vmesh + compute_cells + unpermute_vmesh must give
// the mesh of a fixed-k build without restarts (the reference's
// compute_mesh, voronoi.cpp:17-65), and hydro::riemann_hll must be the
// consistent, antisymmetric HLL flux.  Prints "SHIM OK" and returns 0 on
// success.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <utility>
#include <vector>

#include "global/allvars.h"
#include "hydro/riemann.h"
#include "knn/knn.h"
#include "voronoi/voronoi.h"

namespace {

double uniform(uint64_t seed, uint64_t i) {  // SplitMix64 of (seed, counter)
    uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * 0x1p-53;
}

int fails = 0;
void expect(bool ok, const char* what) {
    if (!ok) {
        std::printf("FAIL: %s\n", what);
        fails++;
    }
}

}  // namespace

int main() {
    const int n = DIMENSION == 3 ? 20000 : 4000;
    std::vector<POINT_TYPE> pts(n);
    double* p = (double*)pts.data();
    for (int i = 0; i < n * DIMENSION; i++) p[i] = uniform(5, (uint64_t)i);

    // compute_cells on a solved KNN problem, then back to input order
    knn_problem* kp = knn::init(pts.data(), n);
    knn::solve(kp);
    // knn::verify (knn.cpp:350-422): every row against brute force
    expect(knn::verify(kp), "knn::verify of the solved problem");
    {
        // a corrupted row must fail it
        POINT_TYPE* sp = knn::get_points(kp);
        unsigned int* rows = knn::get_knearest(kp);
        std::swap(rows[0], rows[_K_ - 1]);
        expect(!knn::verify_rows((const double*)sp, rows, n, _K_, 1e-8, 1), "verify rejects a swapped row");
        std::free(sp);
        std::free(rows);
    }
    std::vector<Status> stat;
    VMesh* a = voronoi::allocate_vmesh(n, 16 * (hsize_t)n);
    voronoi::compute_cells(n, kp, stat, a);
    expect((int)stat.size() == n, "one status per seed");
    unsigned int* perm = knn::get_permutation(kp);
    voronoi::unpermute_vmesh(a, perm);
    std::free(perm);

    // the same mesh from points (no restarts)
    pg_mesh_params prm = pgpu_host::mesh_params(PG_BOUNDARY_NONE, _K_);
    prm.restart = 0;
    pg_mesh* dm = nullptr;
    pgpu_host::check(pg_mesh_compute(pgpu_host::ctx(), &prm, p, n, &dm), "pg_mesh_compute");
    pg_mesh_info_t info;
    pg_mesh_info(dm, &info);
    expect(info.num_faces == (int64_t)a->num_faces, "face count");
    std::vector<double> vol(n), area(info.num_faces);
    std::vector<int32_t> nbr(info.num_faces), st(n);
    std::vector<int64_t> fc(n);
    pg_vmesh_view v = {};
    v.volumes = vol.data();
    v.face_area = area.data();
    v.neighbor_cell = nbr.data();
    v.face_counts = fc.data();
    v.status = st.data();
    pgpu_host::check(pg_mesh_copy_out(pgpu_host::ctx(), dm, &v), "copy-out");
    bool same = info.num_faces == (int64_t)a->num_faces;
    for (int i = 0; same && i < n; i++)
        same = vol[i] == a->volumes[i] && fc[i] == (int64_t)a->face_counts[i] && st[i] == a->status[i];
    for (int64_t f = 0; same && f < info.num_faces; f++) same = nbr[f] == a->neighbor_cell[f] && area[f] == a->face_area[f];
    expect(same, "compute_cells + unpermute_vmesh equals compute_mesh without restarts");
    pg_mesh_free(&dm);

    // riemann_hll on the faces of the first certified cells
    _gamma_ = 1.4;
    int checked = 0;
    for (int i = 0; i < n && checked < 200; i++) {
        if (a->status[i] != success) continue;
        for (hsize_t j = 0; j < a->face_counts[i] && checked < 200; j++, checked++) {
            prim s;
            s.rho = 0.5 + uniform(9, 4 * checked);
            double* sv = (double*)&s.v;
            for (int d = 0; d < DIMENSION; d++) sv[d] = uniform(10, 4 * checked + d) - 0.5;
            const double P = 0.2 + uniform(11, (uint64_t)checked);
            double v2 = 0;
            for (int d = 0; d < DIMENSION; d++) v2 += sv[d] * sv[d];
            s.E = P / (_gamma_ - 1.0) + 0.5 * s.rho * v2;
            // F(U, U) = Euler flux along the stored normal
            const prim f = hydro::riemann_hll((hsize_t)i, j, s, s, a);
            const double3 nf = a->face_normal[a->face_ptr[i] + j];
            const double nn[3] = {nf.x, nf.y, nf.z};
            double un = 0;
            for (int d = 0; d < DIMENSION; d++) un += sv[d] * nn[d];
            const double* fv = (const double*)&f.v;
            bool ok = std::fabs(f.rho - s.rho * un) <= 1e-12 && std::fabs(f.E - (s.E + P) * un) <= 1e-11;
            for (int d = 0; d < DIMENSION; d++) ok = ok && std::fabs(fv[d] - (s.rho * un * sv[d] + P * nn[d])) <= 1e-11;
            expect(ok, "riemann_hll(U, U) is the Euler flux");
        }
    }
    expect(checked == 200, "enough faces checked");
    voronoi::free_vmesh(a);
    knn::knn_free(&kp);
    if (fails) return 1;
    std::printf("SHIM OK\n");
    return 0;
}
