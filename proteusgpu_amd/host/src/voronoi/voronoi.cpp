#include "voronoi.h"

#include <algorithm>
#include <cstdio>
#include <iostream>
#include <vector>

#if defined(USE_HDF5) && defined(WRITE_KNN_OUTPUT)
#include "../io/input.h"
#include "../io/output.h"
#endif

namespace voronoi {

namespace {

#if defined(USE_HDF5) && defined(WRITE_KNN_OUTPUT)
// knn::write_knn_output (knn.cpp:457-473) for the build's own KNN problem:
// the first icData.seedpos_dims[0] rows of the sorted points, neighbour rows
// and permutation.
void write_knn_dump(pg_mesh* dm, const pg_mesh_info_t& info) {
    const size_t ns = (size_t)info.n_seeds;
    std::vector<double> pts(ns * DIMENSION);
    std::vector<unsigned int> rows(ns * (size_t)_K_), perm(ns);
    pgpu_host::check(pg_mesh_get_knn(pgpu_host::ctx(), dm, pts.data(), rows.data(), perm.data()), "knn dump");
    const int num_points = (int)std::min<size_t>(ns, (size_t)icData.seedpos_dims[0]);
    if (!output.writeKNNFile(input.getParameter("output_knn_file"), pts.data(), rows.data(), perm.data(),
                             num_points, _K_))
        std::exit(EXIT_FAILURE);
}
#endif

VMesh* build(POINT_TYPE* pts, int64_t n, int boundary) {
    std::cout << "VORONOI: Computing Voronoi mesh..." << std::endl;
    pg_mesh_params prm = pgpu_host::mesh_params(boundary, _K_);
    pg_mesh* dm = nullptr;
    int rc = pg_mesh_compute(pgpu_host::ctx(), &prm, (const double*)pts, n, &dm);
    if (rc == PG_ERR_GRID) {  // knn.cpp:111-115
        std::cerr << "KNN: We don't support meshes with less than approx 12700 cells (3D)." << std::endl;
        std::exit(EXIT_FAILURE);
    }
    pgpu_host::check(rc, "voronoi::compute_mesh");
    pg_mesh_info_t info;
    pg_mesh_info(dm, &info);
#if defined(USE_HDF5) && defined(WRITE_KNN_OUTPUT)
    write_knn_dump(dm, info);  // voronoi.cpp:42-44
#endif
    const size_t ns = (size_t)info.n_seeds, nf = (size_t)info.num_faces;
    VMesh* m = (VMesh*)std::calloc(1, sizeof(VMesh));
    m->n_seeds = ns;
    m->num_faces = nf;
    m->n_hydro = (hsize_t)info.n_hydro;
    m->dev = dm;
    m->cell_ids = (hsize_t*)std::malloc(sizeof(hsize_t) * ns);
    m->seeds = (double3*)std::malloc(sizeof(double3) * ns);
    m->volumes = (double*)std::malloc(sizeof(double) * ns);
    m->face_counts = (hsize_t*)std::malloc(sizeof(hsize_t) * ns);
    m->face_ptr = (hsize_t*)std::malloc(sizeof(hsize_t) * ns);
    m->neighbor_cell = (int*)std::malloc(sizeof(int) * (nf ? nf : 1));
    m->face_normal = (double3*)std::malloc(sizeof(double3) * (nf ? nf : 1));
    m->face_area = (double*)std::malloc(sizeof(double) * (nf ? nf : 1));
    m->status = (int*)std::malloc(sizeof(int) * ns);
    m->kcert = (int*)std::malloc(sizeof(int) * ns);
    const size_t ng = ns - (size_t)info.n_hydro;
    m->ghost_ids = (boundary == PG_BOUNDARY_PERIODIC && ng) ? (hsize_t*)std::malloc(sizeof(hsize_t) * ng) : nullptr;
    pg_vmesh_view v = {};
    v.cell_ids = (int64_t*)m->cell_ids;
    v.seeds = (double*)m->seeds;
    v.volumes = m->volumes;
    v.face_counts = (int64_t*)m->face_counts;
    v.face_ptr = (int64_t*)m->face_ptr;
    v.neighbor_cell = m->neighbor_cell;
    v.face_normal = (double*)m->face_normal;
    v.face_area = m->face_area;
    v.ghost_ids = (int64_t*)m->ghost_ids;
    v.status = m->status;
    v.kcert = m->kcert;
    pgpu_host::check(pg_mesh_copy_out(pgpu_host::ctx(), dm, &v), "voronoi copy-out");
    // voronoi.cpp:189-196: report cells that failed
    for (size_t i = 0; i < ns; i++)
        if (m->status[i] != success)
            std::cout << "VORONOI: cell " << i << " failed with status: " << m->status[i] << std::endl;
    return m;
}

}  // namespace

VMesh* compute_mesh(POINT_TYPE* pts_data, int num_points) { return build(pts_data, num_points, PG_BOUNDARY_NONE); }

VMesh* compute_periodic_mesh(POINT_TYPE* pts_data, hsize_t num_points) {
    std::cout << "VORONOI: set up periodic mesh" << std::endl;
    return build(pts_data, (int64_t)num_points, PG_BOUNDARY_PERIODIC);
}

VMesh* compute_reflective_mesh(POINT_TYPE* pts_data, hsize_t num_points) {
    return build(pts_data, (int64_t)num_points, PG_BOUNDARY_REFLECTIVE);
}

void free_vmesh(VMesh* m) {
    if (!m) return;
    std::free(m->cell_ids);
    std::free(m->seeds);
    std::free(m->volumes);
    std::free(m->face_counts);
    std::free(m->face_ptr);
    std::free(m->neighbor_cell);
    std::free(m->face_normal);
    std::free(m->face_area);
    std::free(m->ghost_ids);
    std::free(m->status);
    std::free(m->kcert);
    if (m->dev) pg_mesh_free(&m->dev);
    std::free(m);
}

}  // namespace voronoi
