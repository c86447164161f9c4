#include "voronoi.h"

#include <algorithm>
#include <cstdio>
#include <iostream>
#include <vector>

#if defined(USE_HDF5) && defined(WRITE_KNN_OUTPUT)
#include "../io/input.h"
#include "../io/output.h"
#endif
#ifdef VERIFY
#include "../knn/knn.h"
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

// (Re)fill the host arrays of m from the device mesh dm (copy-out in the
// mesh's current order), sized to the actual cell and face counts.
void fill_vmesh(VMesh* m, pg_mesh* dm) {
    pg_mesh_info_t info;
    pgpu_host::check(pg_mesh_info(dm, &info), "pg_mesh_info");
    const size_t ns = (size_t)info.n_seeds, nf = (size_t)info.num_faces;
    m->n_seeds = ns;
    m->num_faces = nf;
    m->n_hydro = (hsize_t)info.n_hydro;
    m->dev = dm;
    m->cell_ids = (hsize_t*)std::realloc(m->cell_ids, sizeof(hsize_t) * ns);
    m->seeds = (double3*)std::realloc(m->seeds, sizeof(double3) * ns);
    m->volumes = (double*)std::realloc(m->volumes, sizeof(double) * ns);
    m->face_counts = (hsize_t*)std::realloc(m->face_counts, sizeof(hsize_t) * ns);
    m->face_ptr = (hsize_t*)std::realloc(m->face_ptr, sizeof(hsize_t) * ns);
    m->neighbor_cell = (int*)std::realloc(m->neighbor_cell, sizeof(int) * (nf ? nf : 1));
    m->face_normal = (double3*)std::realloc(m->face_normal, sizeof(double3) * (nf ? nf : 1));
    m->face_area = (double*)std::realloc(m->face_area, sizeof(double) * (nf ? nf : 1));
    m->status = (int*)std::realloc(m->status, sizeof(int) * ns);
    m->kcert = (int*)std::realloc(m->kcert, sizeof(int) * ns);
    const size_t ng = ns - (size_t)info.n_hydro;
    std::free(m->ghost_ids);
    m->ghost_ids = (info.n_hydro && ng) ? (hsize_t*)std::malloc(sizeof(hsize_t) * ng) : nullptr;
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
}

VMesh* build(POINT_TYPE* pts, int64_t n, int boundary) {
    std::cout << "VORONOI: Computing Voronoi mesh..." << std::endl;
    pg_mesh_params prm = pgpu_host::mesh_params(boundary, _K_);
    pg_mesh* dm = nullptr;
    int rc = pg_mesh_compute(pgpu_host::ctx(), &prm, (const double*)pts, n, &dm);
    if (rc == PG_ERR_GRID) {  // knn.cpp:32-36
        std::cerr << "KNN: We don't support meshes with less than approx 12700 cells (3D)." << std::endl;
        std::exit(EXIT_FAILURE);
    }
    pgpu_host::check(rc, "voronoi::compute_mesh");
#ifdef VERIFY
    {
        // voronoi.cpp:39-41: brute-force check of the build's first-pass rows
        pg_mesh_info_t vi;
        pg_mesh_info(dm, &vi);
        const size_t ns = (size_t)vi.n_seeds;
        std::vector<double> vp(ns * DIMENSION);
        std::vector<unsigned int> vr(ns * (size_t)_K_);
        pgpu_host::check(pg_mesh_get_knn(pgpu_host::ctx(), dm, vp.data(), vr.data(), nullptr), "knn rows");
        if (!knn::verify_rows(vp.data(), vr.data(), (int)ns, _K_)) std::exit(EXIT_FAILURE);
    }
#endif
#if defined(USE_HDF5) && defined(WRITE_KNN_OUTPUT)
    pg_mesh_info_t info;
    pg_mesh_info(dm, &info);
    write_knn_dump(dm, info);  // voronoi.cpp:42-44
#endif
    VMesh* m = (VMesh*)std::calloc(1, sizeof(VMesh));
    fill_vmesh(m, dm);
    // voronoi.cpp:104-110: report cells that failed
    for (size_t i = 0; i < (size_t)m->n_seeds; i++)
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

VMesh* allocate_vmesh(hsize_t n_seeds, hsize_t initial_face_capacity) {
    VMesh* m = (VMesh*)std::calloc(1, sizeof(VMesh));
    m->n_seeds = n_seeds;
    m->cell_ids = (hsize_t*)std::malloc(sizeof(hsize_t) * (n_seeds ? n_seeds : 1));
    m->seeds = (double3*)std::malloc(sizeof(double3) * (n_seeds ? n_seeds : 1));
    m->volumes = (double*)std::malloc(sizeof(double) * (n_seeds ? n_seeds : 1));
    m->face_counts = (hsize_t*)std::calloc(n_seeds ? n_seeds : 1, sizeof(hsize_t));
    m->face_ptr = (hsize_t*)std::calloc(n_seeds ? n_seeds : 1, sizeof(hsize_t));
    m->status = (int*)std::malloc(sizeof(int) * (n_seeds ? n_seeds : 1));
    m->kcert = (int*)std::malloc(sizeof(int) * (n_seeds ? n_seeds : 1));
    const hsize_t fc = initial_face_capacity ? initial_face_capacity : 1;
    m->neighbor_cell = (int*)std::malloc(sizeof(int) * fc);
    m->face_normal = (double3*)std::malloc(sizeof(double3) * fc);
    m->face_area = (double*)std::malloc(sizeof(double) * fc);
    return m;
}

void compute_cells(int N_seedpts, knn_problem* knn, std::vector<Status>& stat, VMesh* mesh) {
    if (!knn || !mesh || N_seedpts != knn->len_pts) {
        std::cerr << "VORONOI: compute_cells needs the solved KNN problem of its seeds" << std::endl;
        std::exit(EXIT_FAILURE);
    }
    std::vector<int32_t> st((size_t)N_seedpts);
    pg_mesh* dm = nullptr;
    pgpu_host::check(pg_voronoi_compute(pgpu_host::ctx(), knn->dev, _MAX_P_, _MAX_T_, &dm, st.data()),
                     "voronoi::compute_cells");
    if (mesh->dev) pg_mesh_free(&mesh->dev);
    fill_vmesh(mesh, dm);
    stat.resize((size_t)N_seedpts);
    for (int i = 0; i < N_seedpts; i++) {
        stat[i] = (Status)st[i];
        if (stat[i] != success)  // voronoi.cpp:104-110
            std::cout << "VORONOI: cell " << i << " failed with status: " << stat[i] << std::endl;
    }
}

void unpermute_vmesh(VMesh* mesh, const unsigned int* sorted_to_original) {
    const size_t n = (size_t)mesh->n_seeds;
    std::vector<uint32_t> perm(n);
    if (!mesh->dev ||
        pg_mesh_get_knn(pgpu_host::ctx(), mesh->dev, nullptr, nullptr, perm.data()) != PG_OK ||
        !std::equal(perm.begin(), perm.end(), sorted_to_original)) {
        std::cerr << "VORONOI: unpermute_vmesh needs a mesh of compute_cells and its KNN permutation" << std::endl;
        std::exit(EXIT_FAILURE);
    }
    pgpu_host::check(pg_mesh_unpermute(pgpu_host::ctx(), mesh->dev), "voronoi::unpermute_vmesh");
    fill_vmesh(mesh, mesh->dev);
}

void vmesh_to_meshdata(VMesh* mesh, MeshCellData& md) {
    const int n = (int)mesh->n_hydro;
    md.header = MeshHeader{};
    md.header.n = n;
    md.header.k = _K_;
    md.header.nmax = _MAX_P_;
    md.seeds_dims = {(hsize_t)n, (hsize_t)DIMENSION};
    md.cell_ids.resize(n);
    md.seeds.resize((size_t)n * DIMENSION);
    md.volumes.assign(mesh->volumes, mesh->volumes + n);
    md.face_counts.resize(n);
    for (int i = 0; i < n; i++) {
        md.cell_ids[i] = (int)mesh->cell_ids[i];
        md.face_counts[i] = (int)mesh->face_counts[i];
        const double* s = (const double*)&mesh->seeds[i];
        for (int d = 0; d < DIMENSION; d++) md.seeds[(size_t)i * DIMENSION + d] = s[d];
    }
    const hsize_t nf = mesh->num_faces;
    md.faces.neighbor_cell.assign(mesh->neighbor_cell, mesh->neighbor_cell + nf);
    md.faces.area.assign(mesh->face_area, mesh->face_area + nf);
    md.faces.normal_dims = {nf, (hsize_t)DIMENSION};
    md.faces.normal.resize(nf * DIMENSION);
    for (hsize_t f = 0; f < nf; f++) {
        const double* v = (const double*)&mesh->face_normal[f];
        for (int d = 0; d < DIMENSION; d++) md.faces.normal[f * DIMENSION + d] = v[d];
    }
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
