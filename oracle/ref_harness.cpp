// Test-infrastructure harness around the UNMODIFIED reference sources.
//
// It is compiled (by oracle/Makefile, target `ref`) together with the
// reference's own translation units where they lie under the reference tree,
// and writes golden vectors for tests/golden.  Nothing in the product links
// or loads it; it never travels to a GPU host as a dependency of the product.
//
// Reference entry points driven here:
//   knn::init / knn::solve / get_points / get_knearest / get_permutation
//                                              src/knn/knn.h:36-57
//   voronoi::compute_mesh                      src/voronoi/voronoi.h:68
//   voronoi::compute_periodic_mesh             src/voronoi/periodic_mesh.h:22
//   hydro::init / dt_CFL / hydro_step          src/hydro/finite_volume_solver.h:20-27
//
//   InputHandler::loadParameters / readICFile  src/io/input.h:30-54
//   OutputHandler::writeMeshFile / writeKNNFile src/io/output.h:74-77
//   voronoi::vmesh_to_meshdata                 src/voronoi/voronoi.h:81
//
// Modes knn | mesh | periodic: input file (little endian) int32 dim, int64 n,
//   f64 pts[n*dim], then (periodic only) f64 rho[n], f64 vel[n*dim], f64 E[n].
//   Output: <outdir>/<name>.bin raw arrays plus <outdir>/index.txt lines
//   "<name> <dtype> <count>".
// Mode files <param.txt> <gamma> <cfl> <t_end> <buff>: the reference's own
//   main loop (main.cpp:38-84) with its IC reader and file writers: reads
//   ic_file, builds the periodic mesh (with WRITE_KNN_OUTPUT the build writes
//   output_knn_file), steps to t_end and writes output_mesh_file.
#include <cstdio>
#include <cstdlib>
#include <cstdint>
#include <string>
#include <vector>

#include "global/allvars.h"
#include "io/input.h"
#include "io/output.h"
#include "knn/knn.h"
#include "voronoi/voronoi.h"
#include "voronoi/periodic_mesh.h"
#include "hydro/finite_volume_solver.h"

namespace {

std::string g_out;
FILE* g_index = nullptr;

template <class T>
void dump(const char* name, const char* dtype, const T* data, size_t count) {
    std::string path = g_out + "/" + name + ".bin";
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) { std::perror(path.c_str()); std::exit(2); }
    if (count) std::fwrite(data, sizeof(T), count, f);
    std::fclose(f);
    std::fprintf(g_index, "%s %s %zu\n", name, dtype, count);
}

void dump_mesh(const VMesh* m) {
    hsize_t n = m->n_seeds, nf = m->num_faces;
    std::vector<int64_t> cid(n), fc(n), fp(n);
    std::vector<double> seeds(3 * n), nrm(3 * nf);
    for (hsize_t i = 0; i < n; i++) {
        cid[i] = (int64_t)m->cell_ids[i];
        fc[i] = (int64_t)m->face_counts[i];
        fp[i] = (int64_t)m->face_ptr[i];
        seeds[3 * i] = m->seeds[i].x;
        seeds[3 * i + 1] = m->seeds[i].y;
        seeds[3 * i + 2] = m->seeds[i].z;
    }
    for (hsize_t f = 0; f < nf; f++) {
        nrm[3 * f] = m->face_normal[f].x;
        nrm[3 * f + 1] = m->face_normal[f].y;
        nrm[3 * f + 2] = m->face_normal[f].z;
    }
    int64_t hdr[3] = {(int64_t)n, (int64_t)nf, (int64_t)m->n_hydro};
    dump("mesh_header", "i8", hdr, 3);
    dump("cell_ids", "i8", cid.data(), n);
    dump("seeds", "f8", seeds.data(), 3 * n);
    dump("volumes", "f8", m->volumes, n);
    dump("face_counts", "i8", fc.data(), n);
    dump("face_ptr", "i8", fp.data(), n);
    dump("neighbor_cell", "i4", m->neighbor_cell, nf);
    dump("face_normal", "f8", nrm.data(), 3 * nf);
    dump("face_area", "f8", m->face_area, nf);
    if (m->ghost_ids && n > m->n_hydro) {
        std::vector<int64_t> g(n - m->n_hydro);
        for (hsize_t i = 0; i < n - m->n_hydro; i++) g[i] = (int64_t)m->ghost_ids[i];
        dump("ghost_ids", "i8", g.data(), g.size());
    }
}

#ifdef USE_HDF5
int run_files(int argc, char** argv) {
    if (argc < 7) {
        std::fprintf(stderr, "usage: %s files <param.txt> <gamma> <cfl> <t_end> <buff>\n", argv[0]);
        return 2;
    }
    input = InputHandler(argv[2]);
    if (!input.loadParameters()) return 2;
    if (!input.readICFile(input.getParameter("ic_file"), icData)) return 3;
    _gamma_ = std::atof(argv[3]);
    const double cfl = std::atof(argv[4]), t_end = std::atof(argv[5]);
    buff = std::atof(argv[6]);
    output = OutputHandler(input.getParameter("output_directory"));
    if (!output.initialize()) return 2;
    const int n = (int)icData.seedpos_dims[0];
    primvars* pv = hydro::init(n);
    VMesh* mesh = voronoi::compute_periodic_mesh((POINT_TYPE*)icData.seedpos.data(), (hsize_t)n);
    double t = 0.0;
    while (t < t_end) {
        double dt = hydro::dt_CFL(cfl, mesh, pv);
        if (t + dt > t_end) dt = t_end - t;
        hydro::hydro_step(dt, mesh, pv);
        t += dt;
    }
    MeshCellData md;
    voronoi::vmesh_to_meshdata(mesh, md);
    const bool ok = output.writeMeshFile(input.getParameter("output_mesh_file"), md, pv, n);
    voronoi::free_vmesh(mesh);
    hydro::free_prim(&pv);
    return ok ? 0 : 3;
}
#endif

}  // namespace

int main(int argc, char** argv) {
#ifdef USE_HDF5
    if (argc > 1 && std::string(argv[1]) == "files") return run_files(argc, argv);
#endif
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s knn|mesh|periodic <input.bin> <outdir> [gamma cfl nsteps [buff]]\n", argv[0]);
        return 2;
    }
    std::string mode = argv[1];
    g_out = argv[3];
    FILE* in = std::fopen(argv[2], "rb");
    if (!in) { std::perror(argv[2]); return 2; }
    int32_t dim = 0;
    int64_t n = 0;
    if (std::fread(&dim, 4, 1, in) != 1 || std::fread(&n, 8, 1, in) != 1) return 2;
    if (dim != DIMENSION) {
        std::fprintf(stderr, "input dim %d but harness built for %d\n", dim, DIMENSION);
        return 2;
    }
    std::vector<double> pts(n * dim);
    if (std::fread(pts.data(), 8, pts.size(), in) != pts.size()) return 2;
    g_index = std::fopen((g_out + "/index.txt").c_str(), "w");
    if (!g_index) { std::perror("index"); return 2; }
    int32_t kk = _K_;
    dump("K", "i4", &kk, 1);

    if (mode == "knn") {
        knn_problem* kp = knn::init((POINT_TYPE*)pts.data(), (int)n);
        knn::solve(kp);
        POINT_TYPE* sp = knn::get_points(kp);
        unsigned int* kn = knn::get_knearest(kp);
        unsigned int* pm = knn::get_permutation(kp);
        int32_t ng = kp->N_grid;
        dump("N_grid", "i4", &ng, 1);
        dump("sorted_points", "f8", (double*)sp, (size_t)n * dim);
        dump("knearest", "u4", kn, (size_t)n * _K_);
        dump("permutation", "u4", pm, (size_t)n);
        free(sp); free(kn); free(pm);
        knn::knn_free(&kp);
    } else if (mode == "mesh") {
        VMesh* m = voronoi::compute_mesh((POINT_TYPE*)pts.data(), (int)n);
        dump_mesh(m);
        voronoi::free_vmesh(m);
    } else if (mode == "periodic") {
        double gam = argc > 4 ? std::atof(argv[4]) : 5.0 / 3.0;
        double cfl = argc > 5 ? std::atof(argv[5]) : 0.4;
        int nsteps = argc > 6 ? std::atoi(argv[6]) : 0;
        if (argc > 7) buff = std::atof(argv[7]);  // global ghost-layer width, allvars.cpp:13
        icData.rho.resize(n);
        icData.vel.resize(n * dim);
        icData.Energy.resize(n);
        if (std::fread(icData.rho.data(), 8, n, in) != (size_t)n) return 2;
        if (std::fread(icData.vel.data(), 8, n * dim, in) != (size_t)(n * dim)) return 2;
        if (std::fread(icData.Energy.data(), 8, n, in) != (size_t)n) return 2;
        _gamma_ = gam;
        primvars* pv = hydro::init((int)n);
        VMesh* m = voronoi::compute_periodic_mesh((POINT_TYPE*)pts.data(), (hsize_t)n);
        dump_mesh(m);
        std::vector<double> dts;
        for (int s = 0; s < nsteps; s++) {
            double dt = hydro::dt_CFL(cfl, m, pv);
            dts.push_back(dt);
            hydro::hydro_step(dt, m, pv);
        }
        dump("dts", "f8", dts.data(), dts.size());
        dump("rho", "f8", pv->rho, n);
        dump("vel", "f8", (double*)pv->v, n * dim);
        dump("E", "f8", pv->E, n);
        voronoi::free_vmesh(m);
        hydro::free_prim(&pv);
    } else {
        std::fprintf(stderr, "unknown mode %s\n", mode.c_str());
        return 2;
    }
    std::fclose(g_index);
    std::fclose(in);
    return 0;
}
