// Driver with the reference's flow (src/main.cpp:32-84, begrun.cpp:14-45):
// banner, param.txt, IC file, mesh, CFL-limited HLL steps to t_end, snapshot.
//
// param.txt keys of the reference: ic_file, output_directory,
// output_mesh_file, output_knn_file.  Keys the reference hard-codes
// (main.cpp:46-47, allvars.cpp:13-14) are optional here: t_end (0.1),
// cfl (0.4), gamma (5/3), buff (0.04), boundary (periodic|reflective),
// move_mesh (false: the mesh is built once as in the reference; true: the
// generators move with the fluid and the mesh is rebuilt every step, all
// resident in HBM through pg_sim_*), regularize_mesh (false; true with
// move_mesh: generators also drift towards their cell centroids, which keeps
// cells round on long runs).  Learned k (resident runs): learned_k
// (false), mlp_weights (float32 parameter file to load; when it does not
// exist the classifier is trained on the GPU on the first mesh and, with
// mlp_weights_out, saved), mlp_train_steps (400).  perturb (0): amplitude
// of the degeneracy fallback, in units of the mean spacing (1e-6 suits
// lattice ICs, whose cells the reference leaves failed).
#include <cstdio>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "global/allvars.h"
#include "hydro/finite_volume_solver.h"
#include "io/input.h"
#include "io/output.h"
#include "knn/knn.h"
#include "voronoi/voronoi.h"

namespace {

void banner() {
    std::cout << "ProteusGPU hot path on MI355X (" << pg_version() << ")\n"
              << "moving-mesh hydrodynamics: grid KNN, meshless Voronoi, HLL\n";
#ifdef dim_2D
    std::cout << "BEGRUN: Running in 2D mode" << std::endl;
#else
    std::cout << "BEGRUN: Running in 3D mode" << std::endl;
#endif
}

int run_resident(int n, double t_end, double cfl, int boundary, bool move, bool regularize) {
    pg_sim_params sp;
    sp.mesh = pgpu_host::mesh_params(boundary, _K_);
    sp.mesh.restart = 1;
    sp.gamma = _gamma_;
    sp.cfl = cfl;
#ifdef PHYSICAL_FLUX
    sp.flux_mode = PG_FLUX_PHYSICAL;
#else
    sp.flux_mode = move ? PG_FLUX_PHYSICAL : PG_FLUX_REFERENCE;
#endif
    sp.move_mesh = !move ? PG_MOVE_STATIC : (regularize ? PG_MOVE_REGULARIZED : PG_MOVE_FLUID);
    pg_sim* sim = nullptr;
    pgpu_host::check(pg_sim_create(pgpu_host::ctx(), &sp, icData.seedpos.data(), icData.rho.data(),
                                   icData.vel.data(), icData.Energy.data(), n, &sim),
                     "pg_sim_create");
    pg_kmodel* km = nullptr;
    if (input.getBool("learned_k", false)) {
        pgpu_host::check(pg_kmodel_create(pgpu_host::ctx(), 1, &km), "pg_kmodel_create");
        const std::string wf = input.getString("mlp_weights", "");
        std::vector<float> w(PG_KMODEL_NPARAM);
        std::ifstream in(wf, std::ios::binary);
        if (!wf.empty() && in.read((char*)w.data(), sizeof(float) * w.size())) {
            pgpu_host::check(pg_kmodel_set_params(pgpu_host::ctx(), km, w.data()), "pg_kmodel_set_params");
            std::cout << "Learned k: weights from " << wf << std::endl;
        } else {
            const pg_mesh* m0 = nullptr;
            pgpu_host::check(pg_sim_mesh(sim, &m0), "pg_sim_mesh");
            double loss = 0;
            pgpu_host::check(pg_kmodel_train_mesh(pgpu_host::ctx(), km, m0, input.getInt("mlp_train_steps", 400), 4096,
                                                  3e-3, 7, &loss),
                             "pg_kmodel_train_mesh");
            std::cout << "Learned k: trained on the first mesh, loss " << loss << std::endl;
            const std::string wo = input.getString("mlp_weights_out", "");
            if (!wo.empty()) {
                pgpu_host::check(pg_kmodel_get_params(pgpu_host::ctx(), km, w.data()), "pg_kmodel_get_params");
                std::ofstream out(wo, std::ios::binary);
                out.write((const char*)w.data(), sizeof(float) * w.size());
            }
        }
        pgpu_host::check(pg_sim_set_kmodel(sim, km), "pg_sim_set_kmodel");
    }
    double t = 0;
    int step = 0;
    while (t < t_end) {
        double dt = 0;
        pgpu_host::check(pg_sim_step(pgpu_host::ctx(), sim, t_end - t, &dt), "pg_sim_step");
        t += dt;
        step++;
        if (step % 3 == 0) std::cout << "Step " << step << "  t = " << t << "  dt = " << dt << std::endl;
    }
    std::cout << "Finished after " << step << " steps at t = " << t << std::endl;
    std::vector<double> pts((size_t)n * DIMENSION), rho(n), vel((size_t)n * DIMENSION), E(n);
    pgpu_host::check(pg_sim_get(pgpu_host::ctx(), sim, pts.data(), rho.data(), vel.data(), E.data()), "pg_sim_get");
    pg_sim_free(&sim);
    if (km) pg_kmodel_free(&km);
#ifdef USE_HDF5
    // snapshot of the final generators' mesh with the final state
    VMesh* mesh = boundary == PG_BOUNDARY_PERIODIC ? voronoi::compute_periodic_mesh((POINT_TYPE*)pts.data(), n)
                                                   : voronoi::compute_reflective_mesh((POINT_TYPE*)pts.data(), n);
    primvars pv{rho.data(), (POINT_TYPE*)vel.data(), E.data()};
    if (!output.writeMeshFile(input.getString("output_mesh_file", "mesh_output.hdf5"), mesh, &pv, n)) return 1;
    voronoi::free_vmesh(mesh);
#endif
    return 0;
}

}  // namespace

int main(int argc, char* argv[]) {
    banner();
#ifdef DRY_RUN
    return 0;  // begrun.cpp:32-34
#endif
    input = InputHandler(argc > 1 ? argv[1] : "param.txt");
    if (!input.loadParameters()) {
        std::cerr << "BEGRUN: Failed to load parameters. Exiting." << std::endl;
        return EXIT_FAILURE;
    }
    try {
#ifdef USE_HDF5
        if (!input.readICFile(input.getParameter("ic_file"), icData)) return EXIT_FAILURE;
#else
        std::cerr << "BEGRUN: built without USE_HDF5; no IC reader" << std::endl;
        return EXIT_FAILURE;
#endif
        output = OutputHandler(input.getParameter("output_directory"));
        if (!output.initialize()) return EXIT_FAILURE;
        const double t_end = input.getDouble("t_end", 0.1);
        const double cfl = input.getDouble("cfl", 0.4);
        _gamma_ = input.getDouble("gamma", 5. / 3.);
        buff = input.getDouble("buff", buff);
        perturb = input.getDouble("perturb", 0.0);
        const std::string bnd = input.getString("boundary", "periodic");
        const int boundary = bnd == "reflective" ? PG_BOUNDARY_REFLECTIVE : PG_BOUNDARY_PERIODIC;
        const bool move = input.getBool("move_mesh", false);
        const int n = (int)icData.seedpos_dims[0];
        if (move || input.getBool("resident", false))
            return run_resident(n, t_end, cfl, boundary, move, input.getBool("regularize_mesh", false));

        // the reference's loop (main.cpp:38-84)
        primvars* primvar = hydro::init(n);
        VMesh* mesh = boundary == PG_BOUNDARY_PERIODIC
                          ? voronoi::compute_periodic_mesh((POINT_TYPE*)icData.seedpos.data(), n)
                          : voronoi::compute_reflective_mesh((POINT_TYPE*)icData.seedpos.data(), n);
        std::cout << "Hydro started" << std::endl;
        double t_sim = 0.0;
        int step = 0;
        while (t_sim < t_end) {
            double dt = hydro::dt_CFL(cfl, mesh, primvar);
            if (t_sim + dt > t_end) dt = t_end - t_sim;
            hydro::hydro_step(dt, mesh, primvar);
            t_sim += dt;
            step++;
            if (step % 3 == 0) std::cout << "Step " << step << "  t = " << t_sim << "  dt = " << dt << std::endl;
        }
        std::cout << "Finished after " << step << " steps at t = " << t_sim << std::endl;
        std::cout << "Hydro finished" << std::endl;
#ifdef USE_HDF5
        if (!output.writeMeshFile(input.getParameter("output_mesh_file"), mesh, primvar, n)) return EXIT_FAILURE;
#endif
        voronoi::free_vmesh(mesh);
        hydro::free_prim(&primvar);
    } catch (const std::exception& e) {
        std::cerr << e.what() << std::endl;
        return EXIT_FAILURE;
    }
    std::cout << "MAIN: Done." << std::endl;
    return 0;
}
