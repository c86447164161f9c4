#include "finite_volume_solver.h"

#include <cstring>
#include <iostream>

#include "../io/input.h"

namespace hydro {

#ifdef PHYSICAL_FLUX
static const int kFlux = PG_FLUX_PHYSICAL;
#else
static const int kFlux = PG_FLUX_REFERENCE;
#endif

primvars* init(int n_hydro) {
    primvars* p = new primvars();
    p->rho = (double*)std::malloc(sizeof(double) * n_hydro);
    p->v = (POINT_TYPE*)std::malloc(sizeof(POINT_TYPE) * n_hydro);
    p->E = (double*)std::malloc(sizeof(double) * n_hydro);
    std::memcpy(p->rho, icData.rho.data(), sizeof(double) * n_hydro);
    std::memcpy(p->v, icData.vel.data(), sizeof(double) * DIMENSION * n_hydro);
    std::memcpy(p->E, icData.Energy.data(), sizeof(double) * n_hydro);
    std::cout << "HYDRO: Initialized primitive variables for " << n_hydro << " particles" << std::endl;
    return p;
}

void free_prim(primvars** p) {
    if (!p || !*p) return;
    std::free((*p)->rho);
    std::free((*p)->v);
    std::free((*p)->E);
    delete *p;
    *p = nullptr;
}

namespace {
pg_state* upload(const VMesh* mesh, const primvars* p) {
    pg_state* st = nullptr;
    pgpu_host::check(pg_state_create(pgpu_host::ctx(), mesh->dev, p->rho, (const double*)p->v, p->E, &st),
                     "hydro state upload");
    return st;
}
}  // namespace

void hydro_step(double dt, const VMesh* mesh, primvars* p) {
    pg_state* st = upload(mesh, p);
    pgpu_host::check(pg_hydro_step(pgpu_host::ctx(), mesh->dev, st, _gamma_, dt, kFlux), "hydro::hydro_step");
    pgpu_host::check(pg_state_copy_out(pgpu_host::ctx(), st, p->rho, (double*)p->v, p->E), "hydro copy-out");
    pg_state_free(&st);
}

double dt_CFL(double CFL, const VMesh* mesh, const primvars* p) {
    pg_state* st = upload(mesh, p);
    double dt = 0;
    pgpu_host::check(pg_dt_cfl(pgpu_host::ctx(), mesh->dev, st, _gamma_, CFL, &dt), "hydro::dt_CFL");
    pg_state_free(&st);
    return dt;
}

// finite_volume_solver.cpp:125-148 (an accessor: ghosts map to their source)
prim get_state_j(hsize_t i, int j, const VMesh* mesh, primvars* p) {
    prim s;
    int nb = mesh->neighbor_cell[mesh->face_ptr[i] + j];
    if (nb < 0) {
        std::cerr << "HYDRO: box-plane face reached a hydro cell (ghost layer too thin)" << std::endl;
        std::exit(EXIT_FAILURE);
    }
    hsize_t idx = (hsize_t)nb;
    if (idx >= mesh->n_hydro) idx = mesh->ghost_ids[idx - mesh->n_hydro];
    s.rho = p->rho[idx];
    s.v = p->v[idx];
    s.E = p->E[idx];
    return s;
}

}  // namespace hydro
