#include "riemann.h"

namespace hydro {

#ifdef PHYSICAL_FLUX
static const int kRiemannFlux = PG_FLUX_PHYSICAL;
#else
static const int kRiemannFlux = PG_FLUX_REFERENCE;
#endif

namespace {
void pack(const prim& s, double* u) {
    const double* v = (const double*)&s.v;
    u[0] = s.rho;
    for (int d = 0; d < DIMENSION; d++) u[1 + d] = v[d];
    u[DIMENSION + 1] = s.E;
}
}  // namespace

prim riemann_hll(hsize_t i, hsize_t j, prim st_l, prim st_r, const VMesh* mesh) {
    const double3 nf = mesh->face_normal[mesh->face_ptr[i] + j];
    const double n[3] = {nf.x, nf.y, nf.z};
    double L[DIMENSION + 2], R[DIMENSION + 2], F[DIMENSION + 2];
    pack(st_l, L);
    pack(st_r, R);
    pgpu_host::check(pg_hll_flux(pgpu_host::ctx(), DIMENSION, _gamma_, 1, L, R, n, nullptr, kRiemannFlux, F),
                     "hydro::riemann_hll");
    prim f;
    double* fv = (double*)&f.v;
    f.rho = F[0];
    for (int d = 0; d < DIMENSION; d++) fv[d] = F[1 + d];
    f.E = F[DIMENSION + 1];
    return f;
}

}  // namespace hydro
