#include "allvars.h"

#include <cstdio>

double buff = (1. / 100.) * 4;  // allvars.cpp:13
double _gamma_ = 5. / 3.;       // allvars.cpp:14
double perturb = 0.0;           // param.txt perturb (degenerate inputs such as lattices)

namespace pgpu_host {

pg_ctx* ctx() {
    static pg_ctx* c = nullptr;
    if (!c) check(pg_ctx_create(0, &c), "pg_ctx_create");
    return c;
}

void check(int rc, const char* what) {
    if (rc != PG_OK) {
        std::fprintf(stderr, "PGPU: %s failed (%d): %s\n", what, rc, pg_last_error());
        std::exit(EXIT_FAILURE);
    }
}

pg_mesh_params mesh_params(int boundary, int k) {
    pg_mesh_params p;
    pg_mesh_params_default(&p);
    p.dim = DIMENSION;
    p.boundary = boundary;
    p.buff = buff;
    p.perturb = perturb;
    p.k = k;
    p.max_p = _MAX_P_;
    p.max_t = _MAX_T_;
#ifdef CERT_RESTART
    p.restart = 1;  // uncertified seeds are recomputed at a larger k
#else
    p.restart = 0;  // reference semantics: they stay failed (voronoi.cpp:104-110)
#endif
    return p;
}

}  // namespace pgpu_host
