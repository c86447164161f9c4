// Types and globals of the host layer, with the reference's names
// (src/global/allvars.h:13-217) so code written against the reference builds
// unchanged.  No HIP or torch types: the compute path is libpgpu.so behind
// include/pgpu.h.
#ifndef PGPU_HOST_ALLVARS_H
#define PGPU_HOST_ALLVARS_H

#include <cstddef>
#include <cstdint>
#include <cstdlib>

#ifdef USE_HDF5
#include "hdf5.h"
#else
typedef unsigned long long hsize_t;
#endif

#include "pgpu.h"

// allvars.h:13-20
enum Status {
    triangle_overflow = PG_TRIANGLE_OVERFLOW,
    vertex_overflow = PG_VERTEX_OVERFLOW,
    inconsistent_boundary = PG_INCONSISTENT_BOUNDARY,
    security_radius_not_reached = PG_SECURITY_RADIUS_NOT_REACHED,
    success = PG_SUCCESS,
    needs_exact_predicates = PG_NEEDS_EXACT_PREDICATES
};

struct double2 {
    double x, y;
};
struct double3 {
    double x, y, z;
};

#if defined(dim_2D)
#define DIMENSION 2
typedef double2 POINT_TYPE;
#else
#define DIMENSION 3
typedef double3 POINT_TYPE;
#endif

// Compile-time defaults of Config.sh (Config.sh:34-38)
#ifndef _K_
#define _K_ 32
#endif
#ifndef _MAX_P_
#define _MAX_P_ 64
#endif
#ifndef _MAX_T_
#define _MAX_T_ 96
#endif

// allvars.h:203-217
struct primvars {
    double* rho;
    POINT_TYPE* v;
    double* E;  // energy per unit volume
};

struct prim {
    double rho = 0;
    POINT_TYPE v = {};
    double E = 0;
};

extern double buff;     // ghost layer width (allvars.cpp:13)
extern double _gamma_;  // adiabatic index (allvars.cpp:14)
extern double perturb;  // degeneracy fallback amplitude (pg_mesh_params.perturb; 0 = off)

namespace pgpu_host {
// Process-wide context on GPU 0, created on first use.  Fatal errors print
// and exit(EXIT_FAILURE), as the reference's fatal paths do.
pg_ctx* ctx();
void check(int rc, const char* what);
pg_mesh_params mesh_params(int boundary, int k);
}  // namespace pgpu_host

#endif
