// hydro:: interface of the reference (src/hydro/finite_volume_solver.h:20-27)
// on libpgpu.  primvars stay host SoA arrays as in the reference, so each
// call crosses PCIe; the HBM-resident loop is pg_sim_* (include/pgpu.h).
#ifndef PGPU_HOST_FVS_H
#define PGPU_HOST_FVS_H

#include "../global/allvars.h"
#include "../voronoi/voronoi.h"

namespace hydro {

primvars* init(int n_hydro);          // finite_volume_solver.h:20, copies icData
void free_prim(primvars** primvar);   // :21
// :23 -- one HLL update on the static mesh.  flux orientation: the reference's
// (PG_FLUX_REFERENCE) unless PHYSICAL_FLUX is defined at compile time.
void hydro_step(double dt, const VMesh* mesh, primvars* primvar);
prim get_state_j(hsize_t i, int j, const VMesh* mesh, primvars* primvar);  // :25
double dt_CFL(double CFL, const VMesh* mesh, const primvars* primvar);     // :26

}  // namespace hydro

#endif
