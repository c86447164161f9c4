// hydro::riemann_hll of the reference (src/hydro/riemann.h:24) on libpgpu.
#ifndef PGPU_HOST_RIEMANN_H
#define PGPU_HOST_RIEMANN_H

#include "../global/allvars.h"
#include "../voronoi/voronoi.h"

namespace hydro {

// riemann.cpp:7-45: the HLL flux across face j of cell i between the states
// st_l (cell i) and st_r (its neighbour), in the lab frame, per unit area:
// rho = mass flux, v = momentum flux, E = energy flux.  Orientation as
// hydro_step (the reference's unless PHYSICAL_FLUX).  One device call per
// face: for callers of the per-face interface; hydro_step evaluates every
// face of the mesh in one kernel.
prim riemann_hll(hsize_t i, hsize_t j, prim st_l, prim st_r, const VMesh* mesh);

}  // namespace hydro

#endif
