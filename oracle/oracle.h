/* ORACLE — test infrastructure, not product code.
 *
 * A from-scratch CPU FP64 restatement of the reference's per-timestep hot path
 * (grid KNN, meshless Voronoi clipping with security-radius certificate, HLL
 * finite-volume update and CFL reduce).  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it, and only as the checker.
 *
 * Every function cites the reference file:line it restates.  Arrays are
 * caller-owned; points are row-major [n][dim] doubles.
 *
 * Pinning (what checks each function): see the header of oracle.cpp.
 */
#ifndef PGPU_ORACLE_H
#define PGPU_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* knn.cpp:22 — grid resolution, float arithmetic as in the reference. */
int or_ngrid(int dim, int64_t n);

/* periodic_mesh.cpp:63-126 — ghost points (point-major, region order of the
 * reference) then scale into [0,1]^dim.  out_pts capacity: n*3^dim rows;
 * ghost_ids capacity n*(3^dim-1).  Returns n_total (hydro + ghosts). */
int64_t or_periodic_ghosts(int dim, const double* pts, int64_t n, double buff,
                           double* out_pts, int64_t* ghost_ids);

/* knn.cpp:119-222 — stable counting sort into the uniform grid. */
int or_grid_sort(int dim, const double* pts, int64_t n, int N_grid, double* sorted,
                 uint32_t* perm, int32_t* counters, int32_t* ptrs);

/* knn.cpp:237-347 — k nearest neighbours (sorted-index space), ascending by
 * (FP64 dist^2, original index).  out: n*k. */
int or_knn(int dim, const double* sorted, int64_t n, int N_grid, const int32_t* counters,
           const int32_t* ptrs, const uint32_t* perm, int k, uint32_t* out);

/* Features of every seed for the learned-k classifier, sorted space
 * (no reference counterpart).  out: n*16 floats; definition in oracle.cpp. */
int or_kfeatures(int dim, const double* sorted, int64_t n, int N_grid, const int32_t* counters,
                 const int32_t* ptrs, float* out);

/* Mesh build.  boundary: 0 = periodic ghosts (periodic_mesh.cpp:44), 1 = the
 * unit box as reflective walls (no ghosts, no scaling), 2 = plain
 * compute_mesh on the given points (voronoi.cpp:17).  k: neighbours per seed.
 * restart != 0: seeds that fail the certificate are recomputed from a
 * neighbour list twice as long, up to k_max.  exact != 0: degenerate
 * conflict tests (needs_exact_predicates) are decided by or_exact_conflict's
 * rule, keys = input index (ghosts: n + ghost index).
 * Returns an opaque handle or NULL. */
void* or_mesh_build(int dim, const double* pts, int64_t n, int boundary, double buff, int k,
                    int max_p, int max_t, int restart, int k_max, int exact);
/* The exact conflict test of the exact-predicate fallback (no reference
 * counterpart): vertex of rows 0..dim-1 against row dim, for the seed at
 * `seed` with tie-break key skey.  Row r: box[r] = 0 and point pts[3r..] with
 * key keys[r] (bisector), or box[r] = +-(axis + 1) for the initial cell's
 * box plane x_axis >= -1e-14 (+) / x_axis <= 1 + 1e-14 (-).  Returns +1
 * (conflict), -1 (no conflict), 0 (a subnormal or non-finite value). */
int or_exact_conflict(int dim, const double* seed, uint32_t skey, const int* box, const double* pts,
                      const uint32_t* keys);
void or_mesh_free(void* h);
/* Build stage times in seconds: [0] ghosts + grid sort, [1] KNN rows, [2]
 * first clip pass, [3] certificate restarts. */
void or_mesh_times(void* h, double* t4);
/* OpenMP threads of the following calls (for 1-core / all-core timings). */
void or_set_threads(int n);
int or_get_threads(void);

/* sizes: [0]=n_seeds [1]=n_hydro [2]=num_faces [3]=n_failed [4]=N_grid [5]=k_used */
void or_mesh_sizes(void* h, int64_t* sizes);

/* Copy one array out of a mesh handle (reference VMesh layout, original
 * order, voronoi.cpp:212-333):
 *   "seeds" f8[n_seeds*3]  "volumes" f8[n_seeds]  "face_counts" i8[n_seeds]
 *   "face_ptr" i8[n_seeds] "neighbor_cell" i4[F] "face_normal" f8[F*3]
 *   "face_area" f8[F] "ghost_ids" i8[n_seeds-n_hydro] "status" i4[n_seeds]
 *   "kcert" i4[n_seeds] (1 + index of the certifying neighbour, 0 if none)
 *   "cell_centroid" f8[n_seeds*3] "face_centroid" f8[F*3]
 *   "sorted_points" f8[n_seeds*dim] "permutation" u4[n_seeds] "knearest" u4[n_seeds*k]
 * Returns 0 on success, -1 for an unknown name. */
int or_mesh_get(void* h, const char* name, void* dst);

/* finite_volume_solver.cpp:151-186 — CFL time step over hydro cells. */
double or_dt_cfl(void* h, double gamma, double cfl, const double* rho, const double* vel,
                 const double* E);

/* finite_volume_solver.cpp:43-123 + riemann.cpp:7-124 — one static-mesh HLL
 * step, in place.  flux_mode 0 = the reference's face orientation (stored
 * inward normal, cell i as left state); 1 = outward normal (physical).
 * Box faces (neighbour -1) are reflective walls. */
int or_hydro_step(void* h, double gamma, double dt, int flux_mode, double* rho, double* vel,
                  double* E);

/* riemann.cpp:7-45 — HLL flux for states (rho, vx, vy, vz, E) across a face
 * with normal n (the frame's +x axis), lab frame result in out[5].
 * wall != 0: the right state is the mirror of the left one. */
void or_hll(int dim, double gamma, const double* UL, const double* UR, const double* n, int wall,
            double* out);

/* Moving mesh (no reference counterpart: the reference never moves its mesh,
 * main.cpp:41-63).  Velocity w[3] of the face between generators ri, rj moving
 * with wi, wj, at face centroid fc. */
void or_face_velocity(int dim, const double* ri, const double* rj, const double* wi, const double* wj,
                      const double* fc, double* w);
/* Area-integrated fluxes (mass, momentum[3], energy) summed over the faces of
 * every hydro cell, out[n_hydro*5].  move = 0: static faces (the sums of
 * or_hydro_step); move = 1: faces moving with the generators (wgen, or the
 * fluid velocity when wgen is NULL), flux solved in the face frame.  Returns
 * the number of box faces met on a periodic mesh. */
int or_cell_fluxes(void* mesh, double gamma, int flux_mode, int move, const double* rho, const double* vel,
                   const double* E, const double* wgen, double* out);
/* Moving-mesh simulation: conserved FP64 state, one step = primitives,
 * (steered velocities), CFL dt, flux update, drift, rebuild (k with
 * certificate restarts, max_p 64, max_t 96).  boundary 0 periodic, 1
 * reflective; move 0 static, 1 with the fluid, 2 regularised; exact as in
 * or_mesh_build.  NULL when the arguments are invalid or the first build
 * leaves cells uncertified. */
void* or_sim_create(int dim, const double* pts, const double* rho, const double* vel, const double* E, int64_t n,
                    int boundary, double buff, int k, double gamma, double cfl, int flux_mode, int move, int exact);
/* Returns the step taken (capped at dt_max > 0), -1 when the rebuild leaves
 * cells uncertified. */
double or_sim_step(void* sim, double dt_max);
void or_sim_get(void* sim, double* pts, double* rho, double* vel, double* E);
void or_sim_totals(void* sim, double* tot5);
/* The current mesh (owned by the simulation). */
void* or_sim_mesh(void* sim);
void or_sim_free(void* sim);

#ifdef __cplusplus
}
#endif
#endif
