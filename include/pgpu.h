/*
 * pgpu.h — C ABI of the MI355X moving-mesh hot path (libpgpu.so).
 *
 * Plain pointers and sizes only; no HIP, HDF5 or torch types.  Points are
 * row-major [n][dim] doubles, velocities [n][dim], ids int64/uint32 as noted.
 * Every call returns PG_OK (0) or a negative PG_ERR_*; pg_last_error() gives a
 * thread-local message.  Nothing here calls exit() or throws.  Handles are
 * owned by the library; host arrays by the caller.  One context per host
 * thread; work is stream ordered on the context's stream and synchronous at
 * every *_get_* / *_copy_out call.
 *
 * Each entry point names the reference interface it replaces (file:line in
 * lucas56098/ProteusGPU).  The C++ shims in proteusgpu_amd/host/src keep the
 * reference's namespaces and signatures on top of these calls.
 */
#ifndef PGPU_H
#define PGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PG_OK 0
#define PG_ERR_ARG (-1)      /* invalid argument                                  */
#define PG_ERR_HIP (-2)      /* HIP runtime error (message in pg_last_error)       */
#define PG_ERR_NOMEM (-3)    /* device or host allocation failed                  */
#define PG_ERR_GRID (-4)     /* N_grid < 16 (knn.cpp:111-115 exits here)           */
#define PG_ERR_CELLS (-5)    /* cells left uncertified after every restart         */
#define PG_ERR_NODEV (-6)    /* no usable GPU                                      */

/* Per-cell status codes: same numbering as the reference (allvars.h:13-20). */
#define PG_TRIANGLE_OVERFLOW 0
#define PG_VERTEX_OVERFLOW 1
#define PG_INCONSISTENT_BOUNDARY 2
#define PG_SECURITY_RADIUS_NOT_REACHED 3
#define PG_SUCCESS 4
#define PG_NEEDS_EXACT_PREDICATES 5

/* Boundary treatment of the unit box. */
#define PG_BOUNDARY_PERIODIC 0   /* ghost copies of width buff (periodic_mesh.cpp:44-184) */
#define PG_BOUNDARY_REFLECTIVE 1 /* box planes are walls, no ghosts (no reference counterpart) */
#define PG_BOUNDARY_NONE 2       /* plain voronoi::compute_mesh on the points (voronoi.cpp:17) */

/* HLL face orientation (riemann.cpp:7-45). */
#define PG_FLUX_REFERENCE 0 /* stored inward normal, cell i as left state, as the reference */
#define PG_FLUX_PHYSICAL 1  /* outward normal: the physically consistent upwind flux */

typedef struct pg_ctx pg_ctx;
typedef struct pg_knn pg_knn;
typedef struct pg_mesh pg_mesh;
typedef struct pg_state pg_state;
typedef struct pg_sim pg_sim;

/* ------------------------------------------------------------------ context */
const char* pg_version(void);
const char* pg_last_error(void);
int pg_device_count(int* count);
/* Replaces the reference's gpu* shim (allvars.h:92-129): one device + stream. */
int pg_ctx_create(int device, pg_ctx** out);
int pg_ctx_destroy(pg_ctx* ctx);
int pg_ctx_sync(pg_ctx* ctx);

/* ---------------------------------------------------------------------- KNN */
/* knn::init (knn.h:36, knn.cpp:16-117): copy points, N_grid, stable
 * counting sort into the grid.  PG_ERR_GRID when N_grid < 16. */
int pg_knn_init(pg_ctx* ctx, int dim, const double* pts, int64_t n, pg_knn** out);
/* knn::solve (knn.h:46, knn.cpp:225-315): k nearest neighbours of every
 * point, sorted-index space, ascending (FP64 dist^2, original index). */
int pg_knn_solve(pg_ctx* ctx, pg_knn* knn, int k);
int pg_knn_info(const pg_knn* knn, int64_t* n, int* N_grid, int* k);
/* knn::get_points / get_knearest / get_permutation (knn.h:55-57). */
int pg_knn_get_points(pg_ctx* ctx, const pg_knn* knn, double* out);        /* n*dim */
int pg_knn_get_knearest(pg_ctx* ctx, const pg_knn* knn, uint32_t* out);    /* n*k   */
int pg_knn_get_permutation(pg_ctx* ctx, const pg_knn* knn, uint32_t* out); /* n     */
/* knn::knn_free (knn.h:54): frees and nulls the handle. */
int pg_knn_free(pg_knn** knn);

/* ------------------------------------------------------------------ Voronoi */
typedef struct {
    int dim;       /* 2 or 3 (Config.sh dim_2D / dim_3D)                          */
    int boundary;  /* PG_BOUNDARY_*                                              */
    double buff;   /* ghost layer width (allvars.cpp:13, default 0.04)           */
    int k;         /* neighbours per seed (Config.sh _K_)                        */
    int max_p;     /* plane capacity per cell (Config.sh _MAX_P_, <= 64)         */
    int max_t;     /* vertex capacity per cell (Config.sh _MAX_T_, <= 128)       */
    int restart;   /* nonzero: recompute uncertified seeds at larger k           */
    int k_max;     /* largest k a restart may use (default 256)                  */
} pg_mesh_params;

void pg_mesh_params_default(pg_mesh_params* p);

typedef struct {
    int64_t n_seeds;    /* hydro + ghost cells                                   */
    int64_t n_hydro;    /* active cells (periodic: the input points)             */
    int64_t num_faces;  /* faces over all cells (VMesh.num_faces)                */
    int64_t n_failed;   /* cells not certified after every restart               */
    int64_t n_restarted;/* seeds recomputed at a larger k                        */
    int N_grid;
    int k_used;         /* largest k any seed needed                             */
} pg_mesh_info_t;

/* voronoi::compute_mesh (voronoi.h:68, voronoi.cpp:17-65) for boundary NONE,
 * voronoi::compute_periodic_mesh (periodic_mesh.h:22) for PERIODIC; REFLECTIVE
 * has no reference counterpart.  Points are host arrays. */
int pg_mesh_compute(pg_ctx* ctx, const pg_mesh_params* params, const double* pts, int64_t n,
                    pg_mesh** out);
int pg_mesh_info(const pg_mesh* mesh, pg_mesh_info_t* info);

/* The reference VMesh (voronoi.h:22-41) after unpermute_vmesh
 * (voronoi.cpp:212-333) and, for periodic meshes, the unscale of
 * periodic_mesh.cpp:170-190.  Caller allocates every array (sizes from
 * pg_mesh_info); any pointer may be NULL to skip that array. */
typedef struct {
    int64_t* cell_ids;      /* n_seeds                                         */
    double* seeds;          /* n_seeds*3 (z = 0 in 2D, as VMesh.seeds)          */
    double* volumes;        /* n_seeds                                         */
    int64_t* face_counts;   /* n_seeds                                         */
    int64_t* face_ptr;      /* n_seeds                                         */
    int32_t* neighbor_cell; /* num_faces, original ids, -1 = box plane         */
    double* face_normal;    /* num_faces*3, unit (s_i - s_j) as the reference   */
    double* face_area;      /* num_faces                                        */
    int64_t* ghost_ids;     /* n_seeds - n_hydro                               */
    int32_t* status;        /* n_seeds, PG_* status codes                       */
    int32_t* kcert;         /* n_seeds, minimal certified k (1 + index)         */
    double* cell_centroid;  /* n_seeds*3 (no reference counterpart)             */
    double* face_centroid;  /* num_faces*3 (no reference counterpart)           */
} pg_vmesh_view;

int pg_mesh_copy_out(pg_ctx* ctx, const pg_mesh* mesh, pg_vmesh_view* out);
/* voronoi::free_vmesh (voronoi.h:65). */
int pg_mesh_free(pg_mesh** mesh);

/* -------------------------------------------------------------------- hydro */
/* hydro::init (finite_volume_solver.h:20, .cpp:8-31): SoA primitives of the
 * n_hydro cells; E is total energy per volume (allvars.h:203-207). */
int pg_state_create(pg_ctx* ctx, const pg_mesh* mesh, const double* rho, const double* vel,
                    const double* E, pg_state** out);
/* hydro::dt_CFL (finite_volume_solver.h:26, .cpp:151-186). */
int pg_dt_cfl(pg_ctx* ctx, const pg_mesh* mesh, const pg_state* st, double gamma, double cfl,
              double* dt);
/* hydro::hydro_step (finite_volume_solver.h:23, .cpp:43-123): one HLL
 * update on the static mesh, in place. */
int pg_hydro_step(pg_ctx* ctx, const pg_mesh* mesh, pg_state* st, double gamma, double dt,
                  int flux_mode);
int pg_state_copy_out(pg_ctx* ctx, const pg_state* st, double* rho, double* vel, double* E);
/* hydro::free_prim (finite_volume_solver.h:21). */
int pg_state_free(pg_state** st);

/* --------------------------------------------------- moving-mesh simulation */
/* The per-timestep loop of main.cpp:50-63 kept resident in HBM: generators,
 * conserved state (FP64 mass, momentum, energy) and the mesh never leave the
 * device between steps.  One pg_sim_step = primitives from the conserved
 * state, CFL dt, HLL update, generator drift (move_mesh) and one full
 * Voronoi rebuild (ghosts, grid, KNN, clip, restarts).  With move_mesh = 0
 * the mesh is static as in the reference driver (main.cpp:41). */
typedef struct {
    pg_mesh_params mesh;
    double gamma;   /* adiabatic index (allvars.cpp:14 hard-codes 5/3)           */
    double cfl;     /* CFL number (main.cpp:47 hard-codes 0.4)                   */
    int flux_mode;  /* PG_FLUX_*                                                 */
    int move_mesh;  /* 1: generators move with the fluid, ALE face fluxes        */
} pg_sim_params;

/* Stage times of the last pg_sim_step, milliseconds (HIP events). */
typedef struct {
    double hydro_ms;   /* primitives + dt + HLL update + drift                 */
    double ghosts_ms;  /* ghost generation (periodic)                          */
    double grid_ms;    /* grid counting sort                                   */
    double knn_ms;     /* KNN of every seed at the configured k                */
    double clip_ms;    /* first clip pass                                      */
    double restart_ms; /* certificate restarts                                 */
    double final_ms;   /* hydro face tables                                    */
    double total_ms;
    int64_t n_restarted;
    int k_used;
} pg_sim_times;

int pg_sim_create(pg_ctx* ctx, const pg_sim_params* params, const double* pts, const double* rho,
                  const double* vel, const double* E, int64_t n, pg_sim** out);
/* dt_max > 0 caps the CFL step (main.cpp:54-55 clamps to t_end). */
int pg_sim_step(pg_ctx* ctx, pg_sim* sim, double dt_max, double* dt_used);
/* nsteps static-mesh HLL updates on the current mesh (no rebuild). */
int pg_sim_hydro_steps(pg_ctx* ctx, pg_sim* sim, int nsteps, double dt_max, double* t_advanced);
int pg_sim_times_last(const pg_sim* sim, pg_sim_times* t);
int pg_sim_mesh(const pg_sim* sim, const pg_mesh** mesh);
/* Primitives (original order) and current generator positions. */
int pg_sim_get(pg_ctx* ctx, const pg_sim* sim, double* pts, double* rho, double* vel, double* E);
/* FP64 totals: [mass, px, py, pz, energy]. */
int pg_sim_totals(pg_ctx* ctx, pg_sim* sim, double* tot5);
int pg_sim_free(pg_sim** sim);

#ifdef __cplusplus
}
#endif
#endif /* PGPU_H */
