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
#define PG_ERR_GRID (-4)     /* N_grid < 16 (knn.cpp:32-36 exits here)             */
#define PG_ERR_CELLS (-5)    /* cells left uncertified after every restart         */
#define PG_ERR_NODEV (-6)    /* no usable GPU                                      */
#define PG_ERR_FAILED (-7)   /* the object failed in an earlier call and is unusable */

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
typedef struct pg_kmodel pg_kmodel;
typedef struct pg_batch pg_batch;

/* ------------------------------------------------------------------ context */
const char* pg_version(void);
const char* pg_last_error(void);
int pg_device_count(int* count);
/* Replaces the reference's gpu* shim (allvars.h:92-129): one device + stream. */
int pg_ctx_create(int device, pg_ctx** out);
int pg_ctx_destroy(pg_ctx* ctx);
int pg_ctx_sync(pg_ctx* ctx);
/* No reference counterpart: streams `bytes` device-to-device `reps` times and
 * returns the mean kernel time (HIP events); measured HBM bandwidth =
 * 2*bytes/ms.  Also the calibration kernel for HBM traffic counters. */
int pg_stream_copy_probe(pg_ctx* ctx, int64_t bytes, int reps, double* ms_per_copy);

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

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): pg_knn_init + pg_knn_solve on DEVICE arrays, with every
 * scratch buffer (grid counts and starts, cell ids, scan partials) in a
 * static device array of PG_NOALLOC_SCRATCH_BYTES that each GPU holds until
 * the process ends.  pts (n*dim, device) in; sorted (n*dim), perm (n) and
 * rows (n*k) device out, equal to pg_knn_get_points / get_permutation /
 * get_knearest of the allocating path; N_grid (host, may be NULL).  About
 * 9.2 bytes of scratch per point: n up to roughly 3.5M points; larger n
 * gives PG_ERR_ARG before anything is launched.  Calls no hipMalloc, hipFree
 * or memset; work is enqueued on the context's stream and the call returns
 * without synchronising.  Every call on a GPU shares that GPU's array, so
 * calls that overlap in time on one GPU (several contexts or threads) race. */
#define PG_NOALLOC_SCRATCH_BYTES (32 * 1024 * 1024)
int pg_knn_solve_noalloc(pg_ctx* ctx, int dim, const double* pts, int64_t n, int k, double* sorted,
                         uint32_t* perm, uint32_t* rows, int* N_grid);

/* Shape of the fast clip kernel for every following build of this process,
 * read once at the start of each build
 * (no reference counterpart; results never depend on it): lanes per seed
 * (8 or 16) and vertex slots per seed (48 or 64; seeds that outgrow them
 * are recomputed by a 128-slot kernel).  Default 16 x 64, or PGPU_CLIP_CFG. */
int pg_set_clip_config(int lanes, int slots);
int pg_get_clip_config(int* lanes, int* slots);

/* Certificate restarts (no reference counterpart; results never depend on
 * it) for every following build and batch creation of this process: 1 =
 * an uncertified seed's cell is saved (3328 bytes per seed of device memory
 * while restarts are enabled) and the restart at a longer row continues it;
 * 0 = the restart clips again from the first neighbour.  Default 1, or
 * PGPU_CLIP_RESUME. */
int pg_set_clip_resume(int on);
int pg_get_clip_resume(int* on);

/* ------------------------------------------------------------------ Voronoi */
typedef struct {
    int dim;       /* 2 or 3 (Config.sh dim_2D / dim_3D)                          */
    int boundary;  /* PG_BOUNDARY_*                                              */
    double buff;   /* ghost layer width (allvars.cpp:13, default 0.04)           */
    int k;         /* neighbours per seed (Config.sh _K_)                        */
    int max_p;     /* plane capacity per cell (Config.sh _MAX_P_, <= 64)         */
    int max_t;     /* vertex capacity per cell (Config.sh _MAX_T_, <= 128)       */
    int restart;   /* nonzero: recompute uncertified seeds at larger k; seeds    *
                    * still uncertified at k_max are finished by the vertex-ball *
                    * completion (their first min(k_max, 64) neighbours, then   *
                    * the points inside their vertices' empty-sphere balls      *
                    * until every ball is empty; kcert = k_max + 1; host-driven *
                    * builds only)                                               */
    int k_max;     /* largest k a restart may use (default and limit 1024;      *
                    * batches: at most 256)                                      */
    /* Degeneracy fallback (no reference counterpart; the reference leaves
     * such cells failed, voronoi.cpp:104-110).  0 (default): reference
     * behaviour.  > 0: when a host-driven build (pg_mesh_compute*, pg_sim_*)
     * leaves cells with PG_NEEDS_EXACT_PREDICATES or PG_INCONSISTENT_BOUNDARY
     * (degenerate inputs such as lattices, whose cospherical neighbours give
     * zero predicates), every generator is displaced by a deterministic
     * offset of at most perturb x n^(-1/dim) per axis and the mesh is built
     * again, with ten times the amplitude and a new offset pattern per retry
     * (three builds at most).  The mesh is then the exact Voronoi mesh of
     * the displaced generators: its seeds are the displaced points, and the
     * zero-area faces of the degenerate mesh become faces of area
     * O(amplitude x spacing).  pg_mesh_info_t.displacement reports the
     * amplitude used.  Batches (pg_batch_*) do not apply it.               */
    double perturb;
    /* Exact-predicate fallback (no reference counterpart; the reference sets
     * needs_exact_predicates, cell.cpp:216-218, and leaves the cell failed,
     * voronoi.cpp:104-110).  0 (default): reference behaviour.  1: cells a
     * build leaves with PG_NEEDS_EXACT_PREDICATES are clipped again with the
     * conflict test evaluated exactly from the generators (fixed-width
     * integers) and exact ties broken by Simulation of Simplicity on power
     * weights ranked by input index, so every cell of the perturbed diagram
     * is certified and neighbouring cells agree face for face; the
     * generators are not moved and the zero-area faces of the degenerate
     * diagram appear as faces of zero area.  Cells whose filtered predicate
     * never needs it are unchanged bit for bit.  Such runs certify a cell
     * only when the certifying neighbour lies beyond twice the largest
     * vertex distance by a relative margin of 1e-9 (a tied neighbour can cut
     * a zero-area face).  Coordinates must lie in [2^-120, 2) or be zero
     * for the exact test (else the cell keeps PG_NEEDS_EXACT_PREDICATES).
     * Applied before perturb.  Batches do not apply it.                     */
    int exact;
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
    int64_t row_entries;/* neighbour-row entries computed (sum of rows x k over   *
                         * the first pass and every restart)                     */
    int64_t n_class[6]; /* learned build: seeds predicted per k bucket            */
    int64_t n_degenerate; /* cells the undisplaced build left with                *
                           * PG_NEEDS_EXACT_PREDICATES or PG_INCONSISTENT_BOUNDARY *
                           * (counted when the degeneracy fallback is on)         */
    double displacement;  /* amplitude of the displacement the mesh was built     *
                           * with (0: the generators as given)                    */
    int64_t n_exact;      /* cells clipped again by the exact-predicate fallback */
    int64_t n_ball;       /* cells the vertex-ball completion finished (kcert =   *
                           * k_max + 1 when it certified them)                    */
} pg_mesh_info_t;

/* voronoi::compute_mesh (voronoi.h:68, voronoi.cpp:17-65) for boundary NONE,
 * voronoi::compute_periodic_mesh (periodic_mesh.h:22) for PERIODIC; REFLECTIVE
 * has no reference counterpart.  Points are host arrays. */
int pg_mesh_compute(pg_ctx* ctx, const pg_mesh_params* params, const double* pts, int64_t n,
                    pg_mesh** out);
int pg_mesh_info(const pg_mesh* mesh, pg_mesh_info_t* info);

/* voronoi::compute_cells (voronoi.h:69, voronoi.cpp:68-111): the cells of a
 * solved KNN problem (pg_knn_solve), clipped with its rows of length k and
 * no certificate restarts (boundary NONE, the box [0,1]^dim).  The mesh is
 * indexed by the problem's sorted ids, as compute_cells leaves it before
 * unpermute_vmesh: pg_mesh_copy_out returns cells in sorted order with
 * neighbour ids in sorted-index space.  status (may be NULL): the per-seed
 * status (PG_* codes, n entries, sorted order), as compute_cells' stat. */
int pg_voronoi_compute(pg_ctx* ctx, const pg_knn* knn, int max_p, int max_t, pg_mesh** out, int32_t* status);
/* voronoi::unpermute_vmesh (voronoi.h:77, voronoi.cpp:212-333) for a mesh of
 * pg_voronoi_compute: later copy-outs return cells in input order with
 * original neighbour ids (the reorder runs on the device at copy-out).
 * PG_ERR_ARG for a mesh already in input order. */
int pg_mesh_unpermute(pg_ctx* ctx, pg_mesh* mesh);

/* The reference VMesh (voronoi.h:22-41) after unpermute_vmesh
 * (voronoi.cpp:212-333) and, for periodic meshes, the unscale of
 * periodic_mesh.cpp:146-166.  Caller allocates every array (sizes from
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
/* The KNN problem of a fixed-k mesh build, as knn::get_points /
 * get_knearest / get_permutation (knn.cpp:439-455) return it for the KNN
 * HDF5 dump (knn.cpp:457-473, output.cpp:305-411): the build's points
 * (ghosts included, scaled into the unit box for periodic meshes) in sorted
 * order (n_seeds*dim), the first pass's neighbour rows of length k in
 * sorted-index space (n_seeds*k) and the permutation (n_seeds).  Any pointer
 * may be NULL.  PG_ERR_ARG for a mesh built with learned k (its rows have
 * per-seed lengths). */
int pg_mesh_get_knn(pg_ctx* ctx, const pg_mesh* mesh, double* points, uint32_t* nearest, uint32_t* permutation);
/* voronoi::free_vmesh (voronoi.h:65). */
int pg_mesh_free(pg_mesh** mesh);

/* -------------------------------------------------------------------- hydro */
/* hydro::init (finite_volume_solver.h:20, .cpp:8-31): SoA primitives of the
 * n_hydro cells; E is total energy per volume (allvars.h:203-207). */
int pg_state_create(pg_ctx* ctx, const pg_mesh* mesh, const double* rho, const double* vel,
                    const double* E, pg_state** out);
/* hydro::dt_CFL (finite_volume_solver.h:27, .cpp:151-186). */
int pg_dt_cfl(pg_ctx* ctx, const pg_mesh* mesh, const pg_state* st, double gamma, double cfl,
              double* dt);
/* hydro::hydro_step (finite_volume_solver.h:23, .cpp:43-123): one HLL
 * update on the static mesh, in place. */
int pg_hydro_step(pg_ctx* ctx, const pg_mesh* mesh, pg_state* st, double gamma, double dt,
                  int flux_mode);
int pg_state_copy_out(pg_ctx* ctx, const pg_state* st, double* rho, double* vel, double* E);
/* hydro::riemann_hll (riemann.h:24, riemann.cpp:7-45) for n independent
 * faces: left and right states (rho, v[dim], E; E total energy per volume)
 * and the stored face normal (3 per face, unit (s_i - s_j) as VMesh keeps
 * it) in host arrays; wall (may be NULL): nonzero marks a reflective wall,
 * whose right state is the mirrored left one.  flux (n*(dim+2)): mass,
 * momentum[dim] and energy flux per unit area in the lab frame.
 * PG_FLUX_REFERENCE takes the stored normal as the face's +x axis, as the
 * reference does; PG_FLUX_PHYSICAL its negation (the outward normal). */
int pg_hll_flux(pg_ctx* ctx, int dim, double gamma, int64_t n, const double* left, const double* right,
                const double* normal, const uint8_t* wall, int flux_mode, double* flux);
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
    int move_mesh;  /* PG_MOVE_*                                                 */
} pg_sim_params;

/* Generator motion of a simulation (pg_sim_params.move_mesh). */
#define PG_MOVE_STATIC 0      /* static mesh (main.cpp:41)                     */
#define PG_MOVE_FLUID 1       /* generators move with the fluid velocity       */
#define PG_MOVE_REGULARIZED 2 /* fluid velocity plus a drift towards the cell
                               * centroid (Springel 2010, MNRAS 401, 791, eq. 63:
                               * eta = 0.25, chi = 1), keeps cells round        */

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
    double predict_ms; /* learned k: features + classifier + bucketing         */
    int64_t row_entries;
} pg_sim_times;

int pg_sim_create(pg_ctx* ctx, const pg_sim_params* params, const double* pts, const double* rho,
                  const double* vel, const double* E, int64_t n, pg_sim** out);
/* dt_max > 0 caps the CFL step (main.cpp:54-55 clamps to t_end).  A step
 * that fails (a rebuild that leaves cells uncertified, a periodic ghost layer
 * thinner than the cells, a HIP error) returns its error after the state was
 * already advanced and the generators moved; the simulation is then
 * unusable: pg_sim_step, pg_sim_hydro_steps, pg_sim_get, pg_sim_totals and
 * pg_sim_mesh return PG_ERR_FAILED without launching anything.  Free it. */
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
/* Learned k for every following rebuild of the simulation (NULL: the fixed
 * k of the parameters).  The model must outlive its use. */
int pg_sim_set_kmodel(pg_sim* sim, pg_kmodel* model);

/* ---------------------------------------------------------------- learned k */
/* No counterpart in the reference (its k is the compile-time _K_, knn.h:20).
 * A classifier 16 features -> 64 -> 64 -> 6 (ReLU; bfloat16 operands, FP32
 * accumulation on MFMA) picks each seed's k among the buckets
 * 16, 24, 32, 48, 64, 96; the security-radius certificate judges every
 * cell and restarts the ones it rejects, so the mesh equals the fixed-k
 * mesh bit for bit.  Parameters: one float32 vector of PG_KMODEL_NPARAM,
 * W1[16][64] b1[64] W2[64][64] b2[64] W3[64][6] b3[6] (z = x W + b). */
#define PG_KMODEL_NFEAT 16
#define PG_KMODEL_NCLASS 6
#define PG_KMODEL_NPARAM 5638

/* Glorot-uniform weights from a SplitMix64 stream of `seed`, zero biases. */
int pg_kmodel_create(pg_ctx* ctx, uint64_t seed, pg_kmodel** out);
int pg_kmodel_set_params(pg_ctx* ctx, pg_kmodel* model, const float* params);
int pg_kmodel_get_params(pg_ctx* ctx, const pg_kmodel* model, float* params);
/* Cost model of the bucket choice: a pass at row length k costs k (KNN row)
 * + r m (m clip steps: the certified k, or k when the certificate fails);
 * a rejected seed repeats at 2k (then 4k, ...).  The bucket whose chain has
 * the least expected cost under the predicted distribution wins.
 * r = clip cost per step / KNN cost per row entry.  Restarts that resume
 * the saved cell waste fewer clip steps than the model charges, so the
 * default r = 1.0 is set by measurement rather than the ratio itself (1M
 * uniform 3D seeds on MI355X, learned-k step: r = 2.8 26.1 ms, 1.5 25.4 ms,
 * 0.8 25.5 ms, 0.3 25.3 ms). */
int pg_kmodel_set_cost(pg_kmodel* model, double r);
/* Host features (n x 16) -> bucket index per row (and the 6 logits per row
 * when logits != NULL). */
int pg_kmodel_predict(pg_ctx* ctx, const pg_kmodel* model, const float* feat, int64_t n, int32_t* cls,
                      float* logits);
/* Mean cross-entropy of the rows and its gradient (PG_KMODEL_NPARAM floats). */
int pg_kmodel_grad(pg_ctx* ctx, const pg_kmodel* model, const float* feat, const int32_t* cls, int64_t n,
                   float* grad, double* loss);
/* One Adam step (beta 0.9 / 0.999, eps 1e-8) on the rows as one batch. */
int pg_kmodel_train_step(pg_ctx* ctx, pg_kmodel* model, const float* feat, const int32_t* cls, int64_t n,
                         double lr, double* loss);
/* `steps` Adam steps on a built mesh: features and labels (the bucket of
 * each seed's minimal certified k) stay on the device; batches of `batch`
 * seeds are drawn by a hash of (seed, step, i).  loss: mean loss of the
 * last step. */
int pg_kmodel_train_mesh(pg_ctx* ctx, pg_kmodel* model, const pg_mesh* mesh, int steps, int64_t batch,
                         double lr, uint64_t seed, double* loss);
int pg_kmodel_free(pg_kmodel** model);

/* Features of every seed of a built mesh, original order (n_seeds x 16). */
int pg_mesh_features(pg_ctx* ctx, const pg_mesh* mesh, float* out);
/* pg_mesh_compute with each seed's first k chosen by `model` (NULL: the
 * fixed k of the parameters). */
int pg_mesh_compute_k(pg_ctx* ctx, const pg_mesh_params* params, const pg_kmodel* model, const double* pts,
                      int64_t n, pg_mesh** out);

/* ------------------------------------------------------------------ batches */
/* No counterpart in the reference (one box per run).  Many independent boxes
 * of uneven size advanced together: each pg_batch_step is one hipGraph
 * launch (captured at the first step) holding the moving-mesh step of
 * every box -- primitives, per-box CFL step, HLL update, drift, ghosts,
 * per-box grids, KNN, clip and the restart chain -- with all per-step
 * counts kept on the device.  Every box is bit-identical to the same box
 * run alone through pg_sim_step.  Boxes are concatenated box-major in the
 * point and state arrays; coordinates of every box lie in its own unit box.
 * Every box needs more than max(k, k_max) + 1 points. */
typedef struct {
    int clip_lanes;      /* lanes per seed of the fast clip kernel: 8 or 16         */
    int clip_slots;      /* vertex slots per seed of the fast clip kernel: 48 or 64 */
    int knn_wave_min_k;  /* rows at least this long use the wave-per-query KNN      */
    int blocks_per_cu;   /* resident blocks per CU of the persistent KNN and clip
                          * kernels (KNN, 2D clip, restart clips), at most their
                          * occupancy (default 64: all)                          */
    int use_graph;       /* 1: capture the step in a hipGraph, 0: plain launches     */
} pg_batch_tuning;

typedef struct {
    int nbox;
    int64_t n_hydro;        /* cells over all boxes                              */
    int64_t n_seeds;        /* cells + ghosts of the current mesh                */
    int64_t seed_capacity;  /* seeds reserved (ghost layers included)            */
    double last_step_ms;
} pg_batch_info_t;

int pg_batch_create(pg_ctx* ctx, const pg_sim_params* params, int nbox, const int64_t* n_per_box,
                    const double* pts, const double* rho, const double* vel, const double* E, pg_batch** out);
/* One step of every box; ms (optional) = GPU time of the step.  A failed
 * step (or autotune) leaves the batch unusable as a failed pg_sim_step does:
 * later steps, autotunes, gets, dts and totals return PG_ERR_FAILED. */
int pg_batch_step(pg_ctx* ctx, pg_batch* batch, double* ms);
int pg_batch_set_tuning(pg_batch* batch, const pg_batch_tuning* t);
/* Learned k for the following steps (NULL: the fixed k of the parameters);
 * the model must outlive its use.  Results do not change. */
int pg_batch_set_kmodel(pg_batch* batch, const pg_kmodel* model);
int pg_batch_get_tuning(const pg_batch* batch, pg_batch_tuning* t);
/* Time every tuning candidate (clip shape x KNN kernel choice x grid size)
 * on `reps` steps from the current state, restore the state and keep the
 * fastest.  Results never depend on the tuning. */
int pg_batch_autotune(pg_ctx* ctx, pg_batch* batch, int reps, pg_batch_tuning* best, double* best_ms);
/* Primitives and generator positions, box-major. */
int pg_batch_get(pg_ctx* ctx, pg_batch* batch, double* pts, double* rho, double* vel, double* E);
/* CFL step each box took in the last step (nbox doubles). */
int pg_batch_dt(pg_ctx* ctx, const pg_batch* batch, double* dt);
/* FP64 totals per box: [mass, px, py, pz, energy] x nbox. */
int pg_batch_totals(pg_ctx* ctx, pg_batch* batch, double* tot);
int pg_batch_info(pg_ctx* ctx, const pg_batch* batch, pg_batch_info_t* info);
int pg_batch_free(pg_batch** batch);

#ifdef __cplusplus
}
#endif
#endif /* PGPU_H */
