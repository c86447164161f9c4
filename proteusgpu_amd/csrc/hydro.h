// Host-side launchers of the hydro and copy-out kernels (internal interface).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pg {

// Conserved, extensive state per hydro cell (original order): mass,
// momentum, total energy (FP64), and the primitives derived from it.
struct SimView {
    int64_t n_hydro;
    double* mass;
    double* mom;  // [n][dim]
    double* en;
    double* rho;  // primitives
    double* vel;  // [n][dim]
    double* E;    // energy per volume
};

// Mesh data for the CFL length of a moving mesh.  count == nullptr: the
// reference's length (radius of the cell volume) alone.
struct CflMesh {
    const uint8_t* count = nullptr;  // faces per sorted cell
    const int32_t* nbr = nullptr;    // sorted neighbour id per face (face-major, stride)
    const double* sorted = nullptr;  // sorted generators (scaled box)
    int64_t stride = 0;
    double scale = 1.0;
};

// Primitives and generator velocity of one sorted cell (64 B), read by the
// face-flux pass of the owner-computes-once update.
struct CellRec {
    double rho, v[3], E, g[3];
};

void launch_q_from_prim(int dim, SimView sv, const int32_t* o2s, const double* vol_u, hipStream_t s);
void launch_prim_from_q(int dim, SimView sv, const int32_t* o2s, const double* vol_u, hipStream_t s);
// Owner-computes-once moving-mesh update.  hcode per sorted seed: hydro
// id, or -1 - (hydro id of the source) for a periodic ghost.  Seeds: n, or
// *n_dev when n_dev is set (batches: n is then the capacity).
// nbr_cm: cell-major neighbour lists (FACE_SLOTS per cell, from the clip
// kernel); solve: per cell, the mask of the faces it solves.
void launch_face_twins(const int64_t* n_dev, int64_t n, int64_t stride, const int32_t* hcode, const uint8_t* count,
                       const int32_t* nbr_cm, uint8_t* twin, uint64_t* solve, hipStream_t s);
struct OwnedUpdate {
    const int64_t* n_dev = nullptr;
    int64_t n = 0, stride = 0, nplane = 0;
    const int32_t* hcode = nullptr;
    const uint8_t* count = nullptr;
    const uint8_t* twin = nullptr;  // launch_face_twins of the same mesh
    const uint64_t* solve = nullptr;
    const int32_t* nbr = nullptr;
    const double *nrm = nullptr, *area_u = nullptr, *cen = nullptr, *sorted = nullptr;
    double scale = 1.0, gam = 1.4;
    int flux_mode = 1, move = 1;
    // time step: *dt_dev capped at dt_max > 0, or per box dt_box[hbox[i]]
    const double* dt_dev = nullptr;
    double dt_max = 0.0;
    const double* dt_box = nullptr;
    const int32_t* hbox = nullptr;
    SimView sv{};
    const double* wgen = nullptr;  // steered generator velocities (nullptr: the fluid's)
    CellRec* rec = nullptr;        // scratch: n records
    double* fbuf = nullptr;        // scratch: (dim + 2) doubles per face slot (slot-major)
};
void launch_hll_owned(int dim, OwnedUpdate u, hipStream_t s);
// HLL fluxes of n independent faces (riemann.cpp:7-45): states (rho,
// v[dim], E) per face, stored normals (3 per face), optional wall flags.
void launch_hll_faces(int dim, int64_t n, double gam, const double* L, const double* R, const double* nrm,
                      const uint8_t* wall, int flux_mode, double* F, hipStream_t s);
// Steered generator velocities (regularised mesh motion), hydro order.
void launch_gen_velocity(int dim, int64_t n_h, const double* x, const int32_t* o2s, const double* vol_u,
                         const double* ccen, int64_t stride, double scale, double gam, const double* rho,
                         const double* vel, const double* E, double* w, hipStream_t s);
void launch_drift(int dim, int64_t n, double* x, const double* vel, const double* dt_dev, double dt_max,
                  int periodic, hipStream_t s);
int64_t totals_blocks(int64_t n);
void launch_totals(int dim, SimView sv, double* part, hipStream_t s);

void launch_finalize(int64_t n, int64_t n_hydro, int64_t stride, const uint32_t* perm,
                     const int64_t* gsrc, const uint8_t* count, const int32_t* nbr,
                     const double* area, const double* vol, double ascale, double vscale,
                     int32_t* hid, double* area_u, double* vol_u, int32_t* o2s, int32_t* hcode, int32_t* bad,
                     hipStream_t s);

void launch_hll_static(int dim, int64_t n, int64_t n_hydro, int64_t stride, int64_t nplane,
                       const uint32_t* perm, const uint8_t* count, const int32_t* hid,
                       const double* nrm, const double* area_u, const double* vol_u, double gam,
                       double dt, int flux_mode, const double* rho, const double* vel,
                       const double* E, double* rho2, double* vel2, double* E2, hipStream_t s);

void launch_dt_cfl(int dim, int64_t n_hydro, const int32_t* o2s, const double* vol_u, double gam,
                   double cfl, const double* rho, const double* vel, const double* E,
                   unsigned long long* out, hipStream_t s, const double* wgen = nullptr, CflMesh cm = {});

void launch_counts_orig(int64_t n, const uint32_t* perm, const uint8_t* count, int32_t* co,
                        hipStream_t s);

void launch_pack_faces(int64_t n, int64_t stride, int64_t nplane, int dim, const uint32_t* perm,
                       const uint8_t* count, const int32_t* ptr_orig, const int32_t* nbr,
                       const double* nrm, const double* area, const double* cen, double ascale,
                       double cscale, int32_t* o_nbr, double* o_nrm, double* o_area, double* o_cen,
                       hipStream_t s);

void launch_cells_orig(int64_t n, int64_t stride, int dim, const uint32_t* perm,
                       const double* sorted, const int32_t* status, const int32_t* kcert,
                       const double* vol, const double* ccen, double scale, double vscale,
                       double* o_seeds, double* o_vol, int32_t* o_status, int32_t* o_kcert,
                       double* o_ccen, hipStream_t s);

// batched boxes (batch.hip)
void launch_bfinalize(const int64_t* n_dev, int64_t cap, int64_t stride, const uint32_t* perm, const int32_t* orig_h,
                      const uint8_t* count, const int32_t* nbr, const double* area, const double* vol, double ascale,
                      double vscale, int periodic, int32_t* hid, double* area_u, double* vol_u, int32_t* o2s,
                      int32_t* hsorted, int32_t* hcode, int32_t* bad, hipStream_t s);
void launch_bdt(int dim, int64_t n_h, int nbox, const int32_t* hbox, const int32_t* o2s, const double* vol_u,
                double gam, double cfl, const double* rho, const double* vel, const double* E, double* dt_box,
                hipStream_t s, const double* wgen = nullptr, CflMesh cm = {});
void launch_bdrift(int dim, int64_t n_h, const int32_t* hbox, double* x, const double* vel, const double* dt_box,
                   int periodic, hipStream_t s);
void launch_btotals(int dim, int nbox, const int64_t* hoff, SimView sv, double* out, hipStream_t s);

}  // namespace pg
