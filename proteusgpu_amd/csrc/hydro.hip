// Hydro on the Voronoi mesh: mesh finalisation, the HLL update
// (finite_volume_solver.cpp:43-148, riemann.cpp:7-124), the CFL reduce
// (finite_volume_solver.cpp:151-186) and the reference-layout copy-out of
// the mesh (voronoi.cpp:212-333 unpermute, periodic_mesh.cpp:146-166 unscale).
//
// Face data come from the clip kernel in face-slot-major padded storage
// (entry (f, sorted cell s) at f*stride + s), so threads of a wave that handle
// consecutive sorted cells read consecutive addresses for every face slot.
#include <algorithm>

#include "common.h"
#include "hydro.h"
#include "hydro_core.h"
#include "kernels.h"

namespace pg {
namespace {

constexpr int TB = 256;
inline unsigned nblk(int64_t n) { return (unsigned)((n + TB - 1) / TB); }
inline unsigned nblk_dt(int64_t n) { return (unsigned)((n + (int64_t)TB * DT_CELLS - 1) / ((int64_t)TB * DT_CELLS)); }

// Per sorted cell: hydro id of every face's neighbour (-1 = box plane,
// which is a reflective wall), unscaled area and cell volume.
__global__ void __launch_bounds__(TB) k_finalize(int64_t n, int64_t n_hydro, int64_t stride,
                                                 const uint32_t* perm, const int64_t* gsrc,
                                                 const uint8_t* count, const int32_t* nbr,
                                                 const double* area, const double* vol, double ascale,
                                                 double vscale, int32_t* hid, double* area_u,
                                                 double* vol_u, int32_t* o2s, int32_t* hcode, int32_t* bad) {
    int64_t s = (int64_t)blockIdx.x * TB + threadIdx.x;
    if (s >= n) return;
    const uint32_t o = perm[s];
    // hydro id, or -1 - source for a periodic ghost (-1 without ghosts)
    hcode[s] = o < n_hydro ? (int32_t)o : (gsrc ? -1 - (int32_t)gsrc[o - n_hydro] : -1);
    if (o >= n_hydro) return;
    o2s[o] = (int32_t)s;
    vol_u[s] = vscale * vol[s];
    const int c = count[s];
    if (c == 0) atomicAdd(bad, 1);
    for (int f = 0; f < c; f++) {
        const int64_t at = (int64_t)f * stride + s;
        const int32_t j = nbr[at];
        int32_t h = -1;
        if (j >= 0) {
            const uint32_t oj = perm[j];
            h = oj < n_hydro ? (int32_t)oj : (int32_t)gsrc[oj - n_hydro];
        } else if (gsrc) {
            // periodic mesh: a hydro cell reaching the box plane means the
            // ghost layer is thinner than the cell (finite_volume_solver.cpp:130-136
            // would read out of bounds here)
            atomicAdd(bad + 1, 1);
        }
        hid[at] = h;
        area_u[at] = ascale * area[at];
    }
}

// One static-mesh HLL update per hydro cell (finite_volume_solver.cpp:53-106):
// every cell sums the fluxes through its own faces, exactly as the
// reference, and writes the new primitives to the second buffer.
template <int DIM>
__global__ void __launch_bounds__(TB)
    k_hll_static(int64_t n, int64_t n_hydro, int64_t stride, int64_t nplane,
                 const uint32_t* __restrict__ perm, const uint8_t* __restrict__ count,
                 const int32_t* __restrict__ hid, const double* __restrict__ nrm,
                 const double* __restrict__ area_u, const double* __restrict__ vol_u, double gam,
                 double dt, int flux_mode, const double* __restrict__ rho,
                 const double* __restrict__ vel, const double* __restrict__ E,
                 double* __restrict__ rho2, double* __restrict__ vel2, double* __restrict__ E2) {
    int64_t s = (int64_t)blockIdx.x * TB + threadIdx.x;
    if (s >= n) return;
    const uint32_t i = perm[s];
    if (i >= n_hydro) return;
    Prim si;
    si.rho = rho[i];
    si.E = E[i];
    si.v[2] = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; d++) si.v[d] = vel[(size_t)DIM * i + d];
    Prim tot = {0.0, {0.0, 0.0, 0.0}, 0.0};
    const int c = count[s];
    const double sgn = flux_mode == 1 ? -1.0 : 1.0;
    for (int f = 0; f < c; f++) {
        const int64_t at = (int64_t)f * stride + s;
        const int32_t j = hid[at];
        const bool wall = j < 0;
        Prim sj = si;
        if (!wall) {
            sj.rho = rho[j];
            sj.E = E[j];
#pragma unroll
            for (int d = 0; d < DIM; d++) sj.v[d] = vel[(size_t)DIM * j + d];
        }
        // flux_mode 1 negates the stored (inward) normal exactly
        const Geom g = geom_of(sgn * nrm[at], sgn * nrm[nplane + at], sgn * nrm[2 * nplane + at]);
        const Prim F = hll<DIM>(gam, si, sj, g, wall);
        const double a = area_u[at];
        tot.rho += F.rho * a;
#pragma unroll
        for (int d = 0; d < DIM; d++) tot.v[d] += F.v[d] * a;
        tot.E += F.E * a;
    }
    const double frac = dt / vol_u[s];
    const double nr = si.rho - frac * tot.rho;
    rho2[i] = nr;
    const double inv = 1.0 / nr;
#pragma unroll
    for (int d = 0; d < DIM; d++) vel2[(size_t)DIM * i + d] = (si.rho * si.v[d] - frac * tot.v[d]) * inv;
    E2[i] = si.E - frac * tot.E;
}

// riemann.cpp:7-45 for n independent faces: left / right states (rho,
// v[DIM], E per face), the stored normal of each face (3 per face), optional
// wall flags (the right state is then the mirrored left one), fluxes out
// (mass, momentum[DIM], energy per face).  sgn = -1 takes the outward
// orientation (flux mode 1), exactly as the cell kernels do.
template <int DIM>
__global__ void __launch_bounds__(TB)
    k_hll_faces(int64_t n, double gam, const double* __restrict__ L, const double* __restrict__ R,
                const double* __restrict__ nrm, const uint8_t* __restrict__ wall, double sgn,
                double* __restrict__ F) {
    const int64_t i = (int64_t)blockIdx.x * TB + threadIdx.x;
    if (i >= n) return;
    constexpr int W = DIM + 2;
    Prim a, b;
    a.v[2] = b.v[2] = 0.0;
    a.rho = L[W * i];
    b.rho = R[W * i];
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        a.v[d] = L[W * i + 1 + d];
        b.v[d] = R[W * i + 1 + d];
    }
    a.E = L[W * i + DIM + 1];
    b.E = R[W * i + DIM + 1];
    const bool w = wall != nullptr && wall[i] != 0;
    const Geom g = geom_of(sgn * nrm[3 * i], sgn * nrm[3 * i + 1], sgn * nrm[3 * i + 2]);
    const Prim f = hll<DIM>(gam, a, w ? a : b, g, w);
    F[W * i] = f.rho;
#pragma unroll
    for (int d = 0; d < DIM; d++) F[W * i + 1 + d] = f.v[d];
    F[W * i + DIM + 1] = f.E;
}

// finite_volume_solver.cpp:151-186: per-cell CFL time step, min over cells.
template <int DIM>
__global__ void __launch_bounds__(TB)
    k_dt_cfl(int64_t n_hydro, const int32_t* __restrict__ o2s, const double* __restrict__ vol_u,
             double gam, double cfl, const double* __restrict__ rho, const double* __restrict__ vel,
             const double* __restrict__ E, unsigned long long* out, const double* __restrict__ wgen, CflMesh cm) {
    const int64_t base = ((int64_t)blockIdx.x * TB + (threadIdx.x & ~63u)) * DT_CELLS + (threadIdx.x & 63u);
    double m = 1e100;
#pragma unroll
    for (int it = 0; it < DT_CELLS; it++) {
        const int64_t i = base + 64 * it;
        if (i < n_hydro) m = smin(cell_cfl_dt<DIM>(i, o2s, vol_u, gam, cfl, rho, vel, E, wgen, cm), m);
    }
    // wave min, block min, one atomic per block
    for (int d = 32; d > 0; d >>= 1) m = smin(__shfl_xor(m, d, 64), m);
    __shared__ double wm[TB / 64];
    if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int w = 1; w < TB / 64; w++) m = smin(wm[w], m);
        report_dt(out, m);
    }
}

// ---------------------------------------------------------------- simulation
// Conserved, extensive state per hydro cell (original order): mass, momentum,
// total energy, all FP64.  Primitives are derived from it and the current
// cell volume before every update.


template <int DIM>
__global__ void __launch_bounds__(TB) k_q_from_prim(SimView sv, const int32_t* o2s, const double* vol_u) {
    int64_t i = (int64_t)blockIdx.x * TB + threadIdx.x;
    if (i >= sv.n_hydro) return;
    const double V = vol_u[o2s[i]];
    const double r = sv.rho[i];
    sv.mass[i] = r * V;
#pragma unroll
    for (int d = 0; d < DIM; d++) sv.mom[(size_t)DIM * i + d] = r * sv.vel[(size_t)DIM * i + d] * V;
    sv.en[i] = sv.E[i] * V;
}

template <int DIM>
__global__ void __launch_bounds__(TB) k_prim_from_q(SimView sv, const int32_t* o2s, const double* vol_u) {
    int64_t i = (int64_t)blockIdx.x * TB + threadIdx.x;
    if (i >= sv.n_hydro) return;
    const double V = vol_u[o2s[i]];
    const double m = sv.mass[i];
    sv.rho[i] = m / V;
#pragma unroll
    for (int d = 0; d < DIM; d++) sv.vel[(size_t)DIM * i + d] = sv.mom[(size_t)DIM * i + d] / m;
    sv.E[i] = sv.en[i] / V;
}

// ---------------------------------------------------------- owner-computes-once
// Moving-mesh (ALE) HLL update of the conserved state.  The face moves with
// w_f = (w_i + w_j)/2 plus the standard correction for the offset of the face
// centroid from the generators' midpoint; the Riemann problem is solved in
// the face frame and transformed back (mass flux unchanged, momentum
// + F_m w_f, energy + F_p.w_f + |w_f|^2 F_m / 2).  With move == 0 this is the
// static flux of k_hll_static applied to extensive variables.  Walls do not
// move (w_f = 0).  No reference counterpart: the reference never moves its mesh.
//
// The update solves each face between two hydro cells once (the reference,
// finite_volume_solver.cpp:69-87, solves every face from both sides).
// A face of sorted cells s < t is owned by s; t finds the same face in s's
// list (twin) and subtracts s's flux instead of solving it again.  Faces to
// walls and periodic ghosts, and a face that the other cell does not list,
// are solved by the cell itself.  Three passes, none with atomics:
//   k_cell_records  primitives + generator velocity per sorted seed (ghosts
//                   take their source cell's), so that neighbour reads are
//                   by sorted id and land near the cell's own
//   k_face_flux     one thread per (face row r, cell): faces r, r+FR, ...
//                   that the cell solves; writes flux x area per face slot
//   k_face_sum      per cell: its own slots plus the negated twins, in face
//                   order (a segmented sum over the cell's face list), and
//                   the update of the conserved state
constexpr int FR = 8;            // threads per cell in k_face_flux (face slot rows)
constexpr uint8_t SOLVE = 0xFF;  // twin code: solved by this cell

// Seeds of one update: n (or *n_dev, a device-side count, for batches);
// hcode[s] = hydro id of sorted seed s, or -1 - (hydro id of its source) for
// a periodic ghost.  Kernels walk the seeds grid-stride.
PG_DEV int64_t seed_count(const int64_t* n_dev, int64_t n) { return n_dev ? *n_dev : n; }

// One thread per cell: twin[at] = face index of face (s, f) in the owner's
// list, or SOLVE; solve[s] = mask of the faces s solves.  The owner's list is
// searched in its cell-major copy (nbr_cm, FACE_SLOTS per cell, written by
// the clip kernel), 16 bytes per load.
__global__ void __launch_bounds__(TB) k_face_twins(const int64_t* n_dev, int64_t n_host, int64_t stride,
                                                   const int32_t* __restrict__ hcode,
                                                   const uint8_t* __restrict__ count,
                                                   const int32_t* __restrict__ nbr_cm, uint8_t* __restrict__ twin,
                                                   uint64_t* __restrict__ solve) {
    const int64_t n = seed_count(n_dev, n_host);
    for (int64_t s = (int64_t)blockIdx.x * TB + threadIdx.x; s < n; s += (int64_t)gridDim.x * TB) {
        if (hcode[s] < 0) continue;
        const int c = count[s];
        const int4* ls = reinterpret_cast<const int4*>(nbr_cm + s * FACE_SLOTS);
        uint64_t mask = 0;
        for (int f0 = 0; f0 < c; f0 += 4) {
            const int4 q = ls[f0 >> 2];
            const int32_t tq[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int f = f0 + e;
                if (f >= c) break;
                const int32_t t = tq[e];
                uint8_t code = SOLVE;
                if (t >= 0 && t < s && hcode[t] >= 0) {
                    const int ct = count[t];
                    const int4* lt = reinterpret_cast<const int4*>(nbr_cm + (int64_t)t * FACE_SLOTS);
                    for (int g0 = 0; g0 < ct && code == SOLVE; g0 += 4) {
                        const int4 v = lt[g0 >> 2];
                        const int32_t vv[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                        for (int h = 0; h < 4; h++)
                            if (g0 + h < ct && vv[h] == (int32_t)s && code == SOLVE) code = (uint8_t)(g0 + h);
                    }
                }
                twin[(int64_t)f * stride + s] = code;
                if (code == SOLVE) mask |= 1ull << f;
            }
        }
        solve[s] = mask;
    }
}

template <int DIM>
__global__ void __launch_bounds__(TB) k_cell_records(const int64_t* n_dev, int64_t n_host,
                                                     const int32_t* __restrict__ hcode, SimView sv,
                                                     const double* __restrict__ wgen, CellRec* __restrict__ rec) {
    const int64_t n = seed_count(n_dev, n_host);
    const double* gv = wgen ? wgen : sv.vel;
    for (int64_t s = (int64_t)blockIdx.x * TB + threadIdx.x; s < n; s += (int64_t)gridDim.x * TB) {
        const int32_t c = hcode[s];
        const int64_t h = c >= 0 ? c : -1 - (int64_t)c;
        CellRec r;
        r.rho = sv.rho[h];
        r.E = sv.E[h];
        r.v[2] = r.g[2] = 0.0;
#pragma unroll
        for (int d = 0; d < DIM; d++) {
            r.v[d] = sv.vel[(size_t)DIM * h + d];
            r.g[d] = gv[(size_t)DIM * h + d];
        }
        rec[s] = r;
    }
}

template <int DIM>
__global__ void __launch_bounds__(TB)
    k_face_flux(const int64_t* n_dev, int64_t n_host, int64_t stride, int64_t nplane, const int32_t* __restrict__ hcode,
                const uint64_t* __restrict__ solve, const int32_t* __restrict__ nbr,
                const double* __restrict__ nrm, const double* __restrict__ area_u, const double* __restrict__ cen,
                const double* __restrict__ sorted, double scale, double gam, int flux_mode, int move,
                const CellRec* __restrict__ rec, double* __restrict__ fbuf) {
    const int64_t n = seed_count(n_dev, n_host);
    const double sgn = flux_mode == 1 ? -1.0 : 1.0;
    for (int64_t s = (int64_t)blockIdx.x * TB + threadIdx.x; s < n; s += (int64_t)gridDim.x * TB) {
        if (hcode[s] < 0) continue;
        // face slots r, r + FR, ... (r = blockIdx.y) that this cell solves:
        // the lanes of a wave read the same slot row of consecutive cells
        const uint64_t m = solve[s] >> blockIdx.y;
        if (!m) continue;
        const CellRec own = rec[s];
        Prim si;
        si.rho = own.rho;
        si.E = own.E;
        double ri[3] = {0, 0, 0}, gi[3] = {0, 0, 0};
#pragma unroll
        for (int d = 0; d < 3; d++) {
            si.v[d] = own.v[d];
            gi[d] = own.g[d];
        }
#pragma unroll
        for (int d = 0; d < DIM; d++) ri[d] = (sorted[(size_t)DIM * s + d] - 0.5) * scale + 0.5;
        for (int f = blockIdx.y; f < 64 && (m >> (f - blockIdx.y)); f += FR) {
            if (!((m >> (f - blockIdx.y)) & 1ull)) continue;
            const int64_t at = (int64_t)f * stride + s;
            const int32_t t = nbr[at];
            const bool wall = t < 0;
            Prim sj = si;
            double rj[3] = {0, 0, 0}, gj[3] = {0, 0, 0}, fc[3] = {0, 0, 0};
            if (!wall) {
                const CellRec o = rec[t];
                sj.rho = o.rho;
                sj.E = o.E;
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    sj.v[d] = o.v[d];
                    gj[d] = o.g[d];
                }
                if (move) {
#pragma unroll
                    for (int d = 0; d < DIM; d++) {
                        rj[d] = (sorted[(size_t)DIM * t + d] - 0.5) * scale + 0.5;
                        fc[d] = (cen[(int64_t)d * nplane + at] - 0.5) * scale + 0.5;
                    }
                }
            }
            double Fm, Fp[3], Fe;
            ale_face<DIM>(gam, sgn, move, si, ri, gi, sj, rj, gj, wall, nrm[at], nrm[nplane + at],
                          nrm[2 * nplane + at], fc, Fm, Fp, Fe);
            const double a = area_u[at];
            double* fo = fbuf + at * (DIM + 2);
            fo[0] = Fm * a;
#pragma unroll
            for (int d = 0; d < DIM; d++) fo[1 + d] = Fp[d] * a;
            fo[1 + DIM] = Fe * a;
        }
    }
}

// dt: the device scalar (capped at dt_max > 0), or per box dt_box[hbox[i]]
template <int DIM>
__global__ void __launch_bounds__(TB)
    k_face_sum(const int64_t* n_dev, int64_t n_host, int64_t stride, int64_t nplane, const int32_t* __restrict__ hcode,
               const uint8_t* __restrict__ count, const uint8_t* __restrict__ twin, const int32_t* __restrict__ nbr,
               const double* __restrict__ fbuf, const double* dt_dev, double dt_max, const double* __restrict__ dt_box,
               const int32_t* __restrict__ hbox, SimView sv) {
    const int64_t n = seed_count(n_dev, n_host);
    double dt0 = 0.0;
    if (dt_dev) {
        dt0 = *dt_dev;
        if (dt_max > 0 && dt0 > dt_max) dt0 = dt_max;
    }
    for (int64_t s = (int64_t)blockIdx.x * TB + threadIdx.x; s < n; s += (int64_t)gridDim.x * TB) {
        const int32_t i = hcode[s];
        if (i < 0) continue;
        const double dt = dt_box ? dt_box[hbox[i]] : dt0;
        double q[DIM + 2];
#pragma unroll
        for (int k = 0; k < DIM + 2; k++) q[k] = 0.0;
        const int c = count[s];
        // four faces' records are loaded together (independent loads in
        // flight), then added in face order
        for (int f0 = 0; f0 < c; f0 += 4) {
            double v[4][DIM + 2];
            bool own[4];
#pragma unroll
            for (int e = 0; e < 4; e++) {
                own[e] = true;
#pragma unroll
                for (int k = 0; k < DIM + 2; k++) v[e][k] = 0.0;
                if (f0 + e >= c) continue;
                const int64_t at = (int64_t)(f0 + e) * stride + s;
                const uint8_t tw = twin[at];
                own[e] = tw == SOLVE;
                const double* fo = fbuf + (own[e] ? at : (int64_t)tw * stride + nbr[at]) * (DIM + 2);
#pragma unroll
                for (int k = 0; k < DIM + 2; k++) v[e][k] = fo[k];
            }
#pragma unroll
            for (int e = 0; e < 4; e++) {
                if (f0 + e >= c) break;
#pragma unroll
                for (int k = 0; k < DIM + 2; k++) {
                    if (own[e]) q[k] += v[e][k];
                    else q[k] -= v[e][k];
                }
            }
        }
        sv.mass[i] -= dt * q[0];
#pragma unroll
        for (int d = 0; d < DIM; d++) sv.mom[(size_t)DIM * i + d] -= dt * q[1 + d];
        sv.en[i] -= dt * q[DIM + 1];
    }
}

// Steered generator velocity (mesh regularisation, Springel 2010, MNRAS 401,
// 791, sec. 4.2): w = v + chi f (c - r) / d, with c the cell centroid, d =
// |c - r|, R the cell radius of the cell volume, c_s the sound speed and
// f = 0 for d < 0.9 eta R, c_s (d - 0.9 eta R) / (0.2 eta R) up to
// 1.1 eta R, c_s beyond (eta = 0.25, chi = 1).  Generators drift towards
// their centroids once cells grow lopsided, which keeps them round.  No
// reference counterpart.
template <int DIM>
__global__ void __launch_bounds__(TB)
    k_gen_velocity(int64_t n_h, const double* __restrict__ x, const int32_t* __restrict__ o2s,
                   const double* __restrict__ vol_u, const double* __restrict__ ccen, int64_t stride, double scale,
                   double gam, const double* __restrict__ rho, const double* __restrict__ vel,
                   const double* __restrict__ E, double* __restrict__ w) {
    const int64_t i = (int64_t)blockIdx.x * TB + threadIdx.x;
    if (i >= n_h) return;
    const int32_t s = o2s[i];
    const double V = vol_u[s];
    const double R = DIM == 2 ? sqrt(V / M_PI) : cbrt(3.0 * V / (4.0 * M_PI));
    Prim st;
    st.rho = rho[i];
    st.E = E[i];
    st.v[2] = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; d++) st.v[d] = vel[(size_t)DIM * i + d];
    const double P = pressure<DIM>(gam, st);
    const double cs = (P > 0.0 && st.rho > 0.0) ? sqrt(gam * P / st.rho) : 0.0;
    double dc[3] = {0, 0, 0}, dist = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        const double c = (ccen[(int64_t)d * stride + s] - 0.5) * scale + 0.5;
        dc[d] = c - x[(size_t)DIM * i + d];
        dist += dc[d] * dc[d];
    }
    dist = sqrt(dist);
    const double eta = 0.25, chi = 1.0;
    const double lo = 0.9 * eta * R, hi = 1.1 * eta * R;
    double f = 0.0;
    if (dist > hi) f = cs;
    else if (dist > lo) f = cs * (dist - lo) / (hi - lo);
#pragma unroll
    for (int d = 0; d < DIM; d++) w[(size_t)DIM * i + d] = st.v[d] + (dist > 0.0 ? chi * f * dc[d] / dist : 0.0);
}

// Generators move with the fluid: x += v dt, wrapped (periodic) or mirrored
// at the walls (reflective box).
template <int DIM>
__global__ void __launch_bounds__(TB) k_drift(int64_t n, double* x, const double* vel, const double* dt_dev,
                                              double dt_max, int periodic) {
    int64_t i = (int64_t)blockIdx.x * TB + threadIdx.x;
    if (i >= n) return;
    double dt = *dt_dev;
    if (dt_max > 0 && dt > dt_max) dt = dt_max;
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        double p = x[(size_t)DIM * i + d] + vel[(size_t)DIM * i + d] * dt;
        if (periodic) {
            p -= floor(p);
            if (p >= 1.0) p = 0.0;
        } else {
            if (p < 0.0) p = -p;
            if (p >= 1.0) p = 2.0 - p;
            p = fmin(fmax(p, 0.0), 0.9999999999999999);
        }
        x[(size_t)DIM * i + d] = p;
    }
}

// FP64 totals of the conserved variables (one atomic-free partial per block).
template <int DIM>
__global__ void __launch_bounds__(TB) k_totals(SimView sv, double* part) {
    __shared__ double red[5][TB];
    int64_t i = (int64_t)blockIdx.x * TB + threadIdx.x;
    double v[5] = {0, 0, 0, 0, 0};
    if (i < sv.n_hydro) {
        v[0] = sv.mass[i];
#pragma unroll
        for (int d = 0; d < DIM; d++) v[1 + d] = sv.mom[(size_t)DIM * i + d];
        v[4] = sv.en[i];
    }
#pragma unroll
    for (int q = 0; q < 5; q++) red[q][threadIdx.x] = v[q];
    __syncthreads();
    for (int w = TB / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
#pragma unroll
            for (int q = 0; q < 5; q++) red[q][threadIdx.x] += red[q][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
#pragma unroll
        for (int q = 0; q < 5; q++) part[(size_t)5 * blockIdx.x + q] = red[q][0];
}

// Copy-out: face counts in original order.
__global__ void __launch_bounds__(TB) k_counts_orig(int64_t n, const uint32_t* perm,
                                                    const uint8_t* count, int32_t* co) {
    int64_t s = (int64_t)blockIdx.x * TB + threadIdx.x;
    if (s >= n) return;
    co[perm[s]] = count[s];
}

// Copy-out: packed reference-layout faces of every cell in original order.
__global__ void __launch_bounds__(TB)
    k_pack_faces(int64_t n, int64_t stride, int64_t nplane, int dim, const uint32_t* perm,
                 const uint8_t* count, const int32_t* ptr_orig, const int32_t* nbr,
                 const double* nrm, const double* area, const double* cen, double ascale, double cscale,
                 int32_t* o_nbr, double* o_nrm, double* o_area, double* o_cen) {
    int64_t s = (int64_t)blockIdx.x * TB + threadIdx.x;
    if (s >= n) return;
    const uint32_t o = perm[s];
    const int c = count[s];
    const int64_t base = ptr_orig[o];
    for (int f = 0; f < c; f++) {
        const int64_t at = (int64_t)f * stride + s;
        const int64_t w = base + f;
        const int32_t j = nbr[at];
        o_nbr[w] = j >= 0 ? (int32_t)perm[j] : j;
        o_nrm[3 * w] = nrm[at];
        o_nrm[3 * w + 1] = nrm[nplane + at];
        o_nrm[3 * w + 2] = nrm[2 * nplane + at];
        o_area[w] = ascale * area[at];
        for (int d = 0; d < 3; d++) {
            double x = cen[(int64_t)d * nplane + at];
            o_cen[3 * w + d] = (d < dim && cscale != 1.0) ? (x - 0.5) * cscale + 0.5 : x;
        }
    }
}

// Copy-out: per-cell arrays in original order (seeds unscaled as the
// reference, periodic_mesh.cpp:155-160; failed cells keep zero seeds).
__global__ void __launch_bounds__(TB)
    k_cells_orig(int64_t n, int64_t stride, int dim, const uint32_t* perm, const double* sorted,
                 const int32_t* status, const int32_t* kcert, const double* vol, const double* ccen,
                 double scale, double vscale, double* o_seeds, double* o_vol, int32_t* o_status,
                 int32_t* o_kcert, double* o_ccen) {
    int64_t s = (int64_t)blockIdx.x * TB + threadIdx.x;
    if (s >= n) return;
    const uint32_t o = perm[s];
    const bool ok = status[s] == SUCCESS;
    for (int d = 0; d < 3; d++) {
        double x = (ok && d < dim) ? sorted[(size_t)dim * s + d] : 0.0;
        double cx = ccen[(int64_t)d * stride + s];
        if (scale != 1.0 && d < dim) {
            x = (x - 0.5) * scale + 0.5;
            cx = (cx - 0.5) * scale + 0.5;
        }
        o_seeds[3 * (int64_t)o + d] = x;
        o_ccen[3 * (int64_t)o + d] = cx;
    }
    o_vol[o] = vscale * vol[s];
    o_status[o] = status[s];
    o_kcert[o] = kcert[s];
}

}  // namespace

void launch_finalize(int64_t n, int64_t n_hydro, int64_t stride, const uint32_t* perm,
                     const int64_t* gsrc, const uint8_t* count, const int32_t* nbr,
                     const double* area, const double* vol, double ascale, double vscale,
                     int32_t* hid, double* area_u, double* vol_u, int32_t* o2s, int32_t* hcode, int32_t* bad,
                     hipStream_t s) {
    if (n <= 0) return;
    k_finalize<<<nblk(n), TB, 0, s>>>(n, n_hydro, stride, perm, gsrc, count, nbr, area, vol, ascale,
                                      vscale, hid, area_u, vol_u, o2s, hcode, bad);
}

void launch_hll_static(int dim, int64_t n, int64_t n_hydro, int64_t stride, int64_t nplane,
                       const uint32_t* perm, const uint8_t* count, const int32_t* hid,
                       const double* nrm, const double* area_u, const double* vol_u, double gam,
                       double dt, int flux_mode, const double* rho, const double* vel,
                       const double* E, double* rho2, double* vel2, double* E2, hipStream_t s) {
    if (n <= 0) return;
    if (dim == 2)
        k_hll_static<2><<<nblk(n), TB, 0, s>>>(n, n_hydro, stride, nplane, perm, count, hid, nrm,
                                               area_u, vol_u, gam, dt, flux_mode, rho, vel, E, rho2,
                                               vel2, E2);
    else
        k_hll_static<3><<<nblk(n), TB, 0, s>>>(n, n_hydro, stride, nplane, perm, count, hid, nrm,
                                               area_u, vol_u, gam, dt, flux_mode, rho, vel, E, rho2,
                                               vel2, E2);
}

void launch_dt_cfl(int dim, int64_t n_hydro, const int32_t* o2s, const double* vol_u, double gam,
                   double cfl, const double* rho, const double* vel, const double* E,
                   unsigned long long* out, hipStream_t s, const double* wgen, CflMesh cm) {
    if (n_hydro <= 0) return;
    if (dim == 2)
        k_dt_cfl<2><<<nblk_dt(n_hydro), TB, 0, s>>>(n_hydro, o2s, vol_u, gam, cfl, rho, vel, E, out, wgen, cm);
    else
        k_dt_cfl<3><<<nblk_dt(n_hydro), TB, 0, s>>>(n_hydro, o2s, vol_u, gam, cfl, rho, vel, E, out, wgen, cm);
}

void launch_gen_velocity(int dim, int64_t n_h, const double* x, const int32_t* o2s, const double* vol_u,
                         const double* ccen, int64_t stride, double scale, double gam, const double* rho,
                         const double* vel, const double* E, double* w, hipStream_t s) {
    if (n_h <= 0) return;
    if (dim == 2) k_gen_velocity<2><<<nblk(n_h), TB, 0, s>>>(n_h, x, o2s, vol_u, ccen, stride, scale, gam, rho, vel, E, w);
    else k_gen_velocity<3><<<nblk(n_h), TB, 0, s>>>(n_h, x, o2s, vol_u, ccen, stride, scale, gam, rho, vel, E, w);
}

void launch_q_from_prim(int dim, SimView sv, const int32_t* o2s, const double* vol_u, hipStream_t s) {
    if (sv.n_hydro <= 0) return;
    if (dim == 2) k_q_from_prim<2><<<nblk(sv.n_hydro), TB, 0, s>>>(sv, o2s, vol_u);
    else k_q_from_prim<3><<<nblk(sv.n_hydro), TB, 0, s>>>(sv, o2s, vol_u);
}

void launch_prim_from_q(int dim, SimView sv, const int32_t* o2s, const double* vol_u, hipStream_t s) {
    if (sv.n_hydro <= 0) return;
    if (dim == 2) k_prim_from_q<2><<<nblk(sv.n_hydro), TB, 0, s>>>(sv, o2s, vol_u);
    else k_prim_from_q<3><<<nblk(sv.n_hydro), TB, 0, s>>>(sv, o2s, vol_u);
}

// grid: every seed once (host count), or a capped grid-stride grid over
// the capacity (device count)
static unsigned own_blocks(const int64_t* n_dev, int64_t n) {
    return n_dev ? std::min<unsigned>(nblk(n), 4096) : nblk(n);
}

void launch_face_twins(const int64_t* n_dev, int64_t n, int64_t stride, const int32_t* hcode, const uint8_t* count,
                       const int32_t* nbr_cm, uint8_t* twin, uint64_t* solve, hipStream_t s) {
    if (n <= 0) return;
    k_face_twins<<<own_blocks(n_dev, n), TB, 0, s>>>(n_dev, n, stride, hcode, count, nbr_cm, twin, solve);
}

void launch_hll_owned(int dim, OwnedUpdate u, hipStream_t s) {
    if (u.n <= 0) return;
    const unsigned gb = own_blocks(u.n_dev, u.n);
    const dim3 gf(gb, FR);
    if (dim == 2) {
        k_cell_records<2><<<gb, TB, 0, s>>>(u.n_dev, u.n, u.hcode, u.sv, u.wgen, u.rec);
        k_face_flux<2><<<gf, TB, 0, s>>>(u.n_dev, u.n, u.stride, u.nplane, u.hcode, u.solve, u.nbr, u.nrm,
                                         u.area_u, u.cen, u.sorted, u.scale, u.gam, u.flux_mode, u.move, u.rec,
                                         u.fbuf);
        k_face_sum<2><<<gb, TB, 0, s>>>(u.n_dev, u.n, u.stride, u.nplane, u.hcode, u.count, u.twin, u.nbr, u.fbuf,
                                        u.dt_dev, u.dt_max, u.dt_box, u.hbox, u.sv);
    } else {
        k_cell_records<3><<<gb, TB, 0, s>>>(u.n_dev, u.n, u.hcode, u.sv, u.wgen, u.rec);
        k_face_flux<3><<<gf, TB, 0, s>>>(u.n_dev, u.n, u.stride, u.nplane, u.hcode, u.solve, u.nbr, u.nrm,
                                         u.area_u, u.cen, u.sorted, u.scale, u.gam, u.flux_mode, u.move, u.rec,
                                         u.fbuf);
        k_face_sum<3><<<gb, TB, 0, s>>>(u.n_dev, u.n, u.stride, u.nplane, u.hcode, u.count, u.twin, u.nbr, u.fbuf,
                                        u.dt_dev, u.dt_max, u.dt_box, u.hbox, u.sv);
    }
}

void launch_hll_faces(int dim, int64_t n, double gam, const double* L, const double* R, const double* nrm,
                      const uint8_t* wall, int flux_mode, double* F, hipStream_t s) {
    if (n <= 0) return;
    const double sgn = flux_mode == 1 ? -1.0 : 1.0;
    if (dim == 2) k_hll_faces<2><<<nblk(n), TB, 0, s>>>(n, gam, L, R, nrm, wall, sgn, F);
    else k_hll_faces<3><<<nblk(n), TB, 0, s>>>(n, gam, L, R, nrm, wall, sgn, F);
}

void launch_drift(int dim, int64_t n, double* x, const double* vel, const double* dt_dev, double dt_max,
                  int periodic, hipStream_t s) {
    if (n <= 0) return;
    if (dim == 2) k_drift<2><<<nblk(n), TB, 0, s>>>(n, x, vel, dt_dev, dt_max, periodic);
    else k_drift<3><<<nblk(n), TB, 0, s>>>(n, x, vel, dt_dev, dt_max, periodic);
}

int64_t totals_blocks(int64_t n) { return (int64_t)nblk(n); }

void launch_totals(int dim, SimView sv, double* part, hipStream_t s) {
    if (sv.n_hydro <= 0) return;
    if (dim == 2) k_totals<2><<<nblk(sv.n_hydro), TB, 0, s>>>(sv, part);
    else k_totals<3><<<nblk(sv.n_hydro), TB, 0, s>>>(sv, part);
}

void launch_counts_orig(int64_t n, const uint32_t* perm, const uint8_t* count, int32_t* co,
                        hipStream_t s) {
    if (n <= 0) return;
    k_counts_orig<<<nblk(n), TB, 0, s>>>(n, perm, count, co);
}

void launch_pack_faces(int64_t n, int64_t stride, int64_t nplane, int dim, const uint32_t* perm,
                       const uint8_t* count, const int32_t* ptr_orig, const int32_t* nbr,
                       const double* nrm, const double* area, const double* cen, double ascale,
                       double cscale, int32_t* o_nbr, double* o_nrm, double* o_area, double* o_cen,
                       hipStream_t s) {
    if (n <= 0) return;
    k_pack_faces<<<nblk(n), TB, 0, s>>>(n, stride, nplane, dim, perm, count, ptr_orig, nbr, nrm, area,
                                        cen, ascale, cscale, o_nbr, o_nrm, o_area, o_cen);
}

void launch_cells_orig(int64_t n, int64_t stride, int dim, const uint32_t* perm,
                       const double* sorted, const int32_t* status, const int32_t* kcert,
                       const double* vol, const double* ccen, double scale, double vscale,
                       double* o_seeds, double* o_vol, int32_t* o_status, int32_t* o_kcert,
                       double* o_ccen, hipStream_t s) {
    if (n <= 0) return;
    k_cells_orig<<<nblk(n), TB, 0, s>>>(n, stride, dim, perm, sorted, status, kcert, vol, ccen, scale,
                                        vscale, o_seeds, o_vol, o_status, o_kcert, o_ccen);
}

}  // namespace pg
