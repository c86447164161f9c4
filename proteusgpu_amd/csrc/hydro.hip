// Hydro on the Voronoi mesh: mesh finalisation, the HLL update
// (finite_volume_solver.cpp:43-148, riemann.cpp:7-124), the CFL reduce
// (finite_volume_solver.cpp:151-186) and the reference-layout copy-out of
// the mesh (voronoi.cpp:212-333 unpermute, periodic_mesh.cpp:146-166 unscale).
//
// Face data come from the clip kernel in face-slot-major padded storage
// (entry (f, sorted cell s) at f*stride + s), so threads of a wave that handle
// consecutive sorted cells read consecutive addresses for every face slot.
#include "common.h"
#include "hydro.h"
#include "hydro_core.h"

namespace pg {
namespace {

constexpr int TB = 256;
inline unsigned nblk(int64_t n) { return (unsigned)((n + TB - 1) / TB); }

// Per sorted cell: hydro id of every face's neighbour (-1 = box plane,
// which is a reflective wall), unscaled area and cell volume.
__global__ void __launch_bounds__(TB) k_finalize(int64_t n, int64_t n_hydro, int64_t stride,
                                                 const uint32_t* perm, const int64_t* gsrc,
                                                 const uint8_t* count, const int32_t* nbr,
                                                 const double* area, const double* vol, double ascale,
                                                 double vscale, int32_t* hid, double* area_u,
                                                 double* vol_u, int32_t* o2s, int32_t* bad) {
    int64_t s = (int64_t)blockIdx.x * TB + threadIdx.x;
    if (s >= n) return;
    const uint32_t o = perm[s];
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
    int64_t i = (int64_t)blockIdx.x * TB + threadIdx.x;
    double dti = 1e100;
    if (i < n_hydro) {
        Prim s;
        s.rho = rho[i];
        s.E = E[i];
        s.v[2] = 0.0;
#pragma unroll
        for (int d = 0; d < DIM; d++) s.v[d] = vel[(size_t)DIM * i + d];
        const double P = pressure<DIM>(gam, s);
        const double c = sqrt(gam * P / s.rho);
        const int32_t si = o2s[i];
        const double R = cfl_length<DIM>(vol_u[si], si, cm);
        double va = sqrt(s.v[0] * s.v[0] + s.v[1] * s.v[1] + s.v[2] * s.v[2]);
        if (wgen) va = fmax(va, rel_speed<DIM>(s.v, wgen + (size_t)DIM * i));
        dti = cfl * (R / (c + va));
    }
    // wave min, then one atomic per wave (non-negative doubles order as bits)
    for (int d = 32; d > 0; d >>= 1) dti = smin(__shfl_xor(dti, d, 64), dti);
    if ((threadIdx.x & 63) == 0) {
        if (dti >= 0.0) atomic_min_pos_double(out, dti);
        else atomicMin(out, 0ull);  // a negative step (bad state) reports 0
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

// Moving-mesh (ALE) HLL update of the conserved state.  The face moves with
// w_f = (w_i + w_j)/2 plus the standard correction for the offset of the face
// centroid from the generators' midpoint; the Riemann problem is solved in
// the face frame and transformed back (mass flux unchanged, momentum
// + F_m w_f, energy + F_p.w_f + |w_f|^2 F_m / 2).  With move == 0 this is the
// static flux of k_hll_static applied to extensive variables.  Walls do not
// move (w_f = 0).  No reference counterpart: the reference never moves its mesh.
template <int DIM>
__global__ void __launch_bounds__(TB)
    k_hll_ale(int64_t n, int64_t n_hydro, int64_t stride, int64_t nplane, const uint32_t* __restrict__ perm,
              const uint8_t* __restrict__ count, const int32_t* __restrict__ hid,
              const int32_t* __restrict__ nbr, const double* __restrict__ nrm,
              const double* __restrict__ area_u, const double* __restrict__ cen,
              const double* __restrict__ sorted, double scale, double gam, const double* dt_dev,
              double dt_max, int flux_mode, int move, SimView sv, const double* __restrict__ wgen) {
    int64_t s = (int64_t)blockIdx.x * TB + threadIdx.x;
    if (s >= n) return;
    const uint32_t i = perm[s];
    if (i >= n_hydro) return;
    double dt = *dt_dev;
    if (dt_max > 0 && dt > dt_max) dt = dt_max;
    double fm, fp[3], fe;
    ale_cell_flux<DIM>(s, i, stride, nplane, count, hid, nbr, nrm, area_u, cen, sorted, scale, gam, flux_mode,
                       move, sv, wgen, fm, fp, fe);
    sv.mass[i] -= dt * fm;
#pragma unroll
    for (int d = 0; d < DIM; d++) sv.mom[(size_t)DIM * i + d] -= dt * fp[d];
    sv.en[i] -= dt * fe;
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
                     int32_t* hid, double* area_u, double* vol_u, int32_t* o2s, int32_t* bad,
                     hipStream_t s) {
    if (n <= 0) return;
    k_finalize<<<nblk(n), TB, 0, s>>>(n, n_hydro, stride, perm, gsrc, count, nbr, area, vol, ascale,
                                      vscale, hid, area_u, vol_u, o2s, bad);
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
        k_dt_cfl<2><<<nblk(n_hydro), TB, 0, s>>>(n_hydro, o2s, vol_u, gam, cfl, rho, vel, E, out, wgen, cm);
    else
        k_dt_cfl<3><<<nblk(n_hydro), TB, 0, s>>>(n_hydro, o2s, vol_u, gam, cfl, rho, vel, E, out, wgen, cm);
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

void launch_hll_ale(int dim, int64_t n, int64_t n_hydro, int64_t stride, int64_t nplane, const uint32_t* perm,
                    const uint8_t* count, const int32_t* hid, const int32_t* nbr, const double* nrm,
                    const double* area_u, const double* cen, const double* sorted, double scale, double gam,
                    const double* dt_dev, double dt_max, int flux_mode, int move, SimView sv, hipStream_t s,
                    const double* wgen) {
    if (n <= 0) return;
    if (dim == 2)
        k_hll_ale<2><<<nblk(n), TB, 0, s>>>(n, n_hydro, stride, nplane, perm, count, hid, nbr, nrm, area_u, cen,
                                            sorted, scale, gam, dt_dev, dt_max, flux_mode, move, sv, wgen);
    else
        k_hll_ale<3><<<nblk(n), TB, 0, s>>>(n, n_hydro, stride, nplane, perm, count, hid, nbr, nrm, area_u, cen,
                                            sorted, scale, gam, dt_dev, dt_max, flux_mode, move, sv, wgen);
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
