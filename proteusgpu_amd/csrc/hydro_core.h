// Per-face Riemann problem and per-cell flux sums of the hydro update
// (device code shared by hydro.hip and batch.hip).
#pragma once

#include "common.h"
#include "hydro.h"

namespace pg {

struct Prim {
    double rho, v[3], E;
};

template <int DIM>
PG_DEV double pressure(double gam, const Prim& s) {
    if (DIM == 2) return (gam - 1) * (s.E - (0.5 * s.rho * (s.v[0] * s.v[0] + s.v[1] * s.v[1])));
    return (gam - 1) *
           (s.E - (0.5 * s.rho * (s.v[0] * s.v[0] + s.v[1] * s.v[1] + s.v[2] * s.v[2])));
}

// riemann.cpp:75-97
struct Geom {
    double n[3], m[3], p[3];
};
PG_DEV Geom geom_of(double n0, double n1, double n2) {
    Geom g;
    const double nn = sqrt(n0 * n0 + n1 * n1 + n2 * n2);
    g.n[0] = n0 / nn;
    g.n[1] = n1 / nn;
    g.n[2] = n2 / nn;
    if (g.n[0] != 0.0 || g.n[1] != 0.0) {
        g.m[0] = -g.n[1];
        g.m[1] = g.n[0];
        g.m[2] = 0.0;
    } else {
        g.m[0] = 1.0;
        g.m[1] = 0.0;
        g.m[2] = 0.0;
    }
    const double mm = sqrt(g.m[0] * g.m[0] + g.m[1] * g.m[1] + g.m[2] * g.m[2]);
    g.m[0] = g.m[0] / mm;
    g.m[1] = g.m[1] / mm;
    g.m[2] = g.m[2] / mm;
    g.p[0] = g.n[1] * g.m[2] - g.n[2] * g.m[1];
    g.p[1] = g.n[2] * g.m[0] - g.n[0] * g.m[2];
    g.p[2] = g.n[0] * g.m[1] - g.n[1] * g.m[0];
    return g;
}
template <int DIM>
PG_DEV void to_face(Prim& s, const Geom& g) {
    const double vx = s.v[0], vy = s.v[1];
    if (DIM == 2) {
        s.v[0] = vx * g.n[0] + vy * g.n[1];
        s.v[1] = vx * g.m[0] + vy * g.m[1];
    } else {
        const double vz = s.v[2];
        s.v[0] = vx * g.n[0] + vy * g.n[1] + vz * g.n[2];
        s.v[1] = vx * g.m[0] + vy * g.m[1] + vz * g.m[2];
        s.v[2] = vx * g.p[0] + vy * g.p[1] + vz * g.p[2];
    }
}
template <int DIM>
PG_DEV void from_face(Prim& s, const Geom& g) {
    const double vx = s.v[0], vy = s.v[1];
    if (DIM == 2) {
        s.v[0] = vx * g.n[0] + vy * g.m[0];
        s.v[1] = vx * g.n[1] + vy * g.m[1];
    } else {
        const double vz = s.v[2];
        s.v[0] = vx * g.n[0] + vy * g.m[0] + vz * g.p[0];
        s.v[1] = vx * g.n[1] + vy * g.m[1] + vz * g.p[1];
        s.v[2] = vx * g.n[2] + vy * g.m[2] + vz * g.p[2];
    }
}
template <int DIM>
PG_DEV Prim euler_flux(double gam, const Prim& s) {
    const double P = pressure<DIM>(gam, s);
    Prim f;
    f.rho = s.rho * s.v[0];
    f.v[0] = s.rho * s.v[0] * s.v[0] + P;
    f.v[1] = s.rho * s.v[0] * s.v[1];
    f.v[2] = DIM == 3 ? s.rho * s.v[0] * s.v[2] : 0.0;
    f.E = (s.E + P) * s.v[0];
    return f;
}

// riemann.cpp:7-45 with the reference's operation order.  wall: the right
// state is the left one mirrored in the face frame, so SL == -SR exactly and
// the mass flux through the wall is exactly zero.
template <int DIM>
PG_DEV Prim hll(double gam, Prim L, Prim R, const Geom& g, bool wall) {
    to_face<DIM>(L, g);
    if (wall) {
        R = L;
        R.v[0] = -L.v[0];
    } else {
        to_face<DIM>(R, g);
    }
    const Prim fl = euler_flux<DIM>(gam, L), fr = euler_flux<DIM>(gam, R);
    const double cl = sqrt((gam * pressure<DIM>(gam, L)) / L.rho);
    const double cr = sqrt((gam * pressure<DIM>(gam, R)) / R.rho);
    const double SL = smin(L.v[0] - cl, R.v[0] - cr);
    const double SR = smax(L.v[0] + cl, R.v[0] + cr);
    Prim F = {0.0, {0.0, 0.0, 0.0}, 0.0};
    if (SL >= 0) {
        F = fl;
    } else if (SL < 0 && SR > 0) {
        F.rho = (SR * fl.rho - SL * fr.rho + SL * SR * (R.rho - L.rho)) / (SR - SL);
#pragma unroll
        for (int d = 0; d < DIM; d++)
            F.v[d] = (SR * fl.v[d] - SL * fr.v[d] + SL * SR * (R.rho * R.v[d] - L.rho * L.v[d])) /
                     (SR - SL);
        F.E = (SR * fl.E - SL * fr.E + SL * SR * (R.E - L.E)) / (SR - SL);
    } else if (SR <= 0) {
        F = fr;
    }
    from_face<DIM>(F, g);
    return F;
}

// CFL length of sorted cell s with volume V: the radius of the volume
// (finite_volume_solver.cpp:170-175); with regularised mesh motion also at
// most the distance to the nearest face neighbour, so that slivers (two
// generators close together: small width, large radius) limit the step as
// well.  Round cells have a nearest neighbour beyond R (2D hexagons 1.9 R,
// 3D cubes 1.6 R), so the limit acts on slivers only.
template <int DIM>
PG_DEV double cfl_length(double V, int32_t s, const CflMesh& cm) {
    double R = DIM == 2 ? sqrt(V / M_PI) : cbrt(3.0 * V / (4.0 * M_PI));
    if (cm.count == nullptr) return R;
    const int c = cm.count[s];
    double dmin = 1e300;
    for (int f = 0; f < c; f++) {
        const int32_t js = cm.nbr[(int64_t)f * cm.stride + s];
        if (js < 0) continue;
        double dd = 0.0;
#pragma unroll
        for (int d = 0; d < DIM; d++) {
            const double t = cm.sorted[(size_t)DIM * js + d] - cm.sorted[(size_t)DIM * s + d];
            dd += t * t;
        }
        dmin = fmin(dmin, dd);
    }
    return fmin(R, sqrt(dmin) * cm.scale);
}

// |v - w| (signal speed of the fluid relative to a moving generator)
template <int DIM>
PG_DEV double rel_speed(const double* v, const double* w) {
    double s = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; d++) s += (v[d] - w[d]) * (v[d] - w[d]);
    return sqrt(s);
}

// CFL step of hydro cell i (finite_volume_solver.cpp:151-186): CFL number
// times the cell's CFL length over its sound plus signal speed (with mesh
// motion, the larger of |v| and |v - w|).  A bad state (negative or NaN
// step) returns -1, which the minimum then reports as a zero step.
template <int DIM>
PG_DEV double cell_cfl_dt(int64_t i, const int32_t* __restrict__ o2s, const double* __restrict__ vol_u,
                          double gam, double cfl, const double* __restrict__ rho,
                          const double* __restrict__ vel, const double* __restrict__ E,
                          const double* __restrict__ wgen, const CflMesh& cm) {
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
    const double dti = cfl * (R / (c + va));
    return dti >= 0.0 ? dti : -1.0;
}

// Minimum step into *out (non-negative doubles order as their bits; a
// negative minimum, a bad state, reports 0).
PG_DEV void report_dt(unsigned long long* out, double m) {
    if (m >= 0.0) atomic_min_pos_double(out, m);
    else atomicMin(out, 0ull);
}

// Consecutive cells per lane of the CFL reductions: a wave takes 64 x 4
// consecutive cells (lane-coalesced loads), and a block of 256 lanes reports
// once per 1024 cells, so a 1M-cell step makes about 1k atomic updates of
// its minimum instead of one per wave.
constexpr int DT_CELLS = 4;

// Flux of one face in the lab frame, per unit area (see k_face_flux): own
// state si at generator ri moving with gi, neighbour state sj at rj moving
// with gj (ignored at walls), stored normal n and face centroid fc (both
// unscaled for rj/ri/fc).  The face moves with w = (gi + gj)/2 plus the
// correction for the offset of fc from the generators' midpoint; the
// Riemann problem is solved in the face frame and transformed back.
template <int DIM>
PG_DEV void ale_face(double gam, double sgn, int move, const Prim& si, const double (&ri)[3],
                     const double (&gi)[3], const Prim& sj, const double (&rj)[3], const double (&gj)[3],
                     bool wall, double n0, double n1, double n2, const double (&fc)[3], double& Fm,
                     double (&Fp)[3], double& Fe) {
    double w[3] = {0.0, 0.0, 0.0};
    if (move && !wall) {
        double dr[3], mid[3];
        double dd = 0.0, proj = 0.0;
#pragma unroll
        for (int d = 0; d < DIM; d++) {
            dr[d] = rj[d] - ri[d];
            mid[d] = 0.5 * (ri[d] + rj[d]);
            dd += dr[d] * dr[d];
            proj += (gi[d] - gj[d]) * (fc[d] - mid[d]);
        }
#pragma unroll
        for (int d = 0; d < DIM; d++) w[d] = 0.5 * (gi[d] + gj[d]) + proj * dr[d] / dd;
    }
    // total energy per volume is frame dependent: boost it with the velocity
    Prim L = si, R = wall ? si : sj;
    if (move && !wall) {
        double ke_l = 0, ke_r = 0, ke_l2 = 0, ke_r2 = 0;
#pragma unroll
        for (int d = 0; d < DIM; d++) {
            ke_l += L.v[d] * L.v[d];
            ke_r += R.v[d] * R.v[d];
            L.v[d] -= w[d];
            R.v[d] -= w[d];
            ke_l2 += L.v[d] * L.v[d];
            ke_r2 += R.v[d] * R.v[d];
        }
        L.E += 0.5 * L.rho * (ke_l2 - ke_l);
        R.E += 0.5 * R.rho * (ke_r2 - ke_r);
    }
    const Geom g = geom_of(sgn * n0, sgn * n1, sgn * n2);
    const Prim F = hll<DIM>(gam, L, R, g, wall);
    Fm = F.rho;
    Fp[0] = F.v[0];
    Fp[1] = F.v[1];
    Fp[2] = F.v[2];
    Fe = F.E;
    if (move && !wall) {
        double wp = 0.0, ww = 0.0;
#pragma unroll
        for (int d = 0; d < DIM; d++) {
            wp += Fp[d] * w[d];
            ww += w[d] * w[d];
        }
        Fe += wp + 0.5 * ww * Fm;
#pragma unroll
        for (int d = 0; d < DIM; d++) Fp[d] += Fm * w[d];
    }
}

}  // namespace pg
