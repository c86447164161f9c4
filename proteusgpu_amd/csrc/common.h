// Shared device helpers for the MI355X (gfx950) kernels of libpgpu.
//
// Every translation unit is compiled with -ffp-contract=off: the FP64
// expressions below reproduce the reference's rounding (x86-64, no FMA) bit
// for bit.  Where a fused multiply-add is wanted (filters whose error bound
// accounts for it) it is written as an explicit fma().
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define PG_DEV __device__ __forceinline__

namespace pg {

constexpr int WAVE = 64;
constexpr uint8_t END = 255;

// Status numbering of the reference (allvars.h:13-20) plus one internal code.
enum : int {
    TRIANGLE_OVERFLOW = 0,
    VERTEX_OVERFLOW = 1,
    INCONSISTENT_BOUNDARY = 2,
    SECURITY_RADIUS_NOT_REACHED = 3,
    SUCCESS = 4,
    NEEDS_EXACT_PREDICATES = 5,
    // internal: the fast kernel ran out of vertex slots before the
    // configured max_t; the seed is recomputed by the large-capacity kernel.
    SLOT_OVERFLOW = 16,
};

// allvars.h:168-195.  Operand order and association are the reference's.
PG_DEV double det2(double a11, double a12, double a21, double a22) { return a11 * a22 - a12 * a21; }
PG_DEV double det3(double a11, double a12, double a13, double a21, double a22, double a23,
                   double a31, double a32, double a33) {
    return a11 * det2(a22, a23, a32, a33) - a21 * det2(a12, a13, a32, a33) +
           a31 * det2(a12, a13, a22, a23);
}
PG_DEV double det4(double a11, double a12, double a13, double a14, double a21, double a22,
                   double a23, double a24, double a31, double a32, double a33, double a34,
                   double a41, double a42, double a43, double a44) {
    double m12 = a21 * a12 - a11 * a22;
    double m13 = a31 * a12 - a11 * a32;
    double m14 = a41 * a12 - a11 * a42;
    double m23 = a31 * a22 - a21 * a32;
    double m24 = a41 * a22 - a21 * a42;
    double m34 = a41 * a32 - a31 * a42;
    double m123 = m23 * a13 - m13 * a23 + m12 * a33;
    double m124 = m24 * a13 - m14 * a23 + m12 * a43;
    double m134 = m34 * a13 - m14 * a33 + m13 * a43;
    double m234 = m34 * a23 - m24 * a33 + m23 * a43;
    return m234 * a14 - m134 * a24 + m124 * a34 - m123 * a44;
}
// std::min / std::max semantics (riemann.cpp:21-22 relies on them).
PG_DEV double smin(double a, double b) { return (b < a) ? b : a; }
PG_DEV double smax(double a, double b) { return (a < b) ? b : a; }

// Workgroups of a launch are dealt to the 8 XCDs round-robin (block b on
// XCD b % 8), each XCD with its own L2.  Renumbering them so that XCD x takes
// a contiguous share of the logical blocks keeps the work of neighbouring
// blocks (neighbouring seeds, sharing neighbour points) in one L2.  A
// bijection of [0, nb) for any nb; only locality depends on it.
PG_DEV int64_t xcd_block(int64_t b, int64_t nb) {
    constexpr int64_t X = 8;
    const int64_t per = nb / X, rem = nb % X;
    const int64_t x = b % X, i = b / X;
    return x * per + (x < rem ? x : rem) + i;
}

// knn.cpp:208-222
template <int DIM>
PG_DEV int grid_cell(const double* p, int N) {
    int i = (int)floor(p[0] * (double)N);
    int j = (int)floor(p[1] * (double)N);
    i = max(0, min(i, N - 1));
    j = max(0, min(j, N - 1));
    if (DIM == 2) return i + j * N;
    int k = (int)floor(p[2] * (double)N);
    k = max(0, min(k, N - 1));
    return i + j * N + k * N * N;
}

// knn.cpp:22: N_grid = round(pow(n / 3.1f, 1.0f / dim)), at least 1.  The
// power is evaluated in FP64 and rounded to float (the correctly rounded
// float power), identically on host and device, so that a box's grid can be
// sized on the GPU exactly as the host sizes it.
__host__ __device__ inline int ngrid_of(int dim, int64_t n) {
    const float x = (float)n / 3.1f;
    const float p = (float)pow((double)x, (double)(1.0f / (float)dim));
    const int N = (int)roundf(p);
    return N < 1 ? 1 : N;
}

// knn.h:64-75
template <int DIM>
PG_DEV double dist2(const double* a, const double* b) {
    double dx = a[0] - b[0];
    double dy = a[1] - b[1];
    if (DIM == 2) return dx * dx + dy * dy;
    double dz = a[2] - b[2];
    return dx * dx + dy * dy + dz * dz;
}

// Positive doubles order like their bit patterns: min-reduce through an
// integer atomic is exact and independent of the order of arrival.
PG_DEV void atomic_min_pos_double(unsigned long long* addr, double v) {
    atomicMin(addr, (unsigned long long)__double_as_longlong(v));
}

}  // namespace pg
