// Ghost regions of the periodic box (device code shared by grid.hip and
// batch.hip).
#pragma once

#include "common.h"

namespace pg {
namespace {

// Region codes per axis: 0 = whole (0,1), 1 = low layer (0,buff) shifted +1,
// 2 = high layer (1-buff,1) shifted -1.  The order is the reference's
// (periodic_mesh.cpp:70-79 in 2D, :83-111 in 3D); it fixes ghost numbering.
__constant__ uint8_t REG2[8][2] = {{1, 0}, {0, 2}, {2, 0}, {0, 1}, {1, 1}, {1, 2}, {2, 2}, {2, 1}};
__constant__ uint8_t REG3[26][3] = {
    {0, 0, 1}, {1, 0, 0}, {0, 2, 0}, {2, 0, 0}, {0, 1, 0}, {0, 0, 2},                        // faces
    {0, 1, 1}, {1, 0, 1}, {0, 2, 1}, {2, 0, 1}, {1, 1, 0}, {1, 2, 0},                        // edges
    {2, 2, 0}, {2, 1, 0}, {1, 0, 2}, {0, 2, 2}, {2, 0, 2}, {0, 1, 2},                        //
    {1, 1, 1}, {1, 2, 1}, {2, 2, 1}, {2, 1, 1}, {1, 1, 2}, {1, 2, 2}, {2, 2, 2}, {2, 1, 2}}; // corners

template <int DIM>
PG_DEV bool in_region(const double* p, const uint8_t* code, double buff) {
    bool in = true;
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        double lo = code[d] == 2 ? 1. - buff : 0.;
        double hi = code[d] == 1 ? buff : 1.;
        in = in && (p[d] > lo && p[d] < hi);  // periodic_mesh.cpp:16-24, strict
    }
    return in;
}

}  // namespace
}  // namespace pg
