// forest.h -- many independent systems in one unit cube (forest.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace lfmm {

// ext[6 s ..]: min x, y, z and max x, y, z of system s (particles
// off[s] .. off[s+1]-1 of pos, 3 doubles each); *bad |= 1 when a coordinate
// or charge is not finite.  Device pointers.
hipError_t launch_forest_extent(const double* pos, const double* q, const int64_t* off, int n_sys, double* ext,
                                int* bad, hipStream_t st);
// out[i] = particle i in the forest: (2a + (x - shift) / scale) / 2^level per
// axis with the system's cell (a, b, c) = (s / side^2, s / side % side,
// s % side) and geo[4 s ..] = (shift x, y, z, scale); w = q[i].
hipError_t launch_forest_map(const double* pos, const double* q, const int64_t* off, int n_sys, int64_t n,
                             const double* geo, int side, int level, double4* out, hipStream_t st);

}  // namespace lfmm
