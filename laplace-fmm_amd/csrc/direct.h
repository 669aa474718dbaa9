// direct.h -- FP64 all-pairs sums (direct.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace lfmm {

// pot[i] (and field[dim * i ...]) over all sources for every target; packed
// (x, y, z, q) records (2D: (x, y, 0, q)).  exclude_self skips j == i.
hipError_t launch_direct(int dim, const double4* tgt, int64_t nt, const double4* src, int64_t ns, bool exclude_self,
                         double* pot, double* field, hipStream_t st);

// e3 <- the field of vector charges from the fields e1, e2, e3 of their
// x, y, z components (n x 3 doubles each).
hipError_t launch_cross_fields(const double* e1, const double* e2, double* e3, int64_t n, hipStream_t st);

}  // namespace lfmm
