// direct.h -- FP64 all-pairs sums (direct.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace lfmm {

// pot[i] (and field[dim * i ...]) over all sources for every target; packed
// (x, y, z, q) records (2D: (x, y, 0, q)).  exclude_self skips j == i.
hipError_t launch_direct(int dim, const double4* tgt, int64_t nt, const double4* src, int64_t ns, bool exclude_self,
                         double* pot, double* field, hipStream_t st);

// 3D sums of few targets over many sources: the sources are cut into
// `splits` chunks (direct3d_split_count: enough workgroups to fill the GPU),
// each chunk's sums go to part (splits * nt records) and are added in chunk
// order.  No self exclusion.
int direct3d_split_count(int64_t nt, int64_t ns);
hipError_t launch_direct3d_split(const double4* tgt, int64_t nt, const double4* src, int64_t ns, double* pot,
                                 double* field, double4* part, int splits, hipStream_t st);

// e3 <- the field of vector charges from the fields e1, e2, e3 of their
// x, y, z components (n x 3 doubles each).
hipError_t launch_cross_fields(const double* e1, const double* e2, double* e3, int64_t n, hipStream_t st);

// Fixed scratch of lfmm_direct3d_noalloc: 32 MiB per GPU, 1,048,576 packed
// records (sources plus targets of one call).
constexpr size_t kNoAllocScratchBytes = (size_t)32 << 20;
// Looks up this device's copy of the scratch array (with `device` current);
// once per device, at context creation.  noalloc_scratch returns it (null
// before the lookup).
hipError_t noalloc_scratch_init(int device);
double4* noalloc_scratch(int device);

}  // namespace lfmm
