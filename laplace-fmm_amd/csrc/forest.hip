// forest.hip -- many independent systems mapped into one unit cube, so that
// one adaptive octree solves all of them (lfmm_fmm3d_batch with
// LFMM_FLAG_ADAPTIVE; the layout is described at ForestSpec).
//
// Per system the tree geometry follows the reference's extent rule
// (src/cpu/morton_tree.cpp:293-301): shift = min - 1e-7 per axis, scale =
// largest extent + 1e-6, x -> u = (x - shift) / scale in [0, 1)^3, the
// coordinates its own tree would have.  The forest places u in the system's
// cell: U = (2a + u_x, 2b + u_y, 2c + u_z) / 2^level.
#include "forest.h"

#include <cmath>

namespace lfmm {

namespace {

// One workgroup per system: exact min and max per axis, and whether any
// coordinate or charge is not finite.
__global__ void __launch_bounds__(256) k_forest_extent(const double* __restrict__ pos, const double* __restrict__ q,
                                                       const int64_t* __restrict__ off, int n_sys,
                                                       double* __restrict__ ext, int* __restrict__ bad)
{
    __shared__ double s[6][256];
    const int sys = blockIdx.x;
    if (sys >= n_sys) return;
    const int64_t b = off[sys], e = off[sys + 1];
    double v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    bool nf = false;
    for (int64_t i = b + threadIdx.x; i < e; i += blockDim.x) {
        for (int a = 0; a < 3; ++a) {
            const double x = pos[3 * i + a];
            nf |= !isfinite(x);
            v[a] = fmin(v[a], x);
            v[3 + a] = fmax(v[3 + a], x);
        }
        nf |= !isfinite(q[i]);
    }
    for (int a = 0; a < 6; ++a) s[a][threadIdx.x] = v[a];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w)
            for (int a = 0; a < 3; ++a) {
                s[a][threadIdx.x] = fmin(s[a][threadIdx.x], s[a][threadIdx.x + w]);
                s[3 + a][threadIdx.x] = fmax(s[3 + a][threadIdx.x], s[3 + a][threadIdx.x + w]);
            }
        __syncthreads();
    }
    if (nf) atomicOr(bad, 1);
    if (threadIdx.x == 0)
        for (int a = 0; a < 6; ++a) ext[6 * sys + a] = s[a][0];
}

// Particle i of system s (off[s] <= i < off[s+1], found by bisection) to its
// place in the forest, packed as (x, y, z, q).
__global__ void __launch_bounds__(256) k_forest_map(const double* __restrict__ pos, const double* __restrict__ q,
                                                    const int64_t* __restrict__ off, int n_sys, int64_t n,
                                                    const double* __restrict__ geo, int side, double cell,
                                                    double4* __restrict__ out)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    int lo = 0, hi = n_sys - 1;  // the last s with off[s] <= i
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= i) lo = mid;
        else hi = mid - 1;
    }
    const int s = lo;
    const double* g = geo + 4 * s;  // shift x, y, z, scale
    const int ca = s / (side * side), cb = (s / side) % side, cc = s % side;
    const double ux = (pos[3 * i] - g[0]) / g[3];
    const double uy = (pos[3 * i + 1] - g[1]) / g[3];
    const double uz = (pos[3 * i + 2] - g[2]) / g[3];
    out[i] = make_double4((2.0 * ca + ux) * cell, (2.0 * cb + uy) * cell, (2.0 * cc + uz) * cell, q[i]);
}

}  // namespace

hipError_t launch_forest_extent(const double* pos, const double* q, const int64_t* off, int n_sys, double* ext,
                                int* bad, hipStream_t st)
{
    if (n_sys > 0) hipLaunchKernelGGL(k_forest_extent, dim3(n_sys), dim3(256), 0, st, pos, q, off, n_sys, ext, bad);
    return hipGetLastError();
}

hipError_t launch_forest_map(const double* pos, const double* q, const int64_t* off, int n_sys, int64_t n,
                             const double* geo, int side, int level, double4* out, hipStream_t st)
{
    if (n > 0)
        hipLaunchKernelGGL(k_forest_map, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, pos, q, off, n_sys, n,
                           geo, side, std::ldexp(1.0, -level), out);
    return hipGetLastError();
}

}  // namespace lfmm
