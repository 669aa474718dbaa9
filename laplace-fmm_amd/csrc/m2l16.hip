// m2l16.hip -- level-wide helpers of the matrix-core M2L.
#include "ops3d.h"

#include <algorithm>

namespace lfmm {

// Largest |v[i]| over `count` doubles, max-combined into *out as float bits
// (non-negative floats order like their bit patterns).  One atomic per wave.
__global__ void __launch_bounds__(256) k_absmax(const double* __restrict__ v, int64_t count, uint32_t* out)
{
    double m = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
        m = fmax(m, fabs(v[i]));
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
    if ((threadIdx.x & 63) == 0) atomicMax(out, __float_as_uint((float)m));
}

__global__ void __launch_bounds__(256) k_sum_parts(const double2* __restrict__ partial, int parts, int64_t count,
                                                   double2* __restrict__ dst, int add)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x) {
        double2 v = add ? dst[i] : make_double2(0.0, 0.0);
        for (int p = 0; p < parts; ++p) {
            const double2 x = partial[(int64_t)p * count + i];
            v.x += x.x;
            v.y += x.y;
        }
        dst[i] = v;
    }
}

hipError_t launch_sum_parts(const double2* partial, int parts, int64_t count, double2* dst, bool add, hipStream_t st)
{
    const int64_t blocks = std::min<int64_t>((count + 255) / 256, 8192);
    if (blocks > 0) hipLaunchKernelGGL(k_sum_parts, dim3((unsigned)blocks), dim3(256), 0, st, partial, parts, count, dst, add ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_absmax(const double* v, int64_t count, uint32_t* out, hipStream_t st)
{
    const int64_t blocks = std::min<int64_t>((count + 255) / 256, 2048);
    if (blocks > 0) hipLaunchKernelGGL(k_absmax, dim3((unsigned)blocks), dim3(256), 0, st, v, count, out);
    return hipGetLastError();
}

}  // namespace lfmm
