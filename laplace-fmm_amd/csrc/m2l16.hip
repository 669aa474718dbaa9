// m2l16.hip -- level-wide helpers of the matrix-core M2L.
#include "ops3d.h"

#include <algorithm>

namespace lfmm {

__global__ void __launch_bounds__(256) k_sum_parts(const double2* __restrict__ partial, int parts, int64_t count,
                                                   double2* __restrict__ dst, int add)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x) {
        double2 v = add ? dst[i] : make_double2(0.0, 0.0);
        // the loads of a group of parts are issued together (independent
        // addresses), then added in part order
        int p = 0;
        for (; p + 8 <= parts; p += 8) {
            double2 x[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) x[j] = partial[(int64_t)(p + j) * count + i];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                v.x += x[j].x;
                v.y += x[j].y;
            }
        }
        for (; p < parts; ++p) {
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

}  // namespace lfmm
