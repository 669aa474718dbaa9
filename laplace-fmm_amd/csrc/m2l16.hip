// m2l16.hip -- level-wide helpers of the matrix-core M2L.
#include "lfmm_internal.h"
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

namespace {
struct ListClasses {
    int coff[9];
};
}  // namespace

// Adds the direction parts of a class-sorted list into the locals: list
// entry i is the class-c child of parent plist[i] (coff[c] <= i <
// coff[c + 1]); its part p is at partial + (p * count + i) * nx doubles, its
// local at dst + child * stride doubles.  Parts are added in order.
__global__ void __launch_bounds__(256) k_sum_list_parts(const double* __restrict__ partial, int parts, int count,
                                                        int nx, ListClasses cls, const int32_t* __restrict__ plist,
                                                        const int32_t* __restrict__ child_begin,
                                                        const uint64_t* __restrict__ child_keys, int64_t stride,
                                                        double* __restrict__ dst)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (int64_t)count * nx) return;
    const int i = (int)(g / nx), j = (int)(g - (int64_t)i * nx);
    int c = 0;
    while (c < 7 && i >= cls.coff[c + 1]) ++c;
    const int P = plist[i];
    int t = -1;
    for (int k = child_begin[P]; k < child_begin[P + 1]; ++k)
        if ((int)(child_keys[k] & 7u) == c) t = k;
    if (t < 0) return;
    double v = dst[(int64_t)t * stride + j];
    for (int p = 0; p < parts; ++p) v += partial[((int64_t)p * count + i) * nx + j];
    dst[(int64_t)t * stride + j] = v;
}

hipError_t launch_sum_list_parts(int N, const M2L16Args& a, hipStream_t st)
{
    if (!a.csorted || a.dsplit <= 1) return hipSuccess;
    ListClasses cls;
    for (int c = 0; c < 9; ++c) cls.coff[c] = a.coff[c];
    const int nx = 2 * tri(N);
    const int64_t stride = 2 * (int64_t)(a.stride > 0 ? a.stride : tri(N));
    const int64_t total = (int64_t)a.plist_count * nx;
    if (total > 0)
        hipLaunchKernelGGL(k_sum_list_parts, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st,
                           (const double*)a.partial, a.dsplit, a.plist_count, nx, cls, a.plist, a.child_begin,
                           a.child_keys, stride, (double*)a.dst);
    return hipGetLastError();
}

}  // namespace lfmm
