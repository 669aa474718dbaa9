// direct.hip -- FP64 all-pairs sums on the GPU (the reference's
// fmm::gpu::ComputeExact, src/common/simple_math.cu:9-22 and :24-130, with the
// pair terms of src/common/simple_math.h:120-151 and their device forms
// :197-208).
//
// One thread per target; sources stream through LDS in tiles of 256 packed
// (x, y, z, q) records, each added in input order.  3D pair:
//   q / max(|r|, 1e-4)  and  q r / max(|r|, 1e-4)^3   (r = x_i - x_j)
// evaluated as q * rsqrt(max(|r|^2, 1e-8)), the reference's device form.
// 2D pair: q * 0.5 log max(|dz|^2, 1e-8) and q conj(dz) / max(|dz|^2, 1e-8).
// With `exclude` the pair j == i is skipped (single-set calls only).
#include "device_common.h"
#include "direct.h"

#include <algorithm>
#include <mutex>

namespace lfmm {

namespace {

constexpr int kTile = 256;

// 1 / sqrt(r2): v_rsq_f64 refined by two Newton steps (full FP64 accuracy)
__device__ __forceinline__ double rsqrt_full(double r2)
{
    double y = __builtin_amdgcn_rsq(r2);
    const double hr = 0.5 * r2;
    y = y * fma(-hr * y, y, 1.5);
    y = y * fma(-hr * y, y, 1.5);
    return y;
}

__global__ void __launch_bounds__(kTile) k_direct3d(const double4* __restrict__ tgt, int64_t nt,
                                                   const double4* __restrict__ src, int64_t ns, int exclude,
                                                   double* __restrict__ pot, double* __restrict__ field)
{
    __shared__ double4 tile[kTile];
    const int64_t i = (int64_t)blockIdx.x * kTile + threadIdx.x;
    const double4 t = i < nt ? tgt[i] : make_double4(0.0, 0.0, 0.0, 0.0);
    double p = 0.0, fx = 0.0, fy = 0.0, fz = 0.0;
    for (int64_t base = 0; base < ns; base += kTile) {
        __syncthreads();
        const int64_t j0 = base + threadIdx.x;
        tile[threadIdx.x] = j0 < ns ? src[j0] : make_double4(0.0, 0.0, 0.0, 0.0);
        __syncthreads();
        const int cnt = (int)min<int64_t>(kTile, ns - base);
        for (int j = 0; j < cnt; ++j) {
            if (exclude && base + j == i) continue;
            const double4 s = tile[j];
            const double dx = t.x - s.x, dy = t.y - s.y, dz = t.z - s.z;
            const double r2 = fmax(fma(dx, dx, fma(dy, dy, dz * dz)), 1.e-8);
            const double ir = rsqrt_full(r2);
            const double qi = s.w * ir;
            p += qi;
            const double q3 = qi * ir * ir;
            fx = fma(q3, dx, fx);
            fy = fma(q3, dy, fy);
            fz = fma(q3, dz, fz);
        }
    }
    if (i < nt) {
        if (pot) pot[i] = p;
        if (field) {
            field[3 * i] = fx;
            field[3 * i + 1] = fy;
            field[3 * i + 2] = fz;
        }
    }
}

__global__ void __launch_bounds__(kTile) k_direct2d(const double4* __restrict__ tgt, int64_t nt,
                                                   const double4* __restrict__ src, int64_t ns, int exclude,
                                                   double* __restrict__ pot, double* __restrict__ field)
{
    __shared__ double4 tile[kTile];
    const int64_t i = (int64_t)blockIdx.x * kTile + threadIdx.x;
    const double4 t = i < nt ? tgt[i] : make_double4(0.0, 0.0, 0.0, 0.0);
    double p = 0.0, fx = 0.0, fy = 0.0;
    for (int64_t base = 0; base < ns; base += kTile) {
        __syncthreads();
        const int64_t j0 = base + threadIdx.x;
        tile[threadIdx.x] = j0 < ns ? src[j0] : make_double4(0.0, 0.0, 0.0, 0.0);
        __syncthreads();
        const int cnt = (int)min<int64_t>(kTile, ns - base);
        for (int j = 0; j < cnt; ++j) {
            if (exclude && base + j == i) continue;
            const double4 s = tile[j];
            // packed 2D records are (x, y, 0, q)
            const double dx = t.x - s.x, dy = t.y - s.y;
            const double r2 = fmax(fma(dx, dx, dy * dy), 1.e-8);
            p = fma(s.w * 0.5, log(r2), p);
            const double qr = s.w / r2;
            fx = fma(qr, dx, fx);
            fy = fma(-qr, dy, fy);
        }
    }
    if (i < nt) {
        if (pot) pot[i] = p;
        if (field) {
            field[2 * i] = fx;
            field[2 * i + 1] = fy;
        }
    }
}

// Few targets, many sources: block (x, y) sums targets x * kTile .. over
// source chunk y into part[y * nt + i] = (pot, field); k_direct3d_join adds
// the chunks in order.  Spreads a sampled check (hundreds of targets against
// millions of sources) over the whole GPU.
__global__ void __launch_bounds__(kTile) k_direct3d_part(const double4* __restrict__ tgt, int64_t nt,
                                                        const double4* __restrict__ src, int64_t ns, int64_t chunk,
                                                        double4* __restrict__ part)
{
    __shared__ double4 tile[kTile];
    const int64_t i = (int64_t)blockIdx.x * kTile + threadIdx.x;
    const double4 t = i < nt ? tgt[i] : make_double4(0.0, 0.0, 0.0, 0.0);
    const int64_t j_beg = (int64_t)blockIdx.y * chunk, j_end = min(ns, j_beg + chunk);
    double p = 0.0, fx = 0.0, fy = 0.0, fz = 0.0;
    for (int64_t base = j_beg; base < j_end; base += kTile) {
        __syncthreads();
        const int64_t j0 = base + threadIdx.x;
        tile[threadIdx.x] = j0 < j_end ? src[j0] : make_double4(0.0, 0.0, 0.0, 0.0);
        __syncthreads();
        const int cnt = (int)min<int64_t>(kTile, j_end - base);
        for (int j = 0; j < cnt; ++j) {
            const double4 s = tile[j];
            const double dx = t.x - s.x, dy = t.y - s.y, dz = t.z - s.z;
            const double r2 = fmax(fma(dx, dx, fma(dy, dy, dz * dz)), 1.e-8);
            const double ir = rsqrt_full(r2);
            const double qi = s.w * ir;
            p += qi;
            const double q3 = qi * ir * ir;
            fx = fma(q3, dx, fx);
            fy = fma(q3, dy, fy);
            fz = fma(q3, dz, fz);
        }
    }
    if (i < nt) part[(int64_t)blockIdx.y * nt + i] = make_double4(p, fx, fy, fz);
}

__global__ void __launch_bounds__(256) k_direct3d_join(const double4* __restrict__ part, int64_t nt, int splits,
                                                      double* __restrict__ pot, double* __restrict__ field)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nt) return;
    double p = 0.0, fx = 0.0, fy = 0.0, fz = 0.0;
    for (int y = 0; y < splits; ++y) {
        const double4 v = part[(int64_t)y * nt + i];
        p += v.x;
        fx += v.y;
        fy += v.z;
        fz += v.w;
    }
    if (pot) pot[i] = p;
    if (field) {
        field[3 * i] = fx;
        field[3 * i + 1] = fy;
        field[3 * i + 2] = fz;
    }
}

// Field of vector charges from the fields of their components: with
// E_c = sum_j q_j^c r / |r|^3, sum_j q_j x r / |r|^3 = (E_2.z - E_3.y,
// E_3.x - E_1.z, E_1.y - E_2.x) (the reference's forces_cross,
// src/cuda/multipole3d.cu:765-776).  Written over e3.
__global__ void __launch_bounds__(256) k_cross_fields(const double* __restrict__ e1, const double* __restrict__ e2,
                                                     double* __restrict__ e3, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double a0 = e1[3 * i], a1 = e1[3 * i + 1], a2 = e1[3 * i + 2];
    const double b0 = e2[3 * i], b1 = e2[3 * i + 1], b2 = e2[3 * i + 2];
    const double c0 = e3[3 * i], c1 = e3[3 * i + 1], c2 = e3[3 * i + 2];
    (void)a0;
    (void)b1;
    (void)c2;
    e3[3 * i] = b2 - c1;
    e3[3 * i + 1] = c0 - a2;
    e3[3 * i + 2] = a1 - b0;
}

}  // namespace

hipError_t launch_cross_fields(const double* e1, const double* e2, double* e3, int64_t n, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_cross_fields, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, e1, e2, e3, n);
    return hipGetLastError();
}

hipError_t launch_direct(int dim, const double4* tgt, int64_t nt, const double4* src, int64_t ns, bool exclude_self,
                         double* pot, double* field, hipStream_t st)
{
    if (nt <= 0) return hipSuccess;
    const dim3 grid((unsigned)((nt + kTile - 1) / kTile));
    if (dim == 3)
        hipLaunchKernelGGL(k_direct3d, grid, dim3(kTile), 0, st, tgt, nt, src, ns, exclude_self ? 1 : 0, pot, field);
    else
        hipLaunchKernelGGL(k_direct2d, grid, dim3(kTile), 0, st, tgt, nt, src, ns, exclude_self ? 1 : 0, pot, field);
    return hipGetLastError();
}

int direct3d_split_count(int64_t nt, int64_t ns)
{
    if (nt <= 0 || ns <= 0) return 1;
    const int64_t tblocks = (nt + kTile - 1) / kTile;
    const int64_t want = (2048 + tblocks - 1) / tblocks;        // about 8 workgroups per CU
    const int64_t most = (ns + 4 * kTile - 1) / (4 * kTile);    // at least 4 tiles per chunk
    return (int)std::max<int64_t>(1, std::min<int64_t>({want, most, 65535}));
}

hipError_t launch_direct3d_split(const double4* tgt, int64_t nt, const double4* src, int64_t ns, double* pot,
                                 double* field, double4* part, int splits, hipStream_t st)
{
    if (nt <= 0) return hipSuccess;
    const int64_t chunk = ((ns + splits - 1) / splits + kTile - 1) / kTile * kTile;
    const dim3 grid((unsigned)((nt + kTile - 1) / kTile), (unsigned)splits);
    hipLaunchKernelGGL(k_direct3d_part, grid, dim3(kTile), 0, st, tgt, nt, src, ns, chunk, part);
    hipLaunchKernelGGL(k_direct3d_join, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, st, part, nt, splits, pot,
                       field);
    return hipGetLastError();
}

// ---- the allocation-free direct sum (lfmm_direct3d_noalloc)
// The packed (x, y, z, q) records of one call's sources and targets live in
// this fixed array, one copy per GPU, held until the process ends.  Host code
// never passes the array's name to a kernel (in host code it is a stand-in,
// not a device address): each device's address comes from
// hipGetSymbolAddress once, with that device current (noalloc_scratch_init,
// from lfmm_create), and is used from then on.
constexpr int64_t kNoAllocRecords = (int64_t)(kNoAllocScratchBytes / sizeof(double4));
static __device__ __align__(256) double4 scratch[kNoAllocRecords];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double4* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

hipError_t noalloc_scratch_init(int device)
{
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (g_scratch_ptr[device]) return hipSuccess;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = (double4*)p;
    return hipSuccess;
}

double4* noalloc_scratch(int device)
{
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> g(g_scratch_mu);
    return g_scratch_ptr[device];
}

}  // namespace lfmm
