// device_common.h -- device helpers shared by the tree and operator kernels
// (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace lfmm {

constexpr int kWave = 64;

// ---- Morton interleave: x is the high bit of each triple, as in the
// reference's MortonCode (src/cpu/morton_tree.cpp:93-99). 21 bits per axis.
__host__ __device__ inline uint64_t spread3(uint64_t v)
{
    v &= 0x1fffffull;
    v = (v | (v << 32)) & 0x1f00000000ffffull;
    v = (v | (v << 16)) & 0x1f0000ff0000ffull;
    v = (v | (v << 8)) & 0x100f00f00f00f00full;
    v = (v | (v << 4)) & 0x10c30c30c30c30c3ull;
    v = (v | (v << 2)) & 0x1249249249249249ull;
    return v;
}

__host__ __device__ inline uint32_t compact3(uint64_t v)
{
    v &= 0x1249249249249249ull;
    v = (v ^ (v >> 2)) & 0x10c30c30c30c30c3ull;
    v = (v ^ (v >> 4)) & 0x100f00f00f00f00full;
    v = (v ^ (v >> 8)) & 0x1f0000ff0000ffull;
    v = (v ^ (v >> 16)) & 0x1f00000000ffffull;
    v = (v ^ (v >> 32)) & 0x1fffffull;
    return (uint32_t)v;
}

__host__ __device__ inline uint64_t morton3(uint32_t ix, uint32_t iy, uint32_t iz)
{
    return (spread3(ix) << 2) | (spread3(iy) << 1) | spread3(iz);
}

// 2D interleave: x is the high bit of each pair (x << 1 | y), as in the
// reference's MortonCode<2> (src/cpu/morton_tree.cpp:84-91). 31 bits per axis.
__host__ __device__ inline uint64_t spread2(uint64_t v)
{
    v &= 0x7fffffffull;
    v = (v | (v << 16)) & 0x0000ffff0000ffffull;
    v = (v | (v << 8)) & 0x00ff00ff00ff00ffull;
    v = (v | (v << 4)) & 0x0f0f0f0f0f0f0f0full;
    v = (v | (v << 2)) & 0x3333333333333333ull;
    v = (v | (v << 1)) & 0x5555555555555555ull;
    return v;
}

__host__ __device__ inline uint32_t compact2(uint64_t v)
{
    v &= 0x5555555555555555ull;
    v = (v ^ (v >> 1)) & 0x3333333333333333ull;
    v = (v ^ (v >> 2)) & 0x0f0f0f0f0f0f0f0full;
    v = (v ^ (v >> 4)) & 0x00ff00ff00ff00ffull;
    v = (v ^ (v >> 8)) & 0x0000ffff0000ffffull;
    v = (v ^ (v >> 16)) & 0x00000000ffffffffull;
    return (uint32_t)v;
}

__host__ __device__ inline uint64_t morton2(uint32_t ix, uint32_t iy) { return (spread2(ix) << 1) | spread2(iy); }

// Box centre exactly as the reference's PointFromMortonCode
// (src/cpu/morton_tree.cpp:122-145): ((2i+1)/2^(l+1)) * scale + shift with
// the product and the sum rounded separately.
__device__ inline double box_centre(uint32_t i, int level, double scale, double shift)
{
    const double frac = (double)(2u * i + 1u) * __longlong_as_double((int64_t)(1023 - (level + 1)) << 52);
    return __dadd_rn(__dmul_rn(frac, scale), shift);
}

// ---- DPP row (16-lane) reductions -------------------------------------
template <int CTRL>
__device__ inline double dpp_d(double v)
{
    const int lo = __double2loint(v), hi = __double2hiint(v);
    const int l2 = __builtin_amdgcn_update_dpp(0, lo, CTRL, 0xf, 0xf, false);
    const int h2 = __builtin_amdgcn_update_dpp(0, hi, CTRL, 0xf, 0xf, false);
    return __hiloint2double(h2, l2);
}

// Sum over the 16 lanes of a DPP row; every lane of the row gets the sum.
__device__ inline double row16_sum(double v)
{
    v += dpp_d<0xB1>(v);   // quad_perm [1,0,3,2]
    v += dpp_d<0x4E>(v);   // quad_perm [2,3,0,1]
    v += dpp_d<0x141>(v);  // row_half_mirror
    v += dpp_d<0x140>(v);  // row_mirror
    return v;
}

__device__ inline double2 cmul(double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ inline double2 cmulc(double2 a, double2 b)  // a * conj(b)
{
    return make_double2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}

// *out = max(*out, m) over the wave, at most one atomic per wave (m >= 0; non-negative
// floats order like their bit patterns).  Call from every lane of the wave.
__device__ inline void wave_absmax(double m, uint32_t* out)
{
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
    if ((threadIdx.x & 63) == 0) {
        // the word only grows: skip the atomic when it already holds at least
        // this value (a stale read only lets a needless atomic through), so a
        // level's waves do not all serialise on one address
        const uint32_t b = __float_as_uint((float)m);
        if (b > __atomic_load_n(out, __ATOMIC_RELAXED)) atomicMax(out, b);
    }
}

// *out = max(*out, m) over the workgroup (at most 1024 threads), at most one
// atomic per workgroup; skipped when the word already holds at least m (it
// only grows).  Call from every thread of the workgroup: a single word
// takes every atomic, so one per wave serialises thousands of them.
__device__ inline void block_max_u32(uint32_t m, uint32_t* out)
{
    __shared__ uint32_t wm[16];
    for (int off = 32; off > 0; off >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, off));
    if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < (int)((blockDim.x + 63) >> 6); ++w) m = max(m, wm[w]);
        if (m > __atomic_load_n(out, __ATOMIC_RELAXED)) atomicMax(out, m);
    }
}

// Find box index by key at a level: dense table when present, otherwise a
// binary search over the sorted unique keys.  Returns -1 when absent.
__device__ inline int find_box(uint64_t key, const int* __restrict__ lut, const uint64_t* __restrict__ keys, int size)
{
    if (lut) return lut[key];
    int lo = 0, hi = size;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        const uint64_t k = keys[mid];
        if (k < key) lo = mid + 1;
        else hi = mid;
    }
    return (lo < size && keys[lo] == key) ? lo : -1;
}

// Blocks are dealt to the 8 XCDs round-robin; renumber them so that each XCD
// walks a contiguous range of work items (a compact region of the tree in
// Morton order), whose shared inputs then stay in that XCD's L2.
__device__ __forceinline__ int xcd_local_block(int b, int nb)
{
    const int per = nb >> 3, rem = nb & 7;
    const int x = b & 7, slot = b >> 3;
    return (x < rem) ? x * (per + 1) + slot : rem * (per + 1) + (x - rem) * per + slot;
}

}  // namespace lfmm
