// ops2d.hip -- FP64 operator kernels of the 2D (complex-series) FMM, gfx950.
//
//   k_p2m2d    P2M, thread per leaf                    (reference Multipole, src/cpu/multipole.cpp:22-36)
//   k_m2m2d    M2M, thread per (parent, coefficient)    (M2M :38-53, Upward :155-177)
//   k_m2l2d    L2L + M2L, thread per target, operators by scalar loads (M2L :66-84, L2L :86-97, Downward :179-231);
//              FP64, or FP32 per-direction sums added in FP64
//   k_leaf2d   L2P + P2P over the 3x3 neighbourhood, thread per particle,
//              results scattered to input order       (ComputePotentials / ComputeForces :233-449)
//
// Expansions are normalised per level (tables2d.cpp), so every operator is a
// fixed complex matrix per integer box offset or child class.
#include "device_common.h"
#include "lfmm_internal.h"
#include "ops2d.h"

#include <type_traits>

namespace lfmm {

__device__ inline double2 cmadd(double2 acc, double2 a, double2 b)
{
    acc.x = fma(a.x, b.x, fma(-a.y, b.y, acc.x));
    acc.y = fma(a.x, b.y, fma(a.y, b.x, acc.y));
    return acc;
}

__global__ void __launch_bounds__(256) k_p2m2d(P2M2dArgs a)
{
    const int leaf = blockIdx.x * blockDim.x + threadIdx.x;
    if (leaf >= a.nleaf) return;
    const uint64_t key = a.leaf_key[leaf];
    const double scale = a.geom[3];
    const double cx = box_centre(compact2(key >> 1), a.level, scale, a.geom[0]);
    const double cy = box_centre(compact2(key), a.level, scale, a.geom[1]);
    const double ih = ldexp(1.0, a.level) / scale;
    double2 acc[kMaxTerms2d];
#pragma unroll
    for (int k = 0; k < kMaxTerms2d; ++k) acc[k] = make_double2(0.0, 0.0);
    const int b = a.leaf_begin[leaf], e = a.leaf_begin[leaf + 1];
    for (int p = b; p < e; ++p) {
        const double4 s = a.src[p];
        const double ux = (s.x - cx) * ih, uy = (s.y - cy) * ih, q = s.w;
        double wx = q * ux, wy = q * uy;  // q u^k
        acc[0].x += q;
#pragma unroll
        for (int k = 1; k < kMaxTerms2d; ++k) {
            if (k < a.N) {
                acc[k].x += wx;
                acc[k].y += wy;
                const double t = wx * ux - wy * uy;
                wy = wx * uy + wy * ux;
                wx = t;
            }
        }
    }
    double2* out = a.outer + (int64_t)leaf * a.N;
    out[0] = acc[0];
#pragma unroll
    for (int k = 1; k < kMaxTerms2d; ++k)
        if (k < a.N) out[k] = make_double2(-acc[k].x / k, -acc[k].y / k);
}

__global__ void __launch_bounds__(256) k_m2m2d(M2M2dArgs a, const double2* __restrict__ mm)
{
    const int64_t gid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int N = a.N;
    if (gid >= (int64_t)a.num_parents * N) return;
    const int P = (int)(gid / N), j = (int)(gid % N);
    double2 acc = make_double2(0.0, 0.0);
    for (int ch = a.child_begin[P]; ch < a.child_begin[P + 1]; ++ch) {
        const int c = (int)(a.child_keys[ch] & 3u);
        const double2* t = mm + ((int64_t)c * N + j) * N;
        const double2* x = a.src + (int64_t)ch * N;
        for (int s = 0; s <= j; ++s) acc = cmadd(acc, t[s], x[s]);
    }
    a.dst[(int64_t)P * N + j] = acc;
}

// L2L + M2L into the class-c children of 256 consecutive parents, one lane
// per target, one workgroup per (class, parent tile, block of LB output
// coefficients).  All lanes of a workgroup share the class and the block, so
// the operator rows they need (LB rows of the class-c L2L and of the M2L of
// each of the class's 27 directions) are staged in LDS once and read as
// broadcasts, while each lane keeps its target's LB accumulators and one
// source expansion in registers.  The blocks multiply the workgroups by about
// 4, which the deepest level (one target per lane, 1,024 waves at 65,536
// targets) needs to hide the source loads.  Per (target, coefficient) the
// products are summed in the order of the reference's Downward (:179-231):
// L2L, then per direction the level term (coefficient 0) and the source
// coefficients in increasing order.
// F32: each direction's sum over the source coefficients is formed in FP32
// (FP32 operator, sources rounded on load) and added to the FP64 accumulator;
// the level term and the L2L stay FP64.
constexpr int m2l2d_lb(int N) { return (N + 3) / 4; }

template <int N, bool F32>
__global__ void __launch_bounds__(256) k_m2l2d(M2L2dArgs a, Tables2d T)
{
    constexpr int LB = m2l2d_lb(N), NB = (N + LB - 1) / LB, ROWS = LB * N;
    using op_t = typename std::conditional<F32, float2, double2>::type;
    __shared__ op_t mop[kDirsPerClass2d * ROWS];
    __shared__ double2 lop[ROWS];
    const int c = blockIdx.x & 3;
    const int rest = (int)(blockIdx.x >> 2);
    const int l0 = (rest % NB) * LB;
    for (int i = threadIdx.x; i < kDirsPerClass2d * ROWS; i += 256) {
        const int k = i / ROWS, r = (i % ROWS) / N, j = i % N;
        const int d = T.class_dirs[c * kDirsPerClass2d + k];
        const int l = l0 + r;
        if constexpr (F32) mop[i] = l < N ? T.m2l32[((int64_t)d * N + l) * N + j] : make_float2(0.f, 0.f);
        else mop[i] = l < N ? T.m2l[((int64_t)d * N + l) * N + j] : make_double2(0.0, 0.0);
    }
    for (int i = threadIdx.x; i < ROWS; i += 256) {
        const int l = l0 + i / N, j = i % N;
        lop[i] = l < N ? T.l2l[((int64_t)c * N + l) * N + j] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    const int P = (rest / NB) * 256 + (int)threadIdx.x;
    int t = -1;
    if (P < a.num_parents)
        for (int i = a.child_begin[P]; i < a.child_begin[P + 1]; ++i)
            if ((int)(a.child_keys[i] & 3u) == c) t = i;
    if (t < 0) return;  // no barriers below: lanes may leave
    const uint64_t key = a.child_keys[t];
    const int tx = (int)compact2(key >> 1), ty = (int)compact2(key);
    const int side = 1 << a.level;
    double2 acc[LB];
#pragma unroll
    for (int r = 0; r < LB; ++r) acc[r] = make_double2(0.0, 0.0);
    if (a.parent_local) {  // L2L from the parent
        const double2* xp = a.parent_local + (int64_t)P * N;
        double2 x[N];
#pragma unroll
        for (int j = 0; j < N; ++j) x[j] = xp[j];
#pragma unroll
        for (int r = 0; r < LB; ++r)
#pragma unroll
            for (int j = 0; j < N; ++j)
                if (j >= l0 + r) acc[r] = cmadd(acc[r], lop[r * N + j], x[j]);
    }
    for (int k = 0; k < kDirsPerClass2d; ++k) {
        const int d = T.class_dirs[c * kDirsPerClass2d + k];
        const int sx = tx + T.dir_vec[2 * d], sy = ty + T.dir_vec[2 * d + 1];
        if (sx < 0 || sy < 0 || sx >= side || sy >= side) continue;
        const int s = find_box(morton2(sx, sy), a.lut, a.keys, a.size);
        if (s < 0) continue;
        const double2* xs = a.src + (int64_t)s * N;
        double2 x[N];
#pragma unroll
        for (int j = 0; j < N; ++j) x[j] = xs[j];
        if (l0 == 0) {  // level term at_0 log h of the normalised M2L
            acc[0].x = fma(x[0].x, a.log_h, acc[0].x);
            acc[0].y = fma(x[0].y, a.log_h, acc[0].y);
        }
        const op_t* m = mop + k * ROWS;
        if constexpr (F32) {
            float xr[N], xi[N];
#pragma unroll
            for (int j = 0; j < N; ++j) {
                xr[j] = (float)x[j].x;
                xi[j] = (float)x[j].y;
            }
#pragma unroll
            for (int r = 0; r < LB; ++r) {
                float ax = 0.f, ay = 0.f;
#pragma unroll
                for (int j = 0; j < N; ++j) {
                    const float2 mv = m[r * N + j];
                    ax = fmaf(mv.x, xr[j], fmaf(-mv.y, xi[j], ax));
                    ay = fmaf(mv.x, xi[j], fmaf(mv.y, xr[j], ay));
                }
                acc[r].x += (double)ax;
                acc[r].y += (double)ay;
            }
        } else {
#pragma unroll
            for (int r = 0; r < LB; ++r)
#pragma unroll
                for (int j = 0; j < N; ++j) acc[r] = cmadd(acc[r], m[r * N + j], x[j]);
        }
    }
    double2* out = a.dst + (int64_t)t * N;
#pragma unroll
    for (int r = 0; r < LB; ++r)
        if (l0 + r < N) out[l0 + r] = acc[r];
}

// Pair term of the reference 2D kernels (src/common/simple_math.h:119-128):
// q log(max(|dz|, 1e-4)) and q conj(dz) / max(|dz|^2, 1e-8), self pair included.
template <bool LOCAL>
__global__ void __launch_bounds__(256) k_leaf2d(Leaf2dArgs a)
{
    const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (s >= a.n) return;
    const int64_t o = a.perm[s];
    if (o >= a.n_out) return;
    const int leaf = a.leaf_of[s];
    const uint64_t key = a.leaf_key[leaf];
    const uint32_t ix = compact2(key >> 1), iy = compact2(key);
    const double4 t = a.src[s];
    double ph = 0.0, fx = 0.0, fy = 0.0;
    if constexpr (LOCAL) {  // L2P by Horner: P(u) = sum bt_k u^k, P'(u)
        const double scale = a.geom[3];
        const double ih = ldexp(1.0, a.level) / scale;
        const double ux = (t.x - box_centre(ix, a.level, scale, a.geom[0])) * ih;
        const double uy = (t.y - box_centre(iy, a.level, scale, a.geom[1])) * ih;
        const double2* b = a.local + (int64_t)leaf * a.N;
        double2 p = b[a.N - 1], d = make_double2(0.0, 0.0);
        for (int k = a.N - 2; k >= 0; --k) {
            d = make_double2(d.x * ux - d.y * uy + p.x, d.x * uy + d.y * ux + p.y);
            const double2 bk = b[k];
            p = make_double2(p.x * ux - p.y * uy + bk.x, p.x * uy + p.y * ux + bk.y);
        }
        ph = p.x;
        fx = d.x * ih;
        fy = d.y * ih;
    }
    const int side = 1 << a.level;
    for (int nb9 = 0; nb9 < 9; ++nb9) {
        const int nx = (int)ix + nb9 / 3 - 1, ny = (int)iy + nb9 % 3 - 1;
        if (nx < 0 || ny < 0 || nx >= side || ny >= side) continue;
        const int nb = (nb9 == 4) ? leaf : find_box(morton2(nx, ny), a.lut, a.leaf_key, a.nleaf);
        if (nb < 0) continue;
        for (int j = a.leaf_begin[nb]; j < a.leaf_begin[nb + 1]; ++j) {
            if (a.exclude_self && j == (int)s) continue;
            const double4 p = a.src[j];
            const double dx = t.x - p.x, dy = t.y - p.y;
            const double nz = fmax(dx * dx + dy * dy, 1.e-8);
            const double qi = p.w / nz;
            ph += p.w * 0.5 * log(nz);
            fx = fma(qi, dx, fx);
            fy = fma(-qi, dy, fy);
        }
    }
    if (a.pot) a.pot[o] = ph;
    if (a.field) {
        a.field[2 * o] = fx;
        a.field[2 * o + 1] = fy;
    }
}

static inline unsigned nblk(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

hipError_t launch_p2m2d(const P2M2dArgs& a, hipStream_t st)
{
    if (a.N < 1 || a.N > kMaxTerms2d) return hipErrorInvalidValue;
    if (a.nleaf > 0) hipLaunchKernelGGL(k_p2m2d, dim3(nblk(a.nleaf, 256)), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_m2m2d(const M2M2dArgs& a, const Tables2d& T, hipStream_t st)
{
    const int64_t n = (int64_t)a.num_parents * a.N;
    if (n > 0) hipLaunchKernelGGL(k_m2m2d, dim3(nblk(n, 256)), dim3(256), 0, st, a, T.m2m);
    return hipGetLastError();
}

template <int N>
hipError_t m2l2d_n(const M2L2dArgs& a, const Tables2d& T, hipStream_t st)
{
    constexpr int NB = (N + m2l2d_lb(N) - 1) / m2l2d_lb(N);
    const dim3 grid((unsigned)(4 * NB * ((a.num_parents + 255) / 256)));
    if (a.f32) hipLaunchKernelGGL((k_m2l2d<N, true>), grid, dim3(256), 0, st, a, T);
    else hipLaunchKernelGGL((k_m2l2d<N, false>), grid, dim3(256), 0, st, a, T);
    return hipGetLastError();
}

hipError_t launch_m2l2d(const M2L2dArgs& a, const Tables2d& T, hipStream_t st)
{
    if (a.num_parents <= 0) return hipSuccess;
    switch (a.N) {
#define X(n) case n: return m2l2d_n<n>(a, T, st);
        X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15)
        X(16) X(17) X(18) X(19) X(20) X(21) X(22) X(23) X(24) X(25) X(26) X(27) X(28) X(29) X(30)
#undef X
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_leaf2d(const Leaf2dArgs& a, hipStream_t st)
{
    if (a.n <= 0) return hipSuccess;
    if (a.local) hipLaunchKernelGGL(k_leaf2d<true>, dim3(nblk(a.n, 256)), dim3(256), 0, st, a);
    else hipLaunchKernelGGL(k_leaf2d<false>, dim3(nblk(a.n, 256)), dim3(256), 0, st, a);
    return hipGetLastError();
}

}  // namespace lfmm
