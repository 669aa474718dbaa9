// p2p.hip -- near-field pair sums (P2P), several targets per lane.
//
// The reference's near field (src/cpu/multipole3d.cpp:964-1151, GPU
// src/cuda/multipole3d.cu:568-576; pair term src/common/simple_math.h:130-139):
// phi_i += q_j / max(r, 1e-4), E_i += q_j (x_i - x_j) / max(r, 1e-4)^3 over
// the 27 neighbour leaves, self pair included.
//
// One lane per target loads a 32-byte source record for every pair it adds.
// Here each lane owns kTPL (default 2) consecutive targets of one leaf (they
// share the neighbour list) and every source record it loads feeds kTPL pair
// terms; 3 and 4 measured no faster than 2 at 4M uniform points, depth 7,
// where the loop is then bound by its FP64 instructions.
// Past a leaf's end a lane repeats its last target (computed, not written).
// Lanes are listed per leaf: lane slot_begin[leaf] + k owns targets
// leaf_begin[leaf] + kTPL k ... + kTPL k + kTPL - 1.
#include "device_common.h"
#include "ops3d.h"
#include "tree.h"

namespace lfmm {

namespace {

// 1 / sqrt(r2): v_rsq_f64 (relative error ~5e-8) refined by one Newton step:
// relative error <= 4.3e-15 (measured over 4M arguments in [1e-9, 10]), 1e-4
// of the FMM's own error at any order, for 3 fewer FP64 operations per pair
// than a second step
__device__ __forceinline__ double rsqrt_nf(double r2)
{
    const double y = __builtin_amdgcn_rsq(r2);
    const double hr = 0.5 * r2;
    return y * fma(-hr * y, y, 1.5);
}

// Near-source ranges of a leaf: 27 per leaf (uniform tree), or the CSR U list
// of the adaptive tree (nbr_off).
__device__ inline void leaf_ranges(const int2* nbr, const uint32_t* nbr_off, int leaf, const int2*& nr, int& cnt)
{
    if (nbr_off) {
        const uint32_t b = nbr_off[leaf];
        nr = nbr + b;
        cnt = (int)(nbr_off[leaf + 1] - b);
    } else {
        nr = nbr + (int64_t)leaf * 27;
        cnt = 27;
    }
}

// cnt[leaf]: lanes of the leaf; total[leaf]: sources in its near ranges
template <int TPL>
__global__ void __launch_bounds__(256) k_pair_counts(const int32_t* __restrict__ leaf_begin, const int2* __restrict__ nbr,
                                                    const uint32_t* __restrict__ nbr_off, int nleaf,
                                                    uint32_t* __restrict__ cnt, int32_t* __restrict__ total)
{
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= nleaf) return;
    cnt[l] = (uint32_t)((leaf_begin[l + 1] - leaf_begin[l] + TPL - 1) / TPL);
    const int2* nr;
    int nc;
    leaf_ranges(nbr, nbr_off, l, nr, nc);
    int t = 0;
    for (int k = 0; k < nc; ++k) {
        const int2 r = nr[k];
        t += max(r.y - r.x, 0);
    }
    total[l] = t;
}

template <int TPL>
__global__ void __launch_bounds__(256) k_pair_lanes(const int32_t* __restrict__ leaf_of,
                                                   const int32_t* __restrict__ leaf_begin,
                                                   const uint32_t* __restrict__ slot_begin, int64_t n,
                                                   int32_t* __restrict__ lanes)
{
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int leaf = leaf_of[s];
    const int64_t k = s - leaf_begin[leaf];
    if (k % TPL == 0) lanes[slot_begin[leaf] + k / TPL] = (int32_t)s;
}

template <int TPL>
__global__ void __launch_bounds__(256) k_p2p_multi(LeafArgs a, const int32_t* __restrict__ lanes,
                                                  const uint32_t* __restrict__ slot_begin,
                                                  const uint32_t* __restrict__ cnt, const int32_t* __restrict__ near_total)
{
    const int64_t i = (int64_t)xcd_local_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const int64_t total = (int64_t)slot_begin[a.nleaf - 1] + cnt[a.nleaf - 1];
    if (i >= total) return;
    const int64_t s0 = lanes[i];
    const int leaf = a.leaf_of[s0];
    const int64_t e = a.leaf_begin[leaf + 1];
    // targets s0 .. s0 + TPL - 1 of the leaf; past its end the last one is
    // repeated (computed, not written)
    double4 t[TPL];
    double ph[TPL], fx[TPL], fy[TPL], fz[TPL];
#pragma unroll
    for (int u = 0; u < TPL; ++u) {
        const int64_t s = min(s0 + u, e - 1);
        t[u] = a.src[s];
        ph[u] = fx[u] = fy[u] = fz[u] = 0.0;
        if (a.far) {  // far field from k_l2p
            const double4 f = a.far[s];
            ph[u] = f.x;
            fx[u] = f.y;
            fy[u] = f.z;
            fz[u] = f.w;
        }
    }
    const int2* nr;
    int nrc;
    leaf_ranges(a.nbr, a.nbr_off, leaf, nr, nrc);
    const int kl = nrc - 1;  // every list holds the leaf itself: nrc >= 1
    const double sc = a.leaf_scale ? a.leaf_scale[leaf] : 1.0;
    const double r2min = 1.e-8 * (sc * sc);  // (1e-4 in the system's units)^2
    // the near ranges walked as one list of total[leaf] sources, the next
    // range's bounds loaded ahead: a wave runs as long as its longest leaf
    // total, not as the sum over the ranges of its longest leaf in each.  The
    // loop runs on a plain count (the range advance never leaves it), which
    // keeps the accumulators in place across iterations.
    int k = 0;
    int2 r = nr[0], rn = nr[min(1, kl)];
    int j = r.x;
    const int nsrc = near_total[leaf];
    for (int it = 0; it < nsrc; ++it) {
        while (j >= r.y && k < kl) {  // total counts these ranges: k stays <= kl
            ++k;
            r = rn;
            rn = nr[min(k + 1, kl)];
            j = r.x;
        }
        const double4 p = a.src[min(j++, (int)a.n - 1)];
#pragma unroll
        for (int u = 0; u < TPL; ++u) {
            const double dx = t[u].x - p.x, dy = t[u].y - p.y, dz = t[u].z - p.z;
            const double r2 = fmax(fma(dx, dx, fma(dy, dy, dz * dz)), r2min);
            const double ir = rsqrt_nf(r2);
            const double qi = p.w * ir;
            const double qi3 = qi * (ir * ir);
            ph[u] += qi;
            fx[u] = fma(dx, qi3, fx[u]);
            fy[u] = fma(dy, qi3, fy[u]);
            fz[u] = fma(dz, qi3, fz[u]);
        }
    }
    // the self pair (clamped at 1e-4) was included; its field term is exactly zero
    const double self = a.exclude_self ? rsqrt_nf(r2min) : 0.0;
    const double sc2 = sc * sc;
#pragma unroll
    for (int u = 0; u < TPL; ++u) {
        if (s0 + u >= e) break;
        const int64_t o = a.perm[s0 + u];
        if (o >= a.n_out) continue;
        if (a.pot) a.pot[o] = (ph[u] - t[u].w * self) * sc;
        if (a.field) {
            a.field[3 * o] = fx[u] * sc2;
            a.field[3 * o + 1] = fy[u] * sc2;
            a.field[3 * o + 2] = fz[u] * sc2;
        }
    }
}

inline int nblk(int64_t n, int t) { return (int)((n + t - 1) / t); }

}  // namespace

#ifndef LFMM_P2P_TPL
#define LFMM_P2P_TPL 2
#endif
constexpr int kTPL = LFMM_P2P_TPL;  // targets per lane

// upper bound on the lane count: sum over leaves of ceil(count / kTPL)
inline int64_t lane_bound(int64_t n, int64_t nleaf) { return (n + (kTPL - 1) * nleaf) / kTPL + 1; }

int64_t p2p_ws_elems(int64_t n, int64_t nleaf) { return lane_bound(n, nleaf) + 3 * nleaf; }

hipError_t launch_p2p(const LeafArgs& a, int32_t* ws, uint32_t* scan_ws, hipStream_t st)
{
    if (a.n <= 0 || a.nleaf <= 0) return hipSuccess;
    // ws: cnt [nleaf], slot_begin [nleaf], total [nleaf], lanes [lane_bound]
    uint32_t* cnt = (uint32_t*)ws;
    uint32_t* slot = cnt + a.nleaf;
    int32_t* total = (int32_t*)(slot + a.nleaf);
    int32_t* lanes = total + a.nleaf;
    hipLaunchKernelGGL((k_pair_counts<kTPL>), dim3(nblk(a.nleaf, 256)), dim3(256), 0, st, a.leaf_begin, a.nbr,
                       a.nbr_off, a.nleaf, cnt, total);
    hipError_t e = exclusive_scan(cnt, slot, a.nleaf, scan_ws, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_pair_lanes<kTPL>), dim3(nblk(a.n, 256)), dim3(256), 0, st, a.leaf_of, a.leaf_begin, slot,
                       a.n, lanes);
    const int64_t bound = lane_bound(a.n, a.nleaf);
    hipLaunchKernelGGL((k_p2p_multi<kTPL>), dim3(nblk(bound, 256)), dim3(256), 0, st, a, lanes, slot, cnt, total);
    return hipGetLastError();
}

}  // namespace lfmm
