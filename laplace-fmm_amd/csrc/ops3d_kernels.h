// ops3d_kernels.h -- FP64 FMM operator kernels for gfx950 (CDNA4, wave64).
// Compiled once per expansion size by ops3d_inst.hip (-DLFMM_N=<terms>).
//
//   k_p2m   P2M, wave per leaf           (reference Multipole, src/cpu/multipole3d.cpp:20-49)
//   k_rtr   M2M and M2L(+L2L)            (reference M2M :115-150, M2L :152-190, L2L :192-226)
//   k_l2p   L2P, lane per target particle, far field in sorted order
//           (reference ComputeLeaves<double> :1156-1210); the pair sums and the
//           scatter to input order are in p2p.hip
//
// Coefficients are normalised per level (lfmm_internal.h), so every operator
// table depends only on an integer box offset.
//
// k_rtr is the rotate / axial-translate / rotate-back operator.  A workgroup
// of 4 waves owns 64 target boxes (one per lane, the same lane in every
// wave) and walks the interaction directions in lock step, so all rotation
// and translation coefficients of a step are wave-uniform (scalar loads).
// Per direction the op state (one expansion per lane) lives in LDS as
// [coefficient][lane] and is produced / consumed in three phases that are
// partitioned over the 4 waves in three different ways:
//   A  forward rotation  Rz(phi), Ry(theta)        by source degree n
//   B  axial translation                            by order m
//   C  inverse rotation Ry(-theta), Rz(-phi), accumulate in registers  by target degree j
// Each wave keeps only the accumulators of its own target degrees.
#pragma once
#include "device_common.h"
#include "lfmm_internal.h"
#include "ops3d.h"

#include <algorithm>
#include <type_traits>
#include <utility>

namespace lfmm {

template <int B, int E, typename F>
__device__ __forceinline__ void sfor(F&& f)
{
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        sfor<B + 1, E>(f);
    }
}

// Work partition of degrees (phases A, C: cost (n+1)^2) and orders (phase B:
// cost (N-m)^2) over NW waves, longest-processing-time first.
template <int N, int NW>
struct Partition {
    int deg[kMaxTerms3d + 1];
    int ord[kMaxTerms3d + 1];
    constexpr Partition() : deg(), ord()
    {
        long la[NW] = {}, lb[NW] = {};
        for (int n = N - 1; n >= 0; --n) {
            int b = 0;
            for (int w = 1; w < NW; ++w)
                if (la[w] < la[b]) b = w;
            deg[n] = b;
            la[b] += (long)(n + 1) * (n + 1);
        }
        for (int m = 0; m < N; ++m) {
            int b = 0;
            for (int w = 1; w < NW; ++w)
                if (lb[w] < lb[b]) b = w;
            ord[m] = b;
            lb[b] += (long)(N - m) * (N - m);
        }
    }
};

// ===================================================================== P2M
// Mt_nm = sum_i q_i K_nm conj(S_n^m(u_i)), u = (x - centre)/h, where
// S_n^m(u) = |u|^n P_n^m(cos theta) e^{i m phi} (Condon-Shortley P) is built
// with the Cartesian recurrences
//   S_m^m = -(2m-1)(ux + i uy) S_{m-1}^{m-1},  S_{m+1}^m = (2m+1) uz S_m^m,
//   (n-m) S_n^m = (2n-1) uz S_{n-1}^m - (n+m-1) |u|^2 S_{n-2}^m.
// A wave handles 4 leaves, one 16-lane row per leaf and one particle per
// lane (rounds of 16 particles).  Harmonics are produced column by column
// (m-major position pos); each lane writes q conj(S) of 16 consecutive
// positions into a wave-private LDS tile, then lane (row, slot) sums entry
// `slot` over the 16 lanes of its row.  Lane (row, slot) thus accumulates
// positions slot, slot + 16, ... of its leaf; K_nm scales the final sums.
template <int N>
constexpr int p2m_pos(int n, int m) { return m * N - m * (m - 1) / 2 + (n - m); }

template <int N>
struct P2MOrder {  // position -> tri index
    int idx[tri(kMaxTerms3d)];
    constexpr P2MOrder() : idx()
    {
        for (int m = 0; m < N; ++m)
            for (int n = m; n < N; ++n) idx[p2m_pos<N>(n, m)] = tri(n) + m;
    }
};

// Kinds of k_p2m units: the leaves of a uniform tree; the leaves of the
// adaptive tree's global list (several levels); the boxes of one level's X
// lists (P2L).
enum { P2M_UNIFORM = 0, P2M_LEAVES = 1, P2M_P2L = 2 };

// P2L by Kelvin inversion: with u the source position relative to the target
// centre (in box units), Lt_nm = sum q K_nm conj(I_n^m(u)) where
// I_n^m(u) = P_n^m(cos theta) e^{i m phi} / |u|^(n+1) is the irregular solid
// harmonic, and I_n^m(u) = S_n^m(u / |u|^2) / |u|.  So the P2M recurrences
// at u / |u|^2 with charge q / |u| give the local expansion.
template <int N, int MODE>
__global__ void __launch_bounds__(256) k_p2m(P2MArgs a, const double* __restrict__ knm)
{
    constexpr int NX = tri(N);
    constexpr int NCH = (NX + 15) / 16;
    constexpr P2MOrder<N> ORD{};
    __shared__ double2 red[4][16][65];  // per wave [entry][lane], rows padded: conflict-free column reads
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int row = lane >> 4, slot = lane & 15;
    const int unit = (blockIdx.x * 4 + w) * 4 + row;
    const bool ok = unit < a.nleaf;
    double cx = 0.0, cy = 0.0, cz = 0.0, ih = 0.0;
    int b = 0, e = 0;      // particle range (P2L: ranges kb .. ke of x_rng)
    int level = a.level;
    int64_t box = unit;    // output expansion
    if (ok) {
        uint64_t key;
        if constexpr (MODE == P2M_P2L) {
            const int t = a.unit_box[unit];
            box = a.g2 + t;
            key = a.leaf_key[box];
            b = (int)a.x_off[t];
            e = (int)a.x_off[t + 1];
        } else {
            key = a.leaf_key[unit];
            if constexpr (MODE == P2M_LEAVES) {
                box = a.leaf_gid[unit];
                level = a.leaf_lvl[unit];
            }
            b = a.leaf_begin[unit];
            e = a.leaf_begin[unit + 1];
        }
        const double scale = a.geom[3];
        cx = box_centre(compact3(key >> 2), level, scale, a.geom[0]);
        cy = box_centre(compact3(key >> 1), level, scale, a.geom[1]);
        cz = box_centre(compact3(key), level, scale, a.geom[2]);
        ih = ldexp(1.0, level) / scale;
    }
    // P2L: this lane walks the concatenated ranges from position `slot` in
    // steps of 16 (cursor k, j); the row's rounds cover the total count
    int k = b, j = 0, total = e - b;
    auto settle = [&]() {
        while (k < e) {
            const int2 r = a.x_rng[k];
            if (j < r.y) break;
            const int over = j - r.y;
            if (++k < e) j = a.x_rng[k].x + over;
        }
    };
    if constexpr (MODE == P2M_P2L) {
        total = 0;
        for (int i = b; i < e; ++i) {
            const int2 r = a.x_rng[i];
            total += r.y - r.x;
        }
        if (b < e) {
            j = a.x_rng[b].x + slot;
            settle();
        }
    }
    int rounds = (total + 15) >> 4;
    for (int off = 32; off > 0; off >>= 1) rounds = max(rounds, __shfl_xor(rounds, off));
    double2 (*R)[65] = red[w];
    double2 acc[NCH];
#pragma unroll
    for (int c = 0; c < NCH; ++c) acc[c] = make_double2(0.0, 0.0);
    for (int rd = 0; rd < rounds; ++rd) {
        double ux = 0.0, uy = 0.0, uz = 0.0, q = 0.0;
        int p = -1;
        if constexpr (MODE == P2M_P2L) {
            if (k < e) p = j;
        } else {
            p = b + slot + 16 * rd;
            if (p >= e) p = -1;
        }
        if (p >= 0) {
            const double4 s4 = a.src[p];
            ux = (s4.x - cx) * ih;
            uy = (s4.y - cy) * ih;
            uz = (s4.z - cz) * ih;
            q = s4.w;
            if constexpr (MODE == P2M_P2L) {
                const double i2 = 1.0 / (ux * ux + uy * uy + uz * uz);
                ux *= i2;
                uy *= i2;
                uz *= i2;
                q *= sqrt(i2);
            }
        }
        if constexpr (MODE == P2M_P2L) {
            j += 16;
            settle();
        }
        const double r2 = ux * ux + uy * uy + uz * uz;
        auto put = [&](auto PI, double2 S) {
            constexpr int pos = decltype(PI)::value;
            R[pos & 15][lane] = make_double2(q * S.x, -q * S.y);  // K_nm is applied to the sums
            // keep harmonics from being computed ahead of their stores (register pressure)
            __builtin_amdgcn_sched_barrier(0);
            if constexpr ((pos & 15) == 15 || pos == NX - 1) {
                // the chunk is complete: lane (row, slot) sums entry `slot` of its row
                asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
                double sx = 0.0, sy = 0.0;
                if (slot <= (pos & 15)) {
#pragma unroll
                    for (int jj = 0; jj < 16; ++jj) {
                        const double2 v = R[slot][row * 16 + jj];
                        sx += v.x;
                        sy += v.y;
                    }
                }
                // materialise the sums here (the loaded entries die now, not at the loop end)
                asm volatile("" : "+v"(sx), "+v"(sy));
                acc[pos >> 4].x += sx;
                acc[pos >> 4].y += sy;
                asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
            }
        };
        double2 smm = make_double2(1.0, 0.0);
        sfor<0, N>([&](auto M) {
            constexpr int m = decltype(M)::value;
            if constexpr (m > 0) {
                const double c = -(double)(2 * m - 1);
                smm = make_double2(c * (ux * smm.x - uy * smm.y), c * (ux * smm.y + uy * smm.x));
            }
            put(std::integral_constant<int, p2m_pos<N>(m, m)>{}, smm);
            if constexpr (m + 1 < N) {
                double2 s2 = smm;
                double2 s1 = make_double2((2 * m + 1) * uz * smm.x, (2 * m + 1) * uz * smm.y);
                put(std::integral_constant<int, p2m_pos<N>(m + 1, m)>{}, s1);
                sfor<m + 2, N>([&](auto NN) {
                    constexpr int n = decltype(NN)::value;
                    constexpr double c1 = (double)(2 * n - 1) / (double)(n - m);
                    constexpr double c2 = (double)(n + m - 1) / (double)(n - m);
                    const double2 s0 = make_double2(c1 * uz * s1.x - c2 * r2 * s2.x, c1 * uz * s1.y - c2 * r2 * s2.y);
                    put(std::integral_constant<int, p2m_pos<N>(n, m)>{}, s0);
                    s2 = s1;
                    s1 = s0;
                });
            }
        });
    }
    double mx = 0.0;
    if (ok) {
        double2* out = a.outer + box * NX;
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            const int pos = c * 16 + slot;
            if (pos < NX) {
                const double kk = knm[ORD.idx[pos]];
                const double2 v = make_double2(kk * acc[c].x, kk * acc[c].y);
                if constexpr (MODE == P2M_P2L) {
                    const double2 o = out[ORD.idx[pos]];
                    out[ORD.idx[pos]] = make_double2(o.x + v.x, o.y + v.y);
                } else {
                    out[ORD.idx[pos]] = v;
                }
                mx = fmax(mx, fmax(fabs(v.x), fabs(v.y)));
            }
        }
    }
    if constexpr (MODE == P2M_UNIFORM) {
        if (a.absmax) wave_absmax(mx, a.absmax);
    } else if constexpr (MODE == P2M_LEAVES) {
        // the leaves of a wave may sit on different levels: one atomic per row
        if (a.absmax) {
            for (int off = 8; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off));
            if (ok && slot == 0) {
                uint32_t* dst = a.absmax + level;
                const uint32_t bits = __float_as_uint((float)mx);
                if (bits > __atomic_load_n(dst, __ATOMIC_RELAXED)) atomicMax(dst, bits);
            }
        }
    }
}

// ===================================================================== rotate-translate-rotate
enum { KIND_M2M = 0, KIND_M2L = 1, KIND_L2L = 2 };  // KIND_L2L: the L2L step of KIND_M2L alone

// rtr_tile(N) (targets per workgroup) is declared in ops3d.h.
enum { TR_M2L = 0, TR_M2M = 1, TR_L2L = 2 };

template <int N, int W, int TK>
__device__ __forceinline__ void rtr_phase_b(const double* __restrict__ tt, double2* lds, int lane)
{
    constexpr Partition<N, 4> PT{};
    constexpr int TW = rtr_tile(N);
    if (lane >= TW) return;
    sfor<0, N>([&](auto M) {
        constexpr int m = decltype(M)::value;
        if constexpr (PT.ord[m] == W) {
            constexpr int L = N - m;
            double2 x[L];
            sfor<0, L>([&](auto I) {
                constexpr int i = decltype(I)::value;
                x[i] = lds[(tri(m + i) + m) * TW + lane];
            });
            const double* t = tt + tr_base(N, m);
            sfor<0, L>([&](auto J) {
                constexpr int j = decltype(J)::value;
                double yr = 0.0, yi = 0.0;
                sfor<0, L>([&](auto I) {
                    constexpr int i = decltype(I)::value;
                    constexpr bool nz = (TK == TR_M2L) || (TK == TR_M2M && i <= j) || (TK == TR_L2L && i >= j);
                    if constexpr (nz) {
                        const double c = t[j * L + i];
                        yr = fma(c, x[i].x, yr);
                        yi = fma(c, x[i].y, yi);
                    }
                });
                lds[(tri(m + j) + m) * TW + lane] = make_double2(yr, yi);
            });
        }
    });
}

// The operator tables arrive as separate __restrict__ kernel arguments so the
// compiler can prove them read-only and fetch the wave-uniform coefficients
// with scalar loads.
#define RTR_TABLE_PARAMS                                                                             \
    const double2* __restrict__ Trot, const double2* __restrict__ Tephi, const double* __restrict__ Ttrans, \
        const int* __restrict__ Tdir_theta, const int* __restrict__ Tdir_trans, const int* __restrict__ Tdir_vec, \
        const int* __restrict__ Tclass_dirs
#define RTR_TABLE_ARGS Trot, Tephi, Ttrans, Tdir_theta, Tdir_trans, Tdir_vec, Tclass_dirs

// Barrier for LDS hand-over between the phases: waits for this wave's LDS
// traffic only, so global prefetch loads stay in flight across it.
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// Offsets of a wave's phase-A source degrees in its prefetch buffer.
template <int N, int W>
struct PrefetchLayout {
    int off[kMaxTerms3d + 1];
    int count;
    constexpr PrefetchLayout() : off(), count(0)
    {
        constexpr Partition<N, 4> PT{};
        for (int n = 0; n < N; ++n) {
            off[n] = count;
            if (PT.deg[n] == W) count += n + 1;
        }
    }
};

struct RtrStep {
    int sidx;                 // source box for this lane (-1: none)
    int dir;                  // direction table index
    int tk;                   // translation kind
    const double2* src;       // source expansion array
};

template <int N, int KIND, int W>
__device__ __forceinline__ void rtr_wave(RTR_TABLE_PARAMS, const RtrLevel& L, double2* lds, int lane)
{
    constexpr int NX = tri(N);
    constexpr int RS = rot_size(N);
    constexpr int TS = tr_size(N);
    constexpr Partition<N, 4> PT{};
    constexpr PrefetchLayout<N, W> PF{};
    constexpr int TW = rtr_tile(N);
    const bool lok = lane < TW;
    const int blk = xcd_local_block(blockIdx.x, gridDim.x);
    // tile-major block order: the 8 child classes (and direction parts) of a
    // tile are adjacent, so the workgroups of one XCD read overlapping sources
    int tile, c, part = 0;
    int pbeg = 0, pend = L.num_parents;  // this block's parent-list range
    if constexpr (KIND != KIND_M2M) {
        if (L.csorted) {
            // class-sorted list: every lane's parent has a class-c child.
            // Blocks stay tile-major with the 8 classes adjacent (tile k of
            // each class covers about the same region, so they share
            // sources in L2); a class with fewer tiles leaves its blocks idle.
            c = blk & 7;
            tile = blk >> 3;
            if (tile >= L.ctile[c + 1] - L.ctile[c]) return;  // the whole block
            pbeg = L.coff[c];
            pend = L.coff[c + 1];
        } else {
            const int per_tile = 8 * L.dsplit;
            tile = blk / per_tile;
            const int r = blk - tile * per_tile;
            part = r >> 3;
            c = r & 7;
        }
    } else {
        // M2M: the 8 child classes of a tile split over dsplit parts (adjacent blocks)
        tile = blk / L.dsplit;
        part = blk - tile * L.dsplit;
        c = 0;
    }
    // parent of this lane: the next of the level, or of the caller's list
    // (per-box orders: the parents whose children take this launch's N)
    const int Pi = pbeg + tile * TW + lane;
    const bool pvalid = lok && Pi < pend;
    const int P = (pvalid && L.plist) ? L.plist[Pi] : Pi;
    // coefficient stride of the expansion arrays (larger than tri(N) when
    // the boxes of a level carry different orders)
    const int64_t SX = L.stride > 0 ? L.stride : NX;
    int cb = 0, ce = 0;
    if (pvalid) {
        cb = L.child_begin[P];
        ce = L.child_begin[P + 1];
    }
    int tgt = -1;
    uint32_t tx = 0, ty = 0, tz = 0;
    if constexpr (KIND != KIND_M2M) {
        for (int i = cb; i < ce; ++i)
            if ((int)(L.child_keys[i] & 7u) == c) tgt = i;
        if (tgt >= 0) {
            const uint64_t k = L.child_keys[tgt];
            tx = compact3(k >> 2);
            ty = compact3(k >> 1);
            tz = compact3(k);
        }
    } else {
        // a parent without children (a leaf of the adaptive tree) keeps its P2M
        tgt = (pvalid && ce > cb) ? P : -1;
    }
    const int side = 1 << L.level;
    double2 acc[NX];
#pragma unroll
    for (int i = 0; i < NX; ++i) acc[i] = make_double2(0.0, 0.0);

    const int has_l2l = (KIND == KIND_L2L || (KIND == KIND_M2L && L.parent_local)) ? 1 : 0;
    const int ndir = (KIND == KIND_M2M) ? 8 : (KIND == KIND_L2L ? 1 : kDirsPerClass + has_l2l);
    // directions [d_beg, d_end) of this part (L.dsplit parts; the L2L step is in part 0)
    const int d_beg = (int)((int64_t)part * ndir / L.dsplit), d_end = (int)((int64_t)(part + 1) * ndir / L.dsplit);

    auto step_of = [&](int d) {
        RtrStep s;
        s.sidx = -1;
        if constexpr (KIND == KIND_M2M) {
            s.dir = kDiagBase + d;
            s.tk = TR_M2M;
            s.src = L.src;
            for (int i = cb; i < ce; ++i)
                if ((int)(L.child_keys[i] & 7u) == d) s.sidx = i;
        } else {
            if (has_l2l && d == 0) {
                s.dir = kDiagBase + (7 - c);
                s.tk = TR_L2L;
                s.src = L.parent_local;
                s.sidx = tgt >= 0 ? P : -1;
            } else {
                s.dir = Tclass_dirs[c * kDirsPerClass + (d - has_l2l)];
                s.tk = TR_M2L;
                s.src = L.src;
                if (tgt >= 0) {
                    const int sx = (int)tx + Tdir_vec[4 * s.dir], sy = (int)ty + Tdir_vec[4 * s.dir + 1],
                              sz = (int)tz + Tdir_vec[4 * s.dir + 2];
                    if (sx >= 0 && sy >= 0 && sz >= 0 && sx < side && sy < side && sz < side)
                        s.sidx = find_box(morton3(sx, sy, sz), L.lut, L.keys, L.size);
                }
            }
        }
        return s;
    };
    // the phase-A source coefficients of the next step, loaded one step ahead
    double2 pf[PF.count > 0 ? PF.count : 1];
    auto prefetch = [&](const RtrStep& s) {
        const double2* sp = s.src + (int64_t)(s.sidx < 0 ? 0 : s.sidx) * SX;
        sfor<0, N>([&](auto NN) {
            constexpr int n = decltype(NN)::value;
            if constexpr (PT.deg[n] == W) {
                sfor<0, n + 1>([&](auto K) {
                    constexpr int k = decltype(K)::value;
                    pf[PF.off[n] + k] = s.sidx >= 0 ? sp[tri(n) + k] : make_double2(0.0, 0.0);
                });
            }
        });
    };

    RtrStep nxt = step_of(d_beg);
    prefetch(nxt);
    for (int d = d_beg; d < d_end; ++d) {
        const RtrStep cur = nxt;
        const bool any = __ballot(cur.sidx >= 0) != 0ull;  // same lanes in every wave: uniform
        const double2* __restrict__ rot = Trot + (int64_t)Tdir_theta[cur.dir] * RS;
        const double2* __restrict__ eph = Tephi + cur.dir * N;
        if (any) {
            // ---- phase A: Rz(phi) then Ry(theta), by source degree
            sfor<0, N>([&](auto NN) {
                constexpr int n = decltype(NN)::value;
                if constexpr (PT.deg[n] == W) {
                    double2 av[n + 1];
                    sfor<0, n + 1>([&](auto K) {
                        constexpr int k = decltype(K)::value;
                        double2 v = pf[PF.off[n] + k];
                        if constexpr (k > 0) v = cmul(v, eph[k]);
                        av[k] = v;
                    });
                    const double2* r = rot + rot_base(n);
                    sfor<0, n + 1>([&](auto M) {
                        constexpr int m = decltype(M)::value;
                        double re = 0.0, im = 0.0;
                        sfor<0, n + 1>([&](auto K) {
                            constexpr int k = decltype(K)::value;
                            const double2 pq = r[m * (n + 1) + k];
                            re = fma(pq.x, av[k].x, re);
                            im = fma(pq.y, av[k].y, im);
                        });
                        if (lok) lds[(tri(n) + m) * TW + lane] = make_double2(re, im);
                    });
                }
            });
            lds_barrier();
            // ---- phase B: axial translation, by order
            const int tsel = (cur.tk == TR_M2L) ? Tdir_trans[cur.dir] : (cur.tk == TR_M2M ? kTransM2M : kTransL2L);
            const double* __restrict__ tt = Ttrans + (int64_t)tsel * TS;
            if constexpr (KIND == KIND_M2M) {
                rtr_phase_b<N, W, TR_M2M>(tt, lds, lane);
            } else if constexpr (KIND == KIND_L2L) {
                rtr_phase_b<N, W, TR_L2L>(tt, lds, lane);
            } else {
                if (cur.tk == TR_L2L) rtr_phase_b<N, W, TR_L2L>(tt, lds, lane);
                else rtr_phase_b<N, W, TR_M2L>(tt, lds, lane);
            }
            lds_barrier();
        }
        // issue the next step's source loads; they complete under phase C
        if (d + 1 < d_end) {
            nxt = step_of(d + 1);
            prefetch(nxt);
        }
        if (any) {
            // ---- phase C: Ry(-theta) then Rz(-phi), accumulate, by target degree
            sfor<0, N>([&](auto JJ) {
                constexpr int j = decltype(JJ)::value;
                if constexpr (PT.deg[j] == W) {
                    double2 yv[j + 1];
                    sfor<0, j + 1>([&](auto M) {
                        constexpr int m = decltype(M)::value;
                        yv[m] = lok ? lds[(tri(j) + m) * TW + lane] : make_double2(0.0, 0.0);
                    });
                    const double2* r = rot + rot_base(j);
                    sfor<0, j + 1>([&](auto MU) {
                        constexpr int mu = decltype(MU)::value;
                        double re = 0.0, im = 0.0;
                        sfor<0, j + 1>([&](auto M) {
                            constexpr int m = decltype(M)::value;
                            const double2 pq = r[mu * (j + 1) + m];
                            if constexpr (((mu - m) & 1) == 0) {
                                re = fma(pq.x, yv[m].x, re);
                                im = fma(pq.y, yv[m].y, im);
                            } else {
                                re = fma(-pq.x, yv[m].x, re);
                                im = fma(-pq.y, yv[m].y, im);
                            }
                        });
                        double2 z = make_double2(re, im);
                        if constexpr (mu > 0) z = cmulc(z, eph[mu]);
                        acc[tri(j) + mu].x += z.x;
                        acc[tri(j) + mu].y += z.y;
                    });
                }
            });
            lds_barrier();
        }
    }
    double mx = 0.0;
    if (tgt >= 0) {
        double2* out = L.dsplit > 1 ? L.partial + ((int64_t)part * L.size + tgt) * NX : L.dst + (int64_t)tgt * SX;
        // coefficients beyond this launch's order are zero
        if constexpr (W == 0)
            for (int64_t i = NX; i < SX; ++i) out[i] = make_double2(0.0, 0.0);
        sfor<0, N>([&](auto JJ) {
            constexpr int j = decltype(JJ)::value;
            if constexpr (PT.deg[j] == W) {
                sfor<0, j + 1>([&](auto MU) {
                    constexpr int mu = decltype(MU)::value;
                    out[tri(j) + mu] = acc[tri(j) + mu];
                    if constexpr (KIND == KIND_M2M) mx = fmax(mx, fmax(fabs(acc[tri(j) + mu].x), fabs(acc[tri(j) + mu].y)));
                });
            }
        });
    }
    if constexpr (KIND == KIND_M2M)
        if (L.absmax) wave_absmax(mx, L.absmax);
}

template <int N, int KIND>
__global__ void __launch_bounds__(256) k_rtr(RTR_TABLE_PARAMS, RtrLevel L)
{
    __shared__ double2 lds[tri(N) * rtr_tile(N)];
    const int lane = threadIdx.x & 63;
    switch (threadIdx.x >> 6) {
    case 0: rtr_wave<N, KIND, 0>(RTR_TABLE_ARGS, L, lds, lane); break;
    case 1: rtr_wave<N, KIND, 1>(RTR_TABLE_ARGS, L, lds, lane); break;
    case 2: rtr_wave<N, KIND, 2>(RTR_TABLE_ARGS, L, lds, lane); break;
    default: rtr_wave<N, KIND, 3>(RTR_TABLE_ARGS, L, lds, lane); break;
    }
}

// ===================================================================== L2P
// L2P with normalised locals: phi = (1/h) Re G(u), G = sum w_m Lt_nm K_nm S_n^m(u),
// field = -(1/h^2) grad_u Re G, using
//   d/dz S_n^m = (n+m) S_{n-1}^m,  (d/dx + i d/dy) S_n^m = S_{n-1}^{m+1},
//   (d/dx - i d/dy) S_n^m = -(n+m)(n+m-1) S_{n-1}^{m-1}  (m >= 1),
//   (d/dx - i d/dy) S_n^0 = conj(S_{n-1}^1).
// The field is the Cartesian gradient, so it stays finite on the box axis
// where the reference's spherical form divides by sin(theta)
// (src/cpu/multipole3d.cpp:1205-1206).
// A compiler memory fence every CF columns of m bounds how far coefficient
// loads are merged and hoisted: CF = 1 keeps each column's loads at their use
// (fewest registers, but the m +- 1 neighbours are loaded again per column);
// larger CF loads more coefficients once, with more registers live.
template <int N, int CF = 1>
__device__ __forceinline__ void l2p_eval(const double2* __restrict__ Lt, const double* __restrict__ knm, double ux, double uy,
                                         double uz, double& phi, double& gx, double& gy, double& gz)
{
    const double r2 = ux * ux + uy * uy + uz * uz;
    double ph = 0.0;
    double2 dz = make_double2(0.0, 0.0), dp = make_double2(0.0, 0.0), dm = make_double2(0.0, 0.0);
    auto coef = [&](auto NI, auto MI) -> double2 {
        constexpr int n = decltype(NI)::value, m = decltype(MI)::value;
        const double2 l = Lt[tri(n) + m];
        const double k = knm[tri(n) + m] * (m == 0 ? 1.0 : 2.0);
        return make_double2(l.x * k, l.y * k);
    };
    auto use = [&](auto PI, auto QI, double2 S) {
        constexpr int p = decltype(PI)::value, q = decltype(QI)::value;
        {
            const double2 c = coef(PI, QI);
            ph += c.x * S.x - c.y * S.y;
        }
        if constexpr (p + 1 < N) {
            const double2 cz = coef(std::integral_constant<int, p + 1>{}, QI);
            const double2 t = cmul(cz, S);
            dz.x += (double)(p + 1 + q) * t.x;
            dz.y += (double)(p + 1 + q) * t.y;
            if constexpr (q >= 1) {
                const double2 cp = coef(std::integral_constant<int, p + 1>{}, std::integral_constant<int, q - 1>{});
                const double2 u = cmul(cp, S);
                dp.x += u.x;
                dp.y += u.y;
            }
            {
                const double2 cm = coef(std::integral_constant<int, p + 1>{}, std::integral_constant<int, q + 1>{});
                const double2 u = cmul(cm, S);
                constexpr double f = -(double)((p + q + 2) * (p + q + 1));
                dm.x += f * u.x;
                dm.y += f * u.y;
            }
            if constexpr (q == 1) {
                const double2 c0 = coef(std::integral_constant<int, p + 1>{}, std::integral_constant<int, 0>{});
                const double2 u = cmul(c0, make_double2(S.x, -S.y));
                dm.x += u.x;
                dm.y += u.y;
            }
        }
    };
    double2 smm = make_double2(1.0, 0.0);
    sfor<0, N>([&](auto M) {
        constexpr int m = decltype(M)::value;
        // compiler memory fence (see CF above; without any, every load is
        // merged and hoisted into one live set)
        if constexpr (m % CF == 0) asm volatile("" ::: "memory");
        if constexpr (m > 0) {
            const double c = -(double)(2 * m - 1);
            smm = make_double2(c * (ux * smm.x - uy * smm.y), c * (ux * smm.y + uy * smm.x));
        }
        use(std::integral_constant<int, m>{}, M, smm);
        if constexpr (m + 1 < N) {
            double2 s2 = smm;
            double2 s1 = make_double2((2 * m + 1) * uz * smm.x, (2 * m + 1) * uz * smm.y);
            use(std::integral_constant<int, m + 1>{}, M, s1);
            sfor<m + 2, N>([&](auto NN) {
                constexpr int n = decltype(NN)::value;
                constexpr double c1 = (double)(2 * n - 1) / (double)(n - m);
                constexpr double c2 = (double)(n + m - 1) / (double)(n - m);
                const double2 s0 = make_double2(c1 * uz * s1.x - c2 * r2 * s2.x, c1 * uz * s1.y - c2 * r2 * s2.y);
                use(NN, M, s0);
                s2 = s1;
                s1 = s0;
            });
        }
    });
    phi = ph;
    gx = 0.5 * (dp.x + dm.x);
    gy = 0.5 * (dp.y - dm.y);
    gz = dz.x;
}

// L2P alone: the far field of every evaluated particle in sorted order.  Run
// before the pair kernel, which then starts from it, so that the pair loop
// is not held to the register budget (and occupancy) of the L2P recurrences.
// ADAPTIVE: leaves sit on several levels (leaf_lvl; no local below level
// local_min),
// and the multipoles of the leaf's W list are evaluated at the particle too
// (M2P).  M2P by Kelvin inversion: with u the particle relative to the source
// centre (box units), r = |u| and f, g = l2p_eval of the multipole at
// u / r^2, the potential is f / r and its gradient
// (g - 2 (u/r)((u/r).g) - f u) / r^3 (I_n^m(u) = S_n^m(u / r^2) / r).
template <int N, bool ADAPTIVE>
__global__ void __launch_bounds__(256) k_l2p(LeafArgs a, const double2* __restrict__ local, const double* __restrict__ knm)
{
    const int64_t s = xcd_local_block(blockIdx.x, gridDim.x) * (int64_t)blockDim.x + threadIdx.x;
    if (s >= a.n) return;
    if (a.perm[s] >= a.n_out) return;
    const int leaf = a.leaf_of[s];
    const uint64_t key = a.leaf_key[leaf];
    const int level = ADAPTIVE ? a.leaf_lvl[leaf] : a.level;
    const double4 t = a.src[s];
    const double scale = a.geom[3];
    double fp = 0.0, fx = 0.0, fy = 0.0, fz = 0.0;
    if (!ADAPTIVE || level >= a.local_min) {
        const uint32_t ix = compact3(key >> 2), iy = compact3(key >> 1), iz = compact3(key);
        const double cx = box_centre(ix, level, scale, a.geom[0]);
        const double cy = box_centre(iy, level, scale, a.geom[1]);
        const double cz = box_centre(iz, level, scale, a.geom[2]);
        const double ih = ldexp(1.0, level) / scale;
        const int64_t box = ADAPTIVE ? a.leaf_gid[leaf] : leaf;
        double lp, gx, gy, gz;
        // uniform tree up to 13 terms: fence every 4 columns (the repeated
        // coefficient loads, which bound this kernel, drop by half, at 166
        // VGPRs for 13 terms); above 13 terms that passes 256 VGPRs, and the
        // adaptive kernel also runs the W-list evaluations: CF = 1 there
        l2p_eval<N, (ADAPTIVE || N > 13) ? 1 : 4>(local + box * tri(N), knm, (t.x - cx) * ih, (t.y - cy) * ih, (t.z - cz) * ih, lp,
                                      gx, gy, gz);
        const double ih2 = ih * ih;
        fp = lp * ih;
        fx = -gx * ih2;
        fy = -gy * ih2;
        fz = -gz * ih2;
    }
    if constexpr (ADAPTIVE) {
        const uint32_t wb = a.w_off[leaf], we = a.w_off[leaf + 1];
        for (uint32_t k = wb; k < we; ++k) {
            const int g = a.w_gid[k];
            const int2 br = a.box_range[g];
            if (br.y - br.x <= a.direct_max) continue;  // summed directly by the pair kernel
            int lv = 0;
            while (lv + 1 < a.depth && g >= a.loff[lv + 1]) ++lv;
            const uint64_t sk = a.keys_all[g];
            const double ih = ldexp(1.0, lv) / scale;
            const double ux = (t.x - box_centre(compact3(sk >> 2), lv, scale, a.geom[0])) * ih;
            const double uy = (t.y - box_centre(compact3(sk >> 1), lv, scale, a.geom[1])) * ih;
            const double uz = (t.z - box_centre(compact3(sk), lv, scale, a.geom[2])) * ih;
            const double i2 = 1.0 / (ux * ux + uy * uy + uz * uz);
            const double ir = sqrt(i2);
            double f, gx, gy, gz;
            l2p_eval<N>(a.outer_all + (int64_t)g * tri(N), knm, ux * i2, uy * i2, uz * i2, f, gx, gy, gz);
            // gradient in u of f(u / r^2) / r
            const double ud = (ux * gx + uy * gy + uz * gz) * i2;  // (u.g) / r^2
            const double ir3 = i2 * ir;
            const double dx = (gx - 2.0 * ux * ud - f * ux) * ir3;
            const double dy = (gy - 2.0 * uy * ud - f * uy) * ir3;
            const double dz = (gz - 2.0 * uz * ud - f * uz) * ir3;
            fp += f * ir * ih;
            const double ih2 = ih * ih;
            fx -= dx * ih2;
            fy -= dy * ih2;
            fz -= dz * ih2;
        }
    }
    a.far[s] = make_double4(fp, fx, fy, fz);
}

// ===================================================================== launchers
template <int N>
hipError_t do_p2m(const P2MArgs& a, hipStream_t st)
{
    const int nb = (a.nleaf + 15) / 16;  // 4 waves x 4 leaves
    if (nb > 0) {
        if (a.leaf_gid) hipLaunchKernelGGL((k_p2m<N, P2M_LEAVES>), dim3(nb), dim3(256), 0, st, a, a.knm);
        else hipLaunchKernelGGL((k_p2m<N, P2M_UNIFORM>), dim3(nb), dim3(256), 0, st, a, a.knm);
    }
    return hipGetLastError();
}

template <int N>
hipError_t do_p2l(const P2MArgs& a, hipStream_t st)
{
    const int nb = (a.nleaf + 15) / 16;
    if (nb > 0) hipLaunchKernelGGL((k_p2m<N, P2M_P2L>), dim3(nb), dim3(256), 0, st, a, a.knm);
    return hipGetLastError();
}

template <int N>
hipError_t do_rtr(int kind, const RtrTables& T, const RtrLevel& L, hipStream_t st)
{
    const int tiles = (L.num_parents + rtr_tile(N) - 1) / rtr_tile(N);
    if (tiles == 0) return hipSuccess;
    if (L.tiles != tiles && kind != KIND_M2M && !L.csorted) return hipErrorInvalidValue;
    if (L.csorted && kind != KIND_M2M) {
        if (L.dsplit != 1 || !L.plist || L.coff[8] != L.num_parents) return hipErrorInvalidValue;
        for (int c = 0; c < 8; ++c)
            if (L.ctile[c + 1] - L.ctile[c] != (L.coff[c + 1] - L.coff[c] + rtr_tile(N) - 1) / rtr_tile(N))
                return hipErrorInvalidValue;
        int mt = 0;
        for (int c = 0; c < 8; ++c) mt = std::max(mt, L.ctile[c + 1] - L.ctile[c]);
        if (mt == 0) return hipSuccess;
        if (kind == KIND_L2L)
            hipLaunchKernelGGL((k_rtr<N, KIND_L2L>), dim3(8 * mt), dim3(256), 0, st, T.rot, T.ephi, T.trans,
                               T.dir_theta, T.dir_trans, T.dir_vec, T.class_dirs, L);
        else
            hipLaunchKernelGGL((k_rtr<N, KIND_M2L>), dim3(8 * mt), dim3(256), 0, st, T.rot, T.ephi, T.trans,
                               T.dir_theta, T.dir_trans, T.dir_vec, T.class_dirs, L);
        return hipGetLastError();
    }
    if (kind == KIND_M2M)
        hipLaunchKernelGGL((k_rtr<N, KIND_M2M>), dim3(tiles * L.dsplit), dim3(256), 0, st, T.rot, T.ephi, T.trans,
                           T.dir_theta, T.dir_trans, T.dir_vec, T.class_dirs, L);
    else if (kind == KIND_L2L)
        hipLaunchKernelGGL((k_rtr<N, KIND_L2L>), dim3(tiles * 8 * L.dsplit), dim3(256), 0, st, T.rot, T.ephi, T.trans,
                           T.dir_theta, T.dir_trans, T.dir_vec, T.class_dirs, L);
    else
        hipLaunchKernelGGL((k_rtr<N, KIND_M2L>), dim3(tiles * 8 * L.dsplit), dim3(256), 0, st, T.rot, T.ephi, T.trans,
                           T.dir_theta, T.dir_trans, T.dir_vec, T.class_dirs, L);
    return hipGetLastError();
}

template <int N>
hipError_t do_l2p(const LeafArgs& a, hipStream_t st)
{
    const int nb = (int)((a.n + 255) / 256);
    if (nb > 0) {
        if (a.leaf_lvl) hipLaunchKernelGGL((k_l2p<N, true>), dim3(nb), dim3(256), 0, st, a, a.local, a.knm);
        else hipLaunchKernelGGL((k_l2p<N, false>), dim3(nb), dim3(256), 0, st, a, a.local, a.knm);
    }
    return hipGetLastError();
}

}  // namespace lfmm
