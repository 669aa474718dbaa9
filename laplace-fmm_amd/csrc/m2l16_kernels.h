// m2l16_kernels.h -- M2L on the matrix cores in split FP16 with FP32
// accumulation (gfx950 v_mfma_f32_32x32x16_f16).
//
// For one level and one child class, the targets of a workgroup share their
// 189 interaction directions (reference M2L :152-190 and Downward :328-349 of
// src/cpu/multipole3d.cpp).  Per direction d the operator is the dense real
// matrix K_d (N^2 x N^2, built on the host from the FP64 rotation and
// translation tables, m2l_dense.cpp), so one direction step of a workgroup
// is the GEMM  L[:, targets] += K_d X[:, sources(targets, d)].
//
// Precision: K_d and the source expansions X are scaled by powers of two and
// split into fp16 hi + lo (22 significant bits).  Each product is
//   hi*hi + hi*lo + lo*hi   (the lo*lo term, ~2^-22 relative, is dropped)
// computed exactly on the matrix cores and accumulated in FP32 over all K and
// all 189 directions (k_m2l16r: rows of degree n <= 5 per direction, then
// summed over the directions).  From 12 terms, operator blocks (32 rows x 16 K) whose
// largest entry is below 2^-11 of the largest operator entry use hi*hi alone,
// and blocks below 2^-22 are skipped: either way the dropped terms are bounded
// by 2^-22 max|K| max|X|, the absolute bound of the split itself.  The result
// is scaled back and added in FP64 to the local expansion, which already holds
// the FP64 L2L of the parent.
//
// Operands (32x32x16 f16 MFMA; lane l, slot s in 0..7):
//   A  row l % 32, K index 8 * (l / 32) + s     (operator, from global/L2)
//   B  col l % 32, K index 8 * (l / 32) + s     (sources, staged in LDS)
//   C  col l % 32, row 8 * (v / 4) + 4 * (l / 32) + v % 4, v = 0..15
// A and B use the same (lane, slot) -> K assignment, so the sum over K does
// not depend on how the hardware numbers K within an instruction.
//
// The same kernel (k_m2l16r<N, CTW, true>) runs the M2M into a level on
// uniform trees: the targets are the parents, the 8 child classes are the
// directions (reference M2M, src/cpu/multipole3d.cpp:115-150), and each
// target's class-c child is read from the children level's class-major
// packing, whose entry for parent t is that child.
#pragma once
#include "ops3d_kernels.h"  // xcd_local_block, lds_barrier

namespace lfmm {

typedef _Float16 m16_h8 __attribute__((ext_vector_type(8)));
typedef float m16_acc __attribute__((ext_vector_type(16)));
typedef unsigned int m16_u4 __attribute__((ext_vector_type(4)));

template <int N>
struct M16 {
    static constexpr int NR = N * N;            // real coordinates of an expansion
    static constexpr int RT = (NR + 31) / 32;   // row tiles of 32
    static constexpr int KC = (NR + 15) / 16;   // K chunks of 16
    static constexpr int RW = (RT + 3) / 4;     // row tiles per wave (at most)
    static constexpr int PIECES = KC * 4;       // 16-byte pieces of a packed source box
};

// Scale of a level's sources: the largest |coefficient| lands below 2^14.
__device__ inline double m16_xscale(uint32_t maxbits)
{
    const float mx = __uint_as_float(maxbits);
    if (!(mx > 0.0f)) return 1.0;
    const int e = max((int)((maxbits >> 23) & 0xff) - 126, -100);  // mx < 2^e
    return ldexp(1.0, 14 - e);
}

// Real coordinate r (< N^2) -> offset of the double inside [tri(N)] double2.
__device__ inline int m16_coef_offset(int r)
{
    const int n = (int)sqrtf((float)r + 0.5f);
    const int q = r - n * n;
    const int m = (q + 1) >> 1;
    const int part = (q != 0 && !(q & 1)) ? 1 : 0;
    return 2 * (tri(n) + m) + part;
}

// Packs a level's normalised multipoles into the B operand layout:
// [box][KC][part hi/lo][half][8 x fp16]; box `size` is all zeros.
template <int N>
__global__ void __launch_bounds__(256) k_m2l16_pack(const double2* __restrict__ outer, int size,
                                                    const uint32_t* __restrict__ xmax, uint4* __restrict__ xp)
{
    using S = M16<N>;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // (box, kc, half)
    if (i >= (int64_t)(size + 1) * S::KC * 2) return;
    const int half = (int)(i & 1);
    const int kc = (int)((i >> 1) % S::KC);
    const int64_t box = (i >> 1) / S::KC;
    const double sc = m16_xscale(*xmax);
    const double* src = (const double*)(outer + box * tri(N));
    m16_h8 hv, lv;
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        const int k = kc * 16 + half * 8 + s;
        double x = 0.0;
        if (box < size && k < S::NR) x = src[m16_coef_offset(k)] * sc;
        const _Float16 h = (_Float16)x;
        hv[s] = h;
        lv[s] = (_Float16)(x - (double)h);
    }
    uint4* dst = xp + (box * S::KC + kc) * 4;
    dst[half] = __builtin_bit_cast(uint4, hv);
    dst[2 + half] = __builtin_bit_cast(uint4, lv);
}

// Packs a level's normalised multipoles class-major for k_m2l16r:
// piece (kc, hi/lo, half) of the class-c child of parent P lands at
// xp[(c * PIECES + 4 kc + 2 hl + half) * (num_parents + 1) + P], zeros where
// the child is absent and in slot num_parents.  For one direction, the
// sources of a tile's targets are one class and a shifted block of parents,
// so a source fragment reads runs of consecutive 16-byte entries.
// A thread packs one (parent, kc, half): the 2 KC threads of a parent are
// consecutive, so together they read each child's expansion in one
// contiguous sweep, and blocks are numbered XCD-locally, so the partial
// 16-byte rows of neighbouring parents merge in one L2.
template <int N>
__global__ void __launch_bounds__(256) k_m2l16_pack_cls(const double2* __restrict__ outer, int num_parents,
                                                        const int32_t* __restrict__ child_begin,
                                                        const uint64_t* __restrict__ child_keys,
                                                        const uint32_t* __restrict__ xmax, uint4* __restrict__ xp,
                                                        int64_t stride)
{
    using S = M16<N>;
    const int64_t np1 = (int64_t)num_parents + 1;
    const int64_t i = (int64_t)xcd_local_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;  // (parent, kc, half)
    if (i >= np1 * S::KC * 2) return;
    const int64_t P = i / (S::KC * 2);
    const int kh = (int)(i - P * (S::KC * 2));
    const int kc = kh >> 1, half = kh & 1;
    const double sc = m16_xscale(*xmax);
    // the parent's children (at most 8, in key order): their classes from
    // independent key loads, so the 8 classes' coefficient loads can all be
    // in flight together
    int cb = 0, ce = 0;
    if (P < num_parents) {
        cb = child_begin[P];
        ce = child_begin[P + 1];
    }
    int kcls[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) kcls[i] = cb + i < ce ? (int)(child_keys[cb + i] & 7u) : -1;
    int off[8];
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        const int k = kc * 16 + half * 8 + s;
        off[s] = k < S::NR ? m16_coef_offset(k) : -1;
    }
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        int child = -1;
#pragma unroll
        for (int i = 0; i < 8; ++i)
            if (kcls[i] == c) child = cb + i;
        m16_h8 hv = {}, lv = {};
        if (child >= 0) {
            const double* src = (const double*)(outer + (int64_t)child * stride);
#pragma unroll
            for (int s = 0; s < 8; ++s) {
                const double x = off[s] >= 0 ? src[off[s]] * sc : 0.0;
                const _Float16 h = (_Float16)x;
                hv[s] = h;
                lv[s] = (_Float16)(x - (double)h);
            }
        }
        uint4* plane = xp + (int64_t)(c * S::PIECES + kc * 4 + half) * np1;
        plane[P] = __builtin_bit_cast(uint4, hv);
        plane[2 * np1 + P] = __builtin_bit_cast(uint4, lv);
    }
}

// Block skipping (the prefix table) pays where the operator blocks are many
// and sparse; below 12 terms the straight-line MFMA stream is faster.
constexpr bool m16_skip(int N) { return N >= 12; }

template <int N, int CT>
__global__ void __launch_bounds__(256) k_m2l16(M2L16Args a, const uint4* __restrict__ kmat, const uint4* __restrict__ xp,
                                               const int* __restrict__ class_dirs, const int* __restrict__ dir_vec,
                                               const int2* __restrict__ kpre)
{
    using S = M16<N>;
    constexpr int RT = S::RT, KC = S::KC, RW = S::RW, NR = S::NR, PIECES = S::PIECES;
    constexpr int TC = 32 * CT;        // targets (columns) per workgroup
    constexpr int TPC = 256 / TC;      // loader threads per column
    constexpr int PPT = (PIECES + TPC - 1) / TPC;  // 16-byte pieces per loader thread
    static_assert(256 % TC == 0, "loader split");
    __shared__ uint4 bs[2][CT][KC][2][64];

    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave index, uniform
    // row tiles w, w + 4, ... of this wave; all valid when RT divides over the 4 waves
    constexpr bool kEven = (RT % 4) == 0;
    const int blk = xcd_local_block(blockIdx.x, gridDim.x);
    // tile-major block order: the 8 child classes (and direction parts) of a
    // tile are adjacent, so the workgroups of one XCD read overlapping sources
    const int tile = blk / (8 * a.dsplit);
    const int part = (blk - tile * 8 * a.dsplit) >> 3;
    const int c = blk & 7;
    const int d_beg = part * kDirsPerClass / a.dsplit, d_end = (part + 1) * kDirsPerClass / a.dsplit;
    const int side = 1 << a.level;

    // target box of column `col`: the class-c child of parent tile * TC + col
    auto target_of = [&](int col, uint32_t& x, uint32_t& y, uint32_t& z) -> int {
        const int P = tile * TC + col;
        int t = -1;
        if (P < a.num_parents) {
            const int cb = a.child_begin[P], ce = a.child_begin[P + 1];
            for (int i = cb; i < ce; ++i)
                if ((int)(a.child_keys[i] & 7u) == c) t = i;
        }
        if (t >= 0) {
            const uint64_t k = a.child_keys[t];
            x = compact3(k >> 2);
            y = compact3(k >> 1);
            z = compact3(k);
        }
        return t;
    };

    // ---- B staging: loader thread (column lcol, part lq) moves PPT pieces
    const int lcol = tid / TPC, lq = tid % TPC;
    uint32_t lx = 0, ly = 0, lz = 0;
    const int ltgt = target_of(lcol, lx, ly, lz);
    auto source_piece = [&](int d) -> const uint4* {
        const int dir = class_dirs[c * kDirsPerClass + d];
        int box = a.size;  // the zero box
        if (ltgt >= 0) {
            const int sx = (int)lx + dir_vec[4 * dir], sy = (int)ly + dir_vec[4 * dir + 1],
                      sz = (int)lz + dir_vec[4 * dir + 2];
            if (sx >= 0 && sy >= 0 && sz >= 0 && sx < side && sy < side && sz < side) {
                const int s = find_box(morton3(sx, sy, sz), a.lut, a.keys, a.size);
                if (s >= 0) box = s;
            }
        }
        return xp + (int64_t)box * PIECES + lq;
    };
    auto stage_slot = [&](int buf, int i) -> uint4* {
        const int pc = lq + TPC * i;
        return &bs[buf][lcol >> 5][pc >> 2][(pc >> 1) & 1][(pc & 1) * 32 + (lcol & 31)];
    };
    constexpr bool kFullPieces = (PIECES % TPC) == 0;
    auto piece_ok = [&](int i) { return kFullPieces || lq + TPC * i < PIECES; };
    // Operator fragments of this wave's row tiles stream through a rolling
    // window: chunk kc + LA is loaded while chunk kc is multiplied (across the
    // step boundary into the next direction).  Only the blocks the precision
    // needs are loaded and multiplied: per (direction, row tile), chunks
    // kc < k3 use hi*hi + hi*lo + lo*hi, k3 <= kc < k1 hi*hi alone, the rest
    // nothing (host table kpre; see pack_m2l16).
    // one row tile per wave: the whole next step is prefetched; two: 4 chunks ahead
    constexpr int LA = RW == 1 ? KC : 4;
    auto dir_of = [&](int d) { return class_dirs[c * kDirsPerClass + d]; };
    auto row_tile = [&](int r) { return kEven ? w + 4 * r : min(w + 4 * r, RT - 1); };
    auto prefix = [&](int dir, int r) -> int2 {
        if (!kEven && w + 4 * r >= RT) return make_int2(0, 0);
        if constexpr (!m16_skip(N)) return make_int2(KC, KC);
        const int2 v = kpre[dir * RT + row_tile(r)];
        // wave-uniform: keep the block bounds in scalar registers (scalar branches)
        return make_int2(__builtin_amdgcn_readfirstlane(v.x), __builtin_amdgcn_readfirstlane(v.y));
    };
    // unneeded fragments are read from one fixed block (cache-resident), so
    // every load is unconditional and the loop stays straight-line code
    auto frag = [&](int dir, int r, int kc, int hl, bool need) -> uint4 {
        const int64_t blk = need ? (int64_t)dir * (RT * KC * 2) + (row_tile(r) * KC + kc) * 2 + hl : 0;
        return kmat[blk * 64 + lane];
    };

    m16_acc acc[RW][CT];
#pragma unroll
    for (int r = 0; r < RW; ++r)
#pragma unroll
        for (int ct = 0; ct < CT; ++ct)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[r][ct][v] = 0.0f;

    // step d_beg: sources staged, first LA operator chunks loaded
    int dcur = dir_of(d_beg);
    int2 pc[RW];
    uint4 hh[RW][LA], hl[RW][LA];
#pragma unroll
    for (int r = 0; r < RW; ++r) {
        pc[r] = prefix(dcur, r);
#pragma unroll
        for (int k = 0; k < LA; ++k) {
            hh[r][k] = frag(dcur, r, k, 0, k < pc[r].y);
            hl[r][k] = frag(dcur, r, k, 1, k < pc[r].x);
        }
    }
    {
        const uint4* p = source_piece(d_beg);
        uint4 t[PPT];
#pragma unroll
        for (int i = 0; i < PPT; ++i) t[i] = piece_ok(i) ? p[TPC * i] : make_uint4(0, 0, 0, 0);
#pragma unroll
        for (int i = 0; i < PPT; ++i)
            if (piece_ok(i)) *stage_slot(0, i) = t[i];
    }
    lds_barrier();
    for (int d = d_beg; d < d_end; ++d) {
        const int buf = (d - d_beg) & 1;
        const int dn = min(d + 1, d_end - 1);  // the last step re-stages its own data (never read)
        const int dnext = dir_of(dn);
        int2 pn[RW];
#pragma unroll
        for (int r = 0; r < RW; ++r) pn[r] = prefix(dnext, r);
        // next step's sources, in flight under the MFMAs
        uint4 nb[PPT];
        {
            const uint4* p = source_piece(dn);
#pragma unroll
            for (int i = 0; i < PPT; ++i) nb[i] = piece_ok(i) ? p[TPC * i] : make_uint4(0, 0, 0, 0);
        }
        uint4 ch[RW][KC], cl[RW][KC];  // this step's chunks LA.. (loaded LA chunks ahead)
        uint4 nh[RW][LA], nl[RW][LA];  // next step's chunks 0..LA-1
        uint4 fb[2][CT][2];
#pragma unroll
        for (int ct = 0; ct < CT; ++ct) {
            fb[0][ct][0] = bs[buf][ct][0][0][lane];
            fb[0][ct][1] = bs[buf][ct][0][1][lane];
        }
#pragma unroll
        for (int kc = 0; kc < KC; ++kc) {
            // issue the operator chunk LA ahead
            const int ka = kc + LA;
#pragma unroll
            for (int r = 0; r < RW; ++r) {
                if (ka < KC) {
                    ch[r][ka] = frag(dcur, r, ka, 0, ka < pc[r].y);
                    cl[r][ka] = frag(dcur, r, ka, 1, ka < pc[r].x);
                } else {
                    nh[r][ka - KC] = frag(dnext, r, ka - KC, 0, ka - KC < pn[r].y);
                    nl[r][ka - KC] = frag(dnext, r, ka - KC, 1, ka - KC < pn[r].x);
                }
            }
            // source fragments of chunk kc + 1 are read while chunk kc multiplies
            if (kc + 1 < KC) {
#pragma unroll
                for (int ct = 0; ct < CT; ++ct) {
                    fb[(kc + 1) & 1][ct][0] = bs[buf][ct][kc + 1][0][lane];
                    fb[(kc + 1) & 1][ct][1] = bs[buf][ct][kc + 1][1][lane];
                }
            }
#pragma unroll
            for (int r = 0; r < RW; ++r) {
                const uint4 ah = kc < LA ? hh[r][kc] : ch[r][kc];
                const uint4 al = kc < LA ? hl[r][kc] : cl[r][kc];
                const m16_h8 xh = __builtin_bit_cast(m16_h8, ah);
                const m16_h8 xl = __builtin_bit_cast(m16_h8, al);
                if (kc < pc[r].y) {
#pragma unroll
                    for (int ct = 0; ct < CT; ++ct) {
                        const m16_h8 bh = __builtin_bit_cast(m16_h8, fb[kc & 1][ct][0]);
                        acc[r][ct] = __builtin_amdgcn_mfma_f32_32x32x16_f16(xh, bh, acc[r][ct], 0, 0, 0);
                    }
                }
                if (kc < pc[r].x) {
#pragma unroll
                    for (int ct = 0; ct < CT; ++ct) {
                        const m16_h8 bl = __builtin_bit_cast(m16_h8, fb[kc & 1][ct][1]);
                        acc[r][ct] = __builtin_amdgcn_mfma_f32_32x32x16_f16(xh, bl, acc[r][ct], 0, 0, 0);
                    }
#pragma unroll
                    for (int ct = 0; ct < CT; ++ct) {
                        const m16_h8 bh = __builtin_bit_cast(m16_h8, fb[kc & 1][ct][0]);
                        acc[r][ct] = __builtin_amdgcn_mfma_f32_32x32x16_f16(xl, bh, acc[r][ct], 0, 0, 0);
                    }
                }
            }
        }
#pragma unroll
        for (int i = 0; i < PPT; ++i)
            if (piece_ok(i)) *stage_slot(buf ^ 1, i) = nb[i];
#pragma unroll
        for (int r = 0; r < RW; ++r) {
            pc[r] = pn[r];
#pragma unroll
            for (int k = 0; k < LA; ++k) {
                hh[r][k] = nh[r][k];
                hl[r][k] = nl[r][k];
            }
        }
        dcur = dnext;
        lds_barrier();
    }

    // ---- epilogue: L += acc / (scale of K * scale of X), in FP64 (or the
    // part's partial sum, summed into L by a reduction, when split)
    const double inv = a.kinv / m16_xscale(*a.xmax);
    double2* const dst = a.dsplit > 1 ? a.partial + (int64_t)part * a.size * tri(N) : a.dst;
#pragma unroll
    for (int ct = 0; ct < CT; ++ct) {
        uint32_t tx, ty, tz;
        const int t = target_of(ct * 32 + (lane & 31), tx, ty, tz);
        if (t < 0) continue;
        double* out = (double*)(dst + (int64_t)t * tri(N));
        if (a.dsplit > 1 && w == 0) {
            // the part's slice is written whole: the imaginary parts of the
            // m = 0 coefficients (no real coordinate) as zeros; lanes l and
            // l + 32 hold the same target
            for (int n = lane >> 5; n < N; n += 2) out[2 * tri(n) + 1] = 0.0;
        }
#pragma unroll
        for (int r = 0; r < RW; ++r) {
            const int rt = w + 4 * r;
            if (!kEven && rt >= RT) continue;
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int row = rt * 32 + (v >> 2) * 8 + (lane >> 5) * 4 + (v & 3);
                if (row < NR) {
                    const int o = m16_coef_offset(row);
                    const double val = (double)acc[r][ct][v] * inv;
                    if (a.dsplit > 1) out[o] = val;
                    else out[o] += val;
                }
            }
        }
    }
}

// ===================================================================== row-complete variant
// k_m2l16r: every wave computes all row tiles of its own CTW column tiles, so
// the four waves of a workgroup need the same operator fragments and share
// them through LDS; each wave loads its own source fragments straight into
// registers.  Per direction, the operator is streamed in NSEG segments of SK
// K chunks (segment s holds chunks s, s + NSEG, s + 2 NSEG, ...: heavy low
// chunks spread over all segments).  Segment i + 1 is loaded (global ->
// registers) while segment i is multiplied and written to the other LDS
// buffer afterwards: one barrier per segment, no wave waits for another's
// rows, and each operator fragment read from LDS feeds 2 * CTW or 4 * CTW
// MFMAs.  Skipping is per (direction, row tile), as in k_m2l16: hi*hi below
// the tile's k1, hi*lo + lo*hi below its k3, under scalar branches on the
// wave-uniform bounds (at 13 terms 96 instead of 131 MFMAs per direction and
// column tile, averaged over the directions the classes use); the sources
// are loaded below the largest bounds of the direction.
template <int N>
struct M16R {
    static constexpr int RT = M16<N>::RT, KC = M16<N>::KC;
    static constexpr int SK = 3;                          // K chunks per segment
    static constexpr int NSEG = (KC + SK - 1) / SK;       // segments per direction
    static constexpr int SF = SK * RT * 2;                // fragments per segment (hi, lo)
    static constexpr int SFW = (SF + 3) / 4;              // fragments staged per wave
};

// Column tiles per wave: two while the accumulators (RT * CTW * 16 floats
// per lane) leave room for the operand pipeline, else one.
constexpr int m16r_ctw(int N) { return m16r_max_rt2 >= (N * N + 31) / 32 ? 2 : 1; }

template <int N, int CTW, bool M2M = false>
__global__ void __launch_bounds__(256, CTW == 1 ? 2 : 1) k_m2l16r(M2L16Args a, const uint4* __restrict__ kmat, const uint4* __restrict__ xp,
                                                const int* __restrict__ class_dirs, const int* __restrict__ dir_vec,
                                                const int2* __restrict__ kpre)
{
    using S = M16<N>;
    using R = M16R<N>;
    constexpr int RT = S::RT, KC = S::KC, NR = S::NR, PIECES = S::PIECES;
    constexpr int SK = R::SK, NSEG = R::NSEG, SF = R::SF, SFW = R::SFW;
    constexpr int TC = 4 * 32 * CTW;  // targets (columns) per workgroup
    // fragments move as native 16-byte vectors (plain loads and stores, no
    // aggregate copies, so the staging arrays stay in registers)
    __shared__ m16_u4 as[2][SF][64];
    const m16_u4* __restrict__ kv = (const m16_u4*)kmat;
    const m16_u4* __restrict__ xv = (const m16_u4*)xp;

    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int blk = xcd_local_block(blockIdx.x, gridDim.x);
    int tile, part, c;
    int pbeg = 0, pend = a.plist ? a.plist_count : a.num_parents;  // this block's parent-list range
    if (a.csorted) {
        // class-sorted list: every column's parent has a class-c child;
        // tile-major blocks with the 8 classes adjacent (see k_rtr), and
        // the direction parts of a tile next to each other
        c = blk & 7;
        const int rest = blk >> 3;
        tile = rest / a.dsplit;
        part = rest - tile * a.dsplit;
        if (tile >= a.ctile[c + 1] - a.ctile[c]) return;  // the whole block
        pbeg = a.coff[c];
        pend = a.coff[c + 1];
    } else {
        tile = blk / (8 * a.dsplit);
        part = (blk - tile * 8 * a.dsplit) >> 3;
        c = blk & 7;
    }
    // M2L: the class's 189 offsets; M2M: the 8 child classes
    const int ndir = M2M ? 8 : kDirsPerClass;
    const int d_beg = part * ndir / a.dsplit, d_end = (part + 1) * ndir / a.dsplit;
    const int side = 1 << a.level;

    // targets of this lane: the class-c child of parent tile * TC + column,
    // column 32 * (CTW * w + ct) + lane % 32
    int tgt[CTW], pidx[CTW];
    uint32_t tx[CTW], ty[CTW], tz[CTW];
#pragma unroll
    for (int ct = 0; ct < CTW; ++ct) {
        const int Pi = pbeg + tile * TC + 32 * (CTW * w + ct) + (lane & 31);
        pidx[ct] = Pi;
        int t = -1;
        if (Pi < pend) {
            const int P = a.plist ? a.plist[Pi] : Pi;
            const int cb = a.child_begin[P], ce = a.child_begin[P + 1];
            for (int i = cb; i < ce; ++i)
                if ((int)(a.child_keys[i] & 7u) == c) t = i;
        }
        tgt[ct] = t;
        tx[ct] = ty[ct] = tz[ct] = 0;
        if (t >= 0) {
            const uint64_t k = a.child_keys[t];
            tx[ct] = compact3(k >> 2);
            ty[ct] = compact3(k >> 1);
            tz[ct] = compact3(k);
        }
    }
    auto dir_of = [&](int d) { return M2M ? d : class_dirs[c * kDirsPerClass + d]; };
    // source of column tile ct in direction dir, as the offset of its class
    // plane entry: class * PIECES * (num_parents + 1) + parent slot (the zero
    // slot num_parents of class 0 when the source or its parent is absent)
    const int64_t np1 = M2M ? a.src_np1 : (int64_t)a.num_parents + 1;
    const int64_t zslot = np1 - 1;  // the zero entry
    const int64_t xpc = a.xpieces > 0 ? a.xpieces : PIECES;  // pieces per packed box
    auto box_of = [&](int dir, int ct) -> int64_t {
        int64_t base = zslot;
        if constexpr (M2M) {
            // the target's child of class dir, packed by its parent (the target)
            if (tgt[ct] >= 0) base = (int64_t)dir * xpc * np1 + tgt[ct];
            return base;
        }
        if (tgt[ct] >= 0) {
            const int sx = (int)tx[ct] + dir_vec[4 * dir], sy = (int)ty[ct] + dir_vec[4 * dir + 1],
                      sz = (int)tz[ct] + dir_vec[4 * dir + 2];
            if (sx >= 0 && sy >= 0 && sz >= 0 && sx < side && sy < side && sz < side) {
                const int p = find_box(morton3(sx >> 1, sy >> 1, sz >> 1), a.plut, a.pkeys, a.num_parents);
                if (p >= 0) base = (int64_t)(((sx & 1) << 2) | ((sy & 1) << 1) | (sz & 1)) * xpc * np1 + p;
            }
        }
        return base;
    };
    // bounds of direction dir (wave-uniform): K chunks below r1[rt] need
    // hi*hi for row tile rt, below r3[rt] all three products; bk1 and bk3
    // are their largest values over the row tiles
    auto prefixes = [&](int dir, int (&r3)[RT], int (&r1)[RT], int& bk3, int& bk1) {
        bk3 = 0;
        bk1 = 0;
#pragma unroll
        for (int rt = 0; rt < RT; ++rt) {
            const int2 v = kpre[dir * RT + rt];
            r3[rt] = __builtin_amdgcn_readfirstlane(v.x);
            r1[rt] = __builtin_amdgcn_readfirstlane(v.y);
            bk3 = max(bk3, r3[rt]);
            bk1 = max(bk1, r1[rt]);
        }
    };
    // operator fragments of segment s of direction dir: wave w loads slot
    // fragments w, w + 4, ...  (slot fragment f = (kci * RT + rt) * 2 + hl);
    // blocks the precision does not need come from one cache-resident block
    struct Stage {
        m16_u4 v[SFW];
    };
    // With an even row-tile count, wave w's fragments w + 4 j are (kci, rt,
    // hl) = (2 j / RT, 2 j % RT + w / 2, w % 2): compile-time in j plus two
    // per-wave constants, so the per-segment address work is an add and a
    // bound test per fragment.
    constexpr bool kEvenRT = (RT % 2) == 0;
    const int w_hl = w & 1, w_rt = w >> 1;
    auto stage_a = [&](int dir, int s, int k3d, int k1d) -> Stage {
        Stage stg;
        if constexpr (kEvenRT) {
            const int kb = w_hl ? k3d : k1d;
            const int64_t base = ((int64_t)(dir * RT + w_rt) * KC + s) * 2 + w_hl;
#pragma unroll
            for (int j = 0; j < SFW; ++j) {
                const int kci = (2 * j) / RT, rt0 = (2 * j) % RT;
                const int kc = s + NSEG * kci;
                const bool need = kc < KC && kc < kb;
                const int64_t b = need ? base + (int64_t)(rt0 * KC + NSEG * kci) * 2 : 0;
                stg.v[j] = kv[b * 64 + lane];
            }
        } else {
#pragma unroll
            for (int j = 0; j < SFW; ++j) {
                const int f = w + 4 * j;
                if ((SF & 3) == 0 || f < SF) {
                    const int kci = f / (2 * RT), rt = (f >> 1) % RT, hl = f & 1;
                    const int kc = s + NSEG * kci;
                    const bool need = kc < KC && kc < (hl ? k3d : k1d);
                    const int64_t b = need ? ((int64_t)(dir * RT + rt) * KC + kc) * 2 + hl : 0;
                    stg.v[j] = kv[b * 64 + lane];
                }
            }
        }
        return stg;
    };
    auto store_a = [&](m16_u4* slot, const Stage& stg) {
#pragma unroll
        for (int j = 0; j < SFW; ++j) {
            const int f = w + 4 * j;
            if ((SF & 3) == 0 || f < SF) slot[f * 64 + lane] = stg.v[j];
        }
    };
    // source fragments of chunk kc (both column tiles): hi below the
    // direction's bound k1 (else the zero box), lo only below its bound k3
    // (the lo products are not issued beyond it)
    auto load_b = [&](const int64_t* box, int kc, int bk3, int bk1, m16_u4 (&b)[CTW][2]) {
        const int64_t ph = (int64_t)(kc * 4 + (lane >> 5)) * np1;
#pragma unroll
        for (int ct = 0; ct < CTW; ++ct) {
            const int64_t bh = kc < bk1 ? box[ct] : zslot;
            b[ct][0] = xv[bh + ph];
        }
        if (kc < bk3) {
#pragma unroll
            for (int ct = 0; ct < CTW; ++ct) b[ct][1] = xv[box[ct] + ph + 2 * np1];
        }
    };
    // steps e = s * SK + kci of a direction in execution order (chunk
    // kc = s + NSEG * kci, absent when kc >= KC).  The ring of source
    // fragments has RB slots (loads LB = RB - 1 steps ahead); RB divides
    // NSTEP, so the ring keeps its slots across directions.
    constexpr int NSTEP = NSEG * SK;
    constexpr int RB = NSTEP % 3 == 0 ? 3 : (NSTEP % 4 == 0 ? 4 : 2);
    constexpr int LB = RB - 1;
    static_assert(NSTEP % RB == 0, "source ring");
    auto kc_of = [](int e) { return e / SK + NSEG * (e % SK); };

    // accumulators start as the product of zero operands, so they are
    // born in (and stay in) accumulation registers across the direction loop
    m16_acc acc[RT][CTW];
    {
        const m16_h8 z8 = {};
        const m16_acc z16 = {};
#pragma unroll
        for (int r = 0; r < RT; ++r)
#pragma unroll
            for (int ct = 0; ct < CTW; ++ct) acc[r][ct] = __builtin_amdgcn_mfma_f32_32x32x16_f16(z8, z8, z16, 0, 0, 0);
    }

    // Row tile 0 (degrees n <= 5, which carry most of the potential inside
    // the target box) restarts from zero every direction and its per-
    // direction sums are added into tot: the FP32 rounding of each MFMA then
    // scales with one direction's contribution, not with the running total
    // over 189 directions (about 5x less error where one source box
    // dominates, as in clustered inputs).
    m16_acc tot[CTW];
#pragma unroll
    for (int ct = 0; ct < CTW; ++ct) tot[ct] = m16_acc{};

    int dcur = dir_of(d_beg);
    int bk3, bk1, k3r[RT], k1r[RT];
    prefixes(dcur, k3r, k1r, bk3, bk1);
    int64_t bx[CTW];
#pragma unroll
    for (int ct = 0; ct < CTW; ++ct) bx[ct] = box_of(dcur, ct);
    Stage stg;
    m16_u4 br[RB][CTW][2];  // source fragments of steps e .. e + LB
    m16_u4* cur = &as[0][0][0];
    m16_u4* nxt = &as[1][0][0];
    stg = stage_a(dcur, 0, bk3, bk1);
#pragma unroll
    for (int e = 0; e < LB; ++e)
        if (kc_of(e) < KC) load_b(bx, kc_of(e), bk3, bk1, br[e]);
    store_a(cur, stg);
    lds_barrier();
    for (int d = d_beg; d < d_end; ++d) {
        const int dnext = dir_of(min(d + 1, d_end - 1));  // the last step re-stages its own data (never read)
        int64_t bxn[CTW];
#pragma unroll
        for (int ct = 0; ct < CTW; ++ct) bxn[ct] = box_of(dnext, ct);
        int bk3n, bk1n, k3n[RT], k1n[RT];
        prefixes(dnext, k3n, k1n, bk3n, bk1n);
        // sources of step e + LB (into ring slot (e + LB) % RB)
        auto issue_b = [&](int e) {
            const int e2 = e + LB;
            if (e2 < NSTEP) {
                if (kc_of(e2) < KC) load_b(bx, kc_of(e2), bk3, bk1, br[e2 % RB]);
            } else if (kc_of(e2 - NSTEP) < KC) {
                load_b(bxn, kc_of(e2 - NSTEP), bk3n, bk1n, br[e2 % RB]);
            }
        };
#pragma unroll
        for (int s = 0; s < NSEG; ++s) {
            // The first source load of the segment goes out before the
            // operator staging: vector loads complete in issue order, so the
            // wait for it (two steps later) does not also wait for the
            // operator fragments, which come from further away.
            issue_b(s * SK);
            // the next segment's operator fragments, in flight under this
            // segment's MFMAs (written to the other LDS buffer at its end)
            if (s + 1 < NSEG) stg = stage_a(dcur, s + 1, bk3, bk1);
            else stg = stage_a(dnext, 0, bk3n, bk1n);
            // Per chunk and row tile: hi*hi below the tile's bound k1, the
            // two lo products below its bound k3 (scalar branches: the
            // bounds are wave-uniform).  The next chunk's hi fragments and
            // this chunk's lo fragments are read from LDS while the hi*hi
            // products run; sources arrive two steps ahead.
            m16_u4 fh[2][RT];
#pragma unroll
            for (int rt = 0; rt < RT; ++rt) fh[0][rt] = cur[(rt * 2) * 64 + lane];
#pragma unroll
            for (int kci = 0; kci < SK; ++kci) {
                const int e = s * SK + kci, kc = kc_of(e);
                if (kci > 0) issue_b(e);
                if (kc < KC) {
                    const m16_u4(&b)[CTW][2] = br[e % RB];
                    const int h = kci & 1;
                    m16_u4 fl[RT];
#pragma unroll
                    for (int rt = 0; rt < RT; ++rt)
                        if (kc < k3r[rt]) fl[rt] = cur[((kci * RT + rt) * 2 + 1) * 64 + lane];
                    if (kci + 1 < SK && kc + NSEG < KC) {
#pragma unroll
                        for (int rt = 0; rt < RT; ++rt) fh[h ^ 1][rt] = cur[(((kci + 1) * RT + rt) * 2) * 64 + lane];
                    }
#pragma unroll
                    for (int rt = 0; rt < RT; ++rt)
                        if (kc < k1r[rt]) {
#pragma unroll
                            for (int ct = 0; ct < CTW; ++ct)
                                acc[rt][ct] = __builtin_amdgcn_mfma_f32_32x32x16_f16(
                                    __builtin_bit_cast(m16_h8, fh[h][rt]), __builtin_bit_cast(m16_h8, b[ct][0]),
                                    acc[rt][ct], 0, 0, 0);
                        }
                    if (kc < bk3) {
#pragma unroll
                        for (int rt = 0; rt < RT; ++rt)
                            if (kc < k3r[rt])
#pragma unroll
                            for (int ct = 0; ct < CTW; ++ct) {
                                acc[rt][ct] = __builtin_amdgcn_mfma_f32_32x32x16_f16(
                                    __builtin_bit_cast(m16_h8, fh[h][rt]), __builtin_bit_cast(m16_h8, b[ct][1]),
                                    acc[rt][ct], 0, 0, 0);
                                acc[rt][ct] = __builtin_amdgcn_mfma_f32_32x32x16_f16(
                                    __builtin_bit_cast(m16_h8, fl[rt]), __builtin_bit_cast(m16_h8, b[ct][0]),
                                    acc[rt][ct], 0, 0, 0);
                            }
                    }
                }
                // keep each chunk's reads and products together (bounded
                // register lifetimes)
                __builtin_amdgcn_sched_barrier(0);
            }
            store_a(nxt, stg);
            m16_u4* t = cur;
            cur = nxt;
            nxt = t;
            lds_barrier();
        }
        dcur = dnext;
        bk3 = bk3n;
        bk1 = bk1n;
#pragma unroll
        for (int rt = 0; rt < RT; ++rt) {
            k3r[rt] = k3n[rt];
            k1r[rt] = k1n[rt];
        }
#pragma unroll
        for (int ct = 0; ct < CTW; ++ct) {
            bx[ct] = bxn[ct];
            tot[ct] += acc[0][ct];
            acc[0][ct] = m16_acc{};
        }
    }

    // ---- epilogue: L += acc / (scale of K * scale of X) in FP64, or the
    // part's partial sum when the directions are split (every double of the
    // part's slice is written: the imaginary parts of the m = 0 coefficients,
    // which have no real coordinate, as zeros)
    const double inv = a.kinv / m16_xscale(*a.xmax);
    // split directions: the part's slice of the level (indexed by target),
    // or with a class-sorted list the part's slice of the list (indexed by
    // list position)
    const int64_t part_size = a.csorted ? a.plist_count : a.size;
    double2* const dst = a.dsplit > 1 ? a.partial + (int64_t)part * part_size * tri(N) : a.dst;
    const int64_t SX = (a.dsplit == 1 && a.stride > 0) ? a.stride : tri(N);
    // written whole: a direction part's partial sum, or an M2M result
    const bool assign = a.dsplit > 1 || M2M;
    double mx = 0.0;  // M2M: largest |coefficient| written
#pragma unroll
    for (int ct = 0; ct < CTW; ++ct) {
        if (tgt[ct] < 0) continue;
        double* out = (double*)(dst + (int64_t)(a.dsplit > 1 && a.csorted ? pidx[ct] : tgt[ct]) * SX);
        if (assign) {
            // lanes l and l + 32 hold the same target
            for (int n = lane >> 5; n < N; n += 2) out[2 * tri(n) + 1] = 0.0;
        }
#pragma unroll
        for (int rt = 0; rt < RT; ++rt) {
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int row = rt * 32 + (v >> 2) * 8 + (lane >> 5) * 4 + (v & 3);
                if (row < NR) {
                    const int o = m16_coef_offset(row);
                    const double val = (double)(rt == 0 ? tot[ct][v] : acc[rt][ct][v]) * inv;
                    if (assign) out[o] = val;
                    else out[o] += val;
                    mx = fmax(mx, fabs(val));
                }
            }
        }
    }
    if constexpr (M2M)
        if (a.absmax) wave_absmax(mx, a.absmax);
}

// Targets per workgroup = 32 * m16_ct(N): two column tiles while the
// double-buffered source stage (64 KiB per tile at N = 16) fits the LDS.
constexpr int m16_ct(int N) { return N <= 16 ? 2 : 1; }

template <int N>
hipError_t do_m2l16_pack(const M2L16Args& a, hipStream_t st)
{
    if constexpr (m16_rowfull(N)) {
        const int64_t packs = ((int64_t)a.num_parents + 1) * M16<N>::KC * 2;
        hipLaunchKernelGGL((k_m2l16_pack_cls<N>), dim3((unsigned)((packs + 255) / 256)), dim3(256), 0, st, a.outer,
                           a.num_parents, a.child_begin, a.child_keys, a.xmax, a.xp,
                           (int64_t)(a.stride > 0 ? a.stride : tri(N)));
    } else {
        const int64_t packs = (int64_t)(a.size + 1) * M16<N>::KC * 2;
        hipLaunchKernelGGL((k_m2l16_pack<N>), dim3((unsigned)((packs + 255) / 256)), dim3(256), 0, st, a.outer,
                           a.size, a.xmax, a.xp);
    }
    return hipGetLastError();
}

template <int N>
hipError_t do_m2l16(const M2L16Args& a, hipStream_t st)
{
    static_assert(m16_targets(N) == (m16_rowfull(N) ? 128 * m16r_ctw(N) : 32 * m16_ct(N)), "tile width");
    const int targets = a.plist ? a.plist_count : a.num_parents;
    const int tiles = (targets + m16_targets(N) - 1) / m16_targets(N);
    if (a.csorted) {
        if (!a.plist || a.dsplit < 1 || a.dsplit > kDirsPerClass || !m16_rowfull(N) || a.coff[8] != a.plist_count ||
            (a.dsplit > 1 && !a.partial))
            return hipErrorInvalidValue;
        for (int c = 0; c < 8; ++c)
            if (a.ctile[c + 1] - a.ctile[c] != (a.coff[c + 1] - a.coff[c] + m16_targets(N) - 1) / m16_targets(N))
                return hipErrorInvalidValue;
        int mt = 0;
        for (int c = 0; c < 8; ++c) mt = mt > a.ctile[c + 1] - a.ctile[c] ? mt : a.ctile[c + 1] - a.ctile[c];
        if constexpr (m16_rowfull(N))
            if (mt > 0)
                hipLaunchKernelGGL((k_m2l16r<N, m16r_ctw(N)>), dim3(8 * mt * a.dsplit), dim3(256), 0, st, a, a.kmat,
                                   (const uint4*)a.xp, a.class_dirs, a.dir_vec, a.kpre);
        return hipGetLastError();
    }
    if (tiles != a.tiles || a.dsplit < 1 || a.dsplit > kDirsPerClass) return hipErrorInvalidValue;
    if (a.m2m && (!m16_rowfull(N) || a.dsplit != 1 || a.plist || a.stride || a.src_np1 < 1)) return hipErrorInvalidValue;
    if (a.plist && (a.dsplit != 1 || !m16_rowfull(N))) return hipErrorInvalidValue;
    if (tiles > 0) {
        if constexpr (m16_rowfull(N)) {
            if (a.m2m)
                hipLaunchKernelGGL((k_m2l16r<N, m16r_ctw(N), true>), dim3(tiles * 8), dim3(256), 0, st, a, a.kmat,
                                   (const uint4*)a.xp, a.class_dirs, a.dir_vec, a.kpre);
            else
                hipLaunchKernelGGL((k_m2l16r<N, m16r_ctw(N)>), dim3(tiles * 8 * a.dsplit), dim3(256), 0, st, a, a.kmat,
                                   (const uint4*)a.xp, a.class_dirs, a.dir_vec, a.kpre);
        }
        else
            hipLaunchKernelGGL((k_m2l16<N, m16_ct(N)>), dim3(tiles * 8 * a.dsplit), dim3(256), 0, st, a, a.kmat,
                               (const uint4*)a.xp, a.class_dirs, a.dir_vec, a.kpre);
    }
    return hipGetLastError();
}

}  // namespace lfmm
