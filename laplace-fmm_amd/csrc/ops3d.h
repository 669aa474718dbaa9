// ops3d.h -- argument blocks and launchers of the 3D operator kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace lfmm {

struct P2MArgs {
    const double4* src;        // sorted (x, y, z, q)
    const int32_t* leaf_begin; // [nleaf + 1]
    const uint64_t* leaf_key;
    int nleaf;
    int level;
    const double* geom;        // {shift x, y, z, scale} (device)
    const double* knm;         // K_nm table, tri(N+1)
    double2* outer;            // normalised multipoles of the leaves
    uint32_t* absmax = nullptr;   // if set: max-combined largest |coefficient| (float bits)
    // adaptive octree, P2M: unit i is leaf i of the global leaf list, written
    // to outer + leaf_gid[i] * NX; leaf_lvl[i] its level (absmax + level)
    const int32_t* leaf_gid = nullptr;
    const int32_t* leaf_lvl = nullptr;
    // adaptive octree, P2L (launch_p2l): unit i is box g2 + unit_box[i] of
    // `level` (key leaf_key[g2 + t], keys by gid); its sources are the
    // particle ranges x_rng[x_off[t] .. x_off[t + 1]); the result is added to
    // outer + (g2 + t) * NX (the locals)
    const int32_t* unit_box = nullptr;
    const uint32_t* x_off = nullptr;
    const int2* x_rng = nullptr;
    int g2 = 0;
};

struct RtrTables {
    const double2* rot;        // [theta][rot_size(N)] (P, Q)
    const double2* ephi;       // [dir][N]
    const double* trans;       // [kNumTrans][tr_size(N)]
    const int* dir_theta;
    const int* dir_trans;
    const int* dir_vec;        // [dir][4]
    const int* class_dirs;     // [8][189]
};

// Targets per workgroup of the rotate-translate-rotate kernel: 64 (one per
// lane) while the FP64 op state of 64 targets fits the 160 KiB LDS, fewer
// lanes above that (N = 18).
constexpr int rtr_tile(int N)
{
    return (N * (N + 1) / 2) * 64 * 16 <= 160 * 1024 ? 64 : (160 * 1024) / ((N * (N + 1) / 2) * 16) / 8 * 8;
}

struct RtrLevel {
    int level;                 // M2L: target level; M2M: parent level
    int num_parents;           // M2L: boxes at level-1; M2M: boxes at level
    const int32_t* child_begin;   // child ranges of those parents
    const uint64_t* child_keys;   // keys of the child level
    const int32_t* lut;        // M2L source lookup at `level` (dense) or null
    const uint64_t* keys;      // keys at `level` (binary search fallback)
    int size;                  // boxes at `level`
    const double2* src;        // M2L: outer[level]; M2M: outer[level+1]
    const double2* parent_local;  // M2L: inner[level-1] (null: no L2L)
    double2* dst;              // M2L: inner[level]; M2M: outer[level]
    // M2L / L2L: target tiles (ceil(num_parents / rtr_tile(N))) and the number
    // of direction parts; with dsplit > 1 part p writes partial + p*size*NX
    // and a reduction forms dst (small levels: more workgroups than tiles).
    int tiles = 0;
    int dsplit = 1;
    double2* partial = nullptr;
    uint32_t* absmax = nullptr;   // M2M: if set, max-combined largest |coefficient| written (float bits)
    // per-box orders: the parents are plist[0 .. num_parents) (indices at
    // level-1) and every expansion array has `stride` coefficients per box
    // (0: tri(N)); requires dsplit == 1
    const int32_t* plist = nullptr;
    int stride = 0;
    // class-sorted parent list (M2L / L2L with plist): the parents with a
    // class-c child are plist[coff[c] .. coff[c+1]), and blocks
    // ctile[c] .. ctile[c+1] serve them (tiles of rtr_tile(N) parents)
    int csorted = 0;
    int coff[9] = {};
    int ctile[9] = {};
};

struct LeafArgs {
    const double4* src;
    const int32_t* leaf_of;    // leaf index of each sorted particle
    const uint64_t* leaf_key;
    const int32_t* leaf_begin;
    int nleaf;
    int level;
    const double* geom;
    const int32_t* lut;        // dense key -> leaf table or null
    const double2* local;      // normalised locals of the leaves
    const double* knm;
    const uint32_t* perm;      // sorted -> input index
    const int2* nbr;           // [nleaf][27] neighbour particle ranges (launch_leaf_ranges)
    int64_t n;
    int64_t n_out;             // outputs for input indices < n_out only (targets first)
    int exclude_self;
    double* pot;
    double* field;
    double4* far = nullptr;    // with locals: per sorted particle (phi, field) of L2P, written by
                               // k_l2p and read by the pair kernel (split launch)
    // adaptive octree: leaves on several levels.  leaf_key / leaf_begin /
    // leaf_of then describe the global leaf list; leaf_lvl its levels; local
    // the locals of all boxes by gid (leaf_gid); nbr / nbr_off the U lists as
    // CSR of particle ranges; w_off / w_gid the W lists, whose multipoles
    // (outer_all, keys_all by gid, level offsets loff) act on the leaf's
    // particles (M2P, in k_l2p)
    const int32_t* leaf_lvl = nullptr;
    const int32_t* leaf_gid = nullptr;
    const uint32_t* nbr_off = nullptr;
    const uint32_t* w_off = nullptr;
    const int32_t* w_gid = nullptr;
    const uint64_t* keys_all = nullptr;
    const double2* outer_all = nullptr;
    const int2* box_range = nullptr;   // particle range of every box (gid): W boxes of at most
    int direct_max = 0;                // direct_max particles are in the near lists, not M2P
    int depth = 0;
    int loff[17] = {};
    int local_min = 2;                 // leaves below this level have no local expansion
    // forest of independent systems: per leaf, its system's length scale s
    // (forest length per system length).  The pair clamp is then 1e-4 s, and
    // the potential and field leave scaled by s and s^2 (the system's units).
    const double* leaf_scale = nullptr;
};

// Matrix-core M2L of one level (m2l16_kernels.h).
struct M2L16Args {
    const uint4* kmat;         // split-fp16 operator fragments [dir][RT][KC][2][64]
    const int2* kpre;          // [dir][RT] (k3, k1): K chunks needing all products / hi*hi
    const double2* outer;      // normalised multipoles of the level
    uint4* xp;                 // packed sources: k_m2l16 [size + 1][KC][2][2] (box `size` is zero);
                               // k_m2l16r [8 classes][KC][2][2][num_parents + 1] (slot num_parents is zero)
    const uint32_t* xmax;      // largest |coefficient| of the level (float bits)
    const int* class_dirs;
    const int* dir_vec;
    int level;                 // target level
    int num_parents;           // boxes at level - 1
    const int32_t* child_begin;
    const uint64_t* child_keys;
    const int32_t* lut;        // dense key -> box table at `level`, or null
    const uint64_t* keys;
    const int32_t* plut;       // the same at level - 1 (parents), or null
    const uint64_t* pkeys;
    int size;                  // boxes at `level`
    double kinv;               // 1 / scale of the packed operators
    double2* dst;              // locals of the level (+= M2L)
    int tiles = 0;             // target tiles of 32 * m16 column tiles
    int dsplit = 1;            // direction parts (partial + p*size*tri(N) when > 1)
    double2* partial = nullptr;
    // per-box orders: target parents plist[0 .. tiles' parents) (indices at
    // level-1; the packing still covers all num_parents) and `stride`
    // coefficients per box in outer and dst (0: tri(N)); needs dsplit == 1
    const int32_t* plist = nullptr;
    int plist_count = 0;
    int stride = 0;
    // class-sorted plist, as RtrLevel (tiles of m16_targets(N) parents);
    // with dsplit > 1 part p of list entry i goes to
    // partial + (p * plist_count + i) * tri(N), added into dst by
    // launch_sum_list_parts
    int csorted = 0;
    int coff[9] = {};
    int ctile[9] = {};
    // k_m2l16r: 16-byte pieces per packed source box when xp was packed at a
    // larger order (0: this kernel's own).  The first t^2 real coordinates of
    // an expansion are those of its order-t truncation, and the operator's
    // columns beyond them are zero, so one packing serves every order up to
    // the packed one.
    int xpieces = 0;
    // M2M mode (k_m2l16r only): the targets are the boxes of `level` (the
    // class-c children of the num_parents boxes above them, as for M2L), the
    // 8 "directions" are the M2M operators from their children of class
    // 0..7 (kmat / kpre / kinv of the M2M pack), and xp is the class-major
    // packing of the children's level by target: entry (class, piece, t) at
    // xp[(class * pieces + piece) * src_np1 + t], slot src_np1 - 1 zero.
    // The result is written (not added) to dst, and its largest |coefficient|
    // goes to *absmax (float bits).
    int m2m = 0;
    int64_t src_np1 = 0;
    uint32_t* absmax = nullptr;
};

// Matrix-core M2L kernel per expansion size: the row-complete k_m2l16r
// (operator shared through LDS, 128 or 256 targets per workgroup) from
// LFMM_M16R_MIN_N terms on, below it k_m2l16 (sources shared through LDS, 64
// or 32 targets per workgroup).
#ifndef LFMM_M16R_MIN_N
#define LFMM_M16R_MIN_N 1
#endif
constexpr bool m16_rowfull(int N) { return N >= LFMM_M16R_MIN_N; }
// k_m2l16r keeps two column tiles per wave up to this many row tiles of 32
constexpr int m16r_max_rt2 = 6;
constexpr int m16_targets(int N)
{
    return m16_rowfull(N) ? ((N * N + 31) / 32 <= m16r_max_rt2 ? 256 : 128) : (N <= 16 ? 64 : 32);
}

hipError_t launch_p2m(int N, const P2MArgs& a, hipStream_t st);
// P2L of one level's X lists (adaptive octree; see P2MArgs)
hipError_t launch_p2l(int N, const P2MArgs& a, hipStream_t st);
hipError_t launch_rtr(int N, int kind, const RtrTables& T, const RtrLevel& L, hipStream_t st);
// L2P alone into a.far (k_l2p), for the pair kernels below
hipError_t launch_l2p(int N, const LeafArgs& a, hipStream_t st);
// Near field with several targets per lane (p2p.hip), starting from a.far when
// set.  ws: p2p_ws_elems(n, nleaf) int32; scan_ws: scan_ws_elems(nleaf).
int64_t p2p_ws_elems(int64_t n, int64_t nleaf);
hipError_t launch_p2p(const LeafArgs& a, int32_t* ws, uint32_t* scan_ws, hipStream_t st);
// nbr[leaf][27] = particle ranges of the 27 neighbour leaves (tree.hip)
hipError_t launch_leaf_ranges(const uint64_t* leaf_key, const int32_t* leaf_begin, int nleaf, int level,
                              const int32_t* lut, int2* nbr, hipStream_t st);
hipError_t launch_m2l16_pack(int N, const M2L16Args& a, hipStream_t st);
hipError_t launch_m2l16(int N, const M2L16Args& a, hipStream_t st);
// dst[i] = (add ? dst[i] : 0) + sum_p partial[p * count + i]   (m2l16.hip)
hipError_t launch_sum_parts(const double2* partial, int parts, int64_t count, double2* dst, bool add, hipStream_t st);
// the direction parts of a class-sorted list (M2L16Args::csorted,
// dsplit > 1), added into the listed targets' locals (m2l16.hip)
hipError_t launch_sum_list_parts(int N, const M2L16Args& a, hipStream_t st);


}  // namespace lfmm
