// tree.h -- host entry points of the GPU octree build (tree.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstdint>
#include <vector>

namespace lfmm {

// dim = 2 (quadtree, x << 1 | y keys) or 3 (octree).
// geom[4] = {shift_x, shift_y, shift_z, scale}; part needs 6 * 1024 doubles.
hipError_t extent(int dim, const double* pos, int64_t stride, int64_t n, double* part, int* nonfinite, double* geom,
                  hipStream_t st);
hipError_t keys(int dim, const double* pos, int64_t stride, int64_t n, int bits, const double* geom, uint64_t* k,
                uint32_t* v, hipStream_t st);
// counts_ws: sum_{l<=max_level} 2^(dim l) uint32; maxv_ws: max_level+1 uint32.  Synchronises.
// *level_out = the smallest level <= max_level whose fullest box holds at
// most leaf_max particles, or max_level + 1 if there is none.
hipError_t occupancy_depth(int dim, const double* pos, int64_t stride, int64_t n, const double* geom, int leaf_max,
                           int max_level, uint32_t* counts_ws, uint32_t* maxv_ws, int* level_out, hipStream_t st);
int64_t scan_ws_elems(int64_t n);
hipError_t exclusive_scan(const uint32_t* in, uint32_t* out, int64_t n, uint32_t* ws, hipStream_t st);
int64_t sort_ws_bytes(int64_t n);
// Stable LSD radix sort of (keys, vals) on the low `bits` bits; results land
// in *keys_res / *vals_res (either the inputs or the workspace).
hipError_t radix_sort(uint64_t* keys, uint32_t* vals, int64_t n, int bits, void* ws, hipStream_t st,
                      uint64_t** keys_res, uint32_t** vals_res);
// (x, y, z, q) records from a strided view (q = 0 when q is null; z = 0 in 2D).
hipError_t pack_points(int dim, const double* pos, int64_t ps, const double* q, int64_t qs, int64_t n, double4* out,
                       hipStream_t st);
// Per-leaf source/target partition of the sorted order (q != 0 first).
hipError_t partition_leaves(const int32_t* begin, int64_t nleaf, uint32_t* order, const double* q, int64_t qs,
                            hipStream_t st);
hipError_t gather_sources(int dim, const double* pos, int64_t pstride, const double* q, int64_t qstride,
                          const uint32_t* order, int64_t n, double4* src, hipStream_t st);
// Run-length compaction of sorted keys at shift sh.  Synchronises (reads the count).
hipError_t run_compact(const uint64_t* keys, int64_t m, int sh, uint32_t* flags, uint32_t* excl, uint32_t* scan_ws,
                       uint64_t* ukeys, int32_t* begin, int32_t* owner, int64_t* count_out, hipStream_t st);
hipError_t build_lut(int dim, const uint64_t* ukeys, int64_t count, int level, int32_t* lut, hipStream_t st);
// Exact M2L and near-field pair counts of a built tree (report only); ws
// holds 2 + depth counters, m2l_per_level (optional) depth entries;
// p2p_reads (optional): the particles of every leaf's near list, summed over
// the leaves.  Synchronises.
hipError_t count_ops(int dim, const std::vector<uint64_t*>& keys, const std::vector<int32_t*>& begin,
                     const std::vector<int32_t*>& lut, const std::vector<int64_t>& cnt, unsigned long long* ws,
                     int64_t* m2l_ops, int64_t* p2p_pairs, int64_t* m2l_per_level, hipStream_t st,
                     int64_t* p2p_reads = nullptr);
// p[0..n) = v with a fill kernel (safe inside stream capture)
hipError_t fill_u32(uint32_t* p, int64_t n, uint32_t v, hipStream_t st);
hipError_t max_run(const int32_t* begin, int64_t count, uint32_t* maxv_dev, uint32_t* out_host, hipStream_t st);
// run_compact with the counts on the device: at most `bound` elements,
// *m_dev of them (null: all `bound`); the number of runs goes to *count_dev.
// No host synchronisation.
hipError_t run_compact_dev(const uint64_t* keys, int64_t bound, const uint32_t* m_dev, int sh, uint32_t* flags,
                           uint32_t* excl, uint32_t* scan_ws, uint64_t* ukeys, int32_t* begin, int32_t* owner,
                           uint32_t* count_dev, hipStream_t st);

}  // namespace lfmm
