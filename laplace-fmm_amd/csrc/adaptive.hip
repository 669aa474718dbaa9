// adaptive.hip -- adaptive octree and its interaction lists on the GPU.
//
// Start: the uniform tree of the deepest level any box needs (every box of
// it holds at most leaf_max particles), built by build_tree with parent
// indices.  A uniform box belongs to the adaptive tree iff its parent holds
// more than leaf_max particles (then so does every ancestor); it is a leaf
// iff it holds at most leaf_max particles or sits on the deepest level.  The
// kept boxes of each level are compacted in key order; leaves are listed
// globally in particle (Morton) order, where they tile [0, n).
//
// Lists (definitions in adaptive.h):
//   U  per leaf: the leaf itself, the leaves of its level or coarser that
//      hold one of its 26 neighbour cells (walk up the levels from the cell),
//      and the finer leaves that found it that way (pairs transposed by a
//      stable sort, so every list has a fixed order)
//   X  per box b (levels >= 2): the coarser leaves that hold a neighbour cell
//      of b's parent and do not touch b
//   W  per leaf: the X pairs transposed (stable sort by leaf)
// Every list keeps a deterministic order, so the sums of a solve do not
// depend on timing.
#include "adaptive.h"
#include "api_internal.h"
#include "device_common.h"
#include "tree.h"

namespace lfmm {

namespace {

inline int nblk(int64_t n, int t) { return (int)((n + t - 1) / t); }

__device__ inline int ad_level(const AdLevels& A, int g)
{
    int l = 0;
    while (l + 1 < A.depth && g >= A.off[l + 1]) ++l;
    return l;
}

__device__ inline int ad_find(const AdLevels& A, int l, uint64_t key)
{
    return find_box(key, A.lut[l], A.keys + A.off[l], A.off[l + 1] - A.off[l]);
}

// The leaf (gid) that holds cell (x, y, z) of level l, at level l or above;
// -2 when a box with children holds the cell at level l (finer leaves below
// it); -1 when no particle lies in the cell.
__device__ inline int ad_cover(const AdLevels& A, int l, int x, int y, int z)
{
    for (int lv = l; lv >= 0; --lv) {
        const int sh = l - lv;
        const int j = ad_find(A, lv, morton3((uint32_t)x >> sh, (uint32_t)y >> sh, (uint32_t)z >> sh));
        if (j >= 0) {
            const int g = A.off[lv] + j;
            if (A.box2leaf[g] >= 0) return g;
            return lv == l ? -2 : -1;
        }
    }
    return -1;
}

__global__ void __launch_bounds__(256) k_ad_ubeg(const int32_t* __restrict__ child_begin, int64_t cnt,
                                                 const int32_t* __restrict__ ubeg_next, int32_t n, int32_t* __restrict__ ubeg)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i > cnt) return;
    ubeg[i] = i < cnt ? ubeg_next[child_begin[i]] : n;
}

__global__ void __launch_bounds__(256) k_ad_keep(const int32_t* __restrict__ uparent, int64_t cnt,
                                                 const int32_t* __restrict__ ubeg_parent, int leaf_max, int always,
                                                 uint32_t* __restrict__ keep)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= cnt) return;
    const int p = uparent[i];
    keep[i] = (always || ubeg_parent[p + 1] - ubeg_parent[p] > leaf_max) ? 1u : 0u;
}

__global__ void __launch_bounds__(256) k_ad_fill(int last, int leafable, const uint64_t* __restrict__ ukeys,
                                                 const int32_t* __restrict__ ubeg, const int32_t* __restrict__ uchild,
                                                 int64_t cnt, const uint32_t* __restrict__ keep,
                                                 const uint32_t* __restrict__ kex, const uint32_t* __restrict__ kex_next,
                                                 int leaf_max, int gid0, uint64_t* __restrict__ keys_all,
                                                 int2* __restrict__ range_all, int32_t* __restrict__ child,
                                                 int32_t* __restrict__ box2leaf)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= cnt || !keep[i]) return;
    const int a = (int)kex[i];
    const int g = gid0 + a;
    const int b = ubeg[i], e = ubeg[i + 1];
    keys_all[g] = ukeys[i];
    range_all[g] = make_int2(b, e);
    // a leaf's uniform children are not kept: its child range is empty
    child[a] = last ? 0 : (int32_t)kex_next[uchild[i]];
    box2leaf[g] = (last || (leafable && e - b <= leaf_max)) ? 0 : -1;
}

__global__ void k_ad_set1(int32_t* __restrict__ p, int32_t v) { *p = v; }

// tag[first particle of a leaf] = its gid
__global__ void __launch_bounds__(256) k_ad_leaf_tag(const int2* __restrict__ range, const int32_t* __restrict__ box2leaf,
                                                     int total, int32_t* __restrict__ tag)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= total || box2leaf[g] < 0) return;
    tag[range[g].x] = g;
}

__global__ void __launch_bounds__(256) k_ad_flags(const int32_t* __restrict__ tag, int64_t n, uint32_t* __restrict__ flags)
{
    const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (s < n) flags[s] = tag[s] >= 0 ? 1u : 0u;
}

__global__ void __launch_bounds__(256) k_ad_leaf_list(AdLevels A, const int32_t* __restrict__ tag,
                                                      const uint32_t* __restrict__ ex, int64_t n,
                                                      int32_t* __restrict__ box2leaf, int32_t* __restrict__ leaf_begin,
                                                      int32_t* __restrict__ leaf_gid, uint64_t* __restrict__ leaf_key,
                                                      int32_t* __restrict__ leaf_lvl, int32_t* __restrict__ leaf_of)
{
    const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int g = tag[s];
    const int i = (int)ex[s] + (g >= 0 ? 1 : 0) - 1;  // particle 0 always starts a leaf
    leaf_of[s] = i;
    if (g >= 0) {
        leaf_begin[i] = (int32_t)s;
        leaf_gid[i] = g;
        leaf_key[i] = A.keys[g];
        leaf_lvl[i] = ad_level(A, g);
        box2leaf[g] = i;
    }
    if (s == n - 1) leaf_begin[i + 1] = (int32_t)n;
}

// U lists, first part: per leaf, the 27 cells around it (offset order) mapped
// to the leaf holding each at its level or coarser; a coarser leaf that holds
// several cells is kept at its first cell.  nbr27 = leaf index or -1.
__global__ void __launch_bounds__(256) k_ad_u1(AdLevels A, const int32_t* __restrict__ leaf_gid,
                                               const int32_t* __restrict__ leaf_lvl, int nleaf,
                                               int32_t* __restrict__ nbr27, uint32_t* __restrict__ nu1,
                                               uint32_t* __restrict__ ncoarse)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nleaf) return;
    const int g = leaf_gid[i], l = leaf_lvl[i];
    const uint64_t key = A.keys[g];
    const int x = (int)compact3(key >> 2), y = (int)compact3(key >> 1), z = (int)compact3(key);
    const int side = 1 << l;
    int found[27];
    uint32_t cnt = 0, cc = 0;
    for (int o = 0; o < 27; ++o) {
        const int cx = x + o / 9 - 1, cy = y + (o / 3) % 3 - 1, cz = z + o % 3 - 1;
        int r = -1;
        if (o == 13) {
            r = g;
        } else if (cx >= 0 && cy >= 0 && cz >= 0 && cx < side && cy < side && cz < side) {
            const int c = ad_cover(A, l, cx, cy, cz);
            if (c >= 0) {
                r = c;
                if (c < A.off[l])  // coarser: drop repeats
                    for (int k = 0; k < o; ++k)
                        if (found[k] == c) r = -1;
            }
        }
        found[o] = r;
        nbr27[(int64_t)i * 27 + o] = r >= 0 ? A.box2leaf[r] : -1;
        if (r >= 0) {
            ++cnt;
            if (r < A.off[l]) ++cc;
        }
    }
    nu1[i] = cnt;
    ncoarse[i] = cc;
}

// (coarse leaf, this leaf) pairs for the transposition, in leaf order
__global__ void __launch_bounds__(256) k_ad_u_pairs(const int32_t* __restrict__ nbr27, const int32_t* __restrict__ leaf_lvl,
                                                    int nleaf, const uint32_t* __restrict__ coff,
                                                    uint64_t* __restrict__ pk, uint32_t* __restrict__ pv,
                                                    uint32_t* __restrict__ finecnt)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nleaf) return;
    uint32_t w = coff[i];
    const int li = leaf_lvl[i];
    for (int o = 0; o < 27; ++o) {
        const int j = nbr27[(int64_t)i * 27 + o];
        if (j >= 0 && leaf_lvl[j] < li) {
            pk[w] = (uint64_t)j;
            pv[w] = (uint32_t)i;
            ++w;
            atomicAdd(&finecnt[j], 1u);
        }
    }
}

__global__ void __launch_bounds__(256) k_ad_add(const uint32_t* __restrict__ a, const uint32_t* __restrict__ b, int64_t n,
                                                uint32_t* __restrict__ out)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i <= n) out[i] = i < n ? a[i] + b[i] : 0u;
}

__global__ void __launch_bounds__(256) k_ad_u_fill(const int32_t* __restrict__ nbr27, const int32_t* __restrict__ leaf_begin,
                                                   int nleaf, const uint32_t* __restrict__ uoff,
                                                   const uint32_t* __restrict__ fine_off,
                                                   const uint32_t* __restrict__ finecnt,
                                                   const uint32_t* __restrict__ fine_leaf, int2* __restrict__ urng)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nleaf) return;
    uint32_t w = uoff[i];
    for (int o = 0; o < 27; ++o) {
        const int j = nbr27[(int64_t)i * 27 + o];
        if (j >= 0) urng[w++] = make_int2(leaf_begin[j], leaf_begin[j + 1]);
    }
    const uint32_t fb = fine_off[i], fe = fb + finecnt[i];
    for (uint32_t k = fb; k < fe; ++k) {
        const int f = (int)fine_leaf[k];
        urng[w++] = make_int2(leaf_begin[f], leaf_begin[f + 1]);
    }
}

// X lists: per box of levels >= 2 (t = gid - g2), the leaves coarser than it
// that hold a neighbour cell of its parent and do not touch it.
__global__ void __launch_bounds__(256) k_ad_x(AdLevels A, int g2, int total, int32_t* __restrict__ x27,
                                              uint32_t* __restrict__ xcnt)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total - g2) return;
    const int g = g2 + t;
    const int l = ad_level(A, g);
    const uint64_t key = A.keys[g];
    const int x = (int)compact3(key >> 2), y = (int)compact3(key >> 1), z = (int)compact3(key);
    const int px = x >> 1, py = y >> 1, pz = z >> 1;
    const int pside = 1 << (l - 1);
    int found[27];
    int cnt = 0;
    for (int o = 0; o < 27; ++o) {
        const int cx = px + o / 9 - 1, cy = py + (o / 3) % 3 - 1, cz = pz + o % 3 - 1;
        if (cx < 0 || cy < 0 || cz < 0 || cx >= pside || cy >= pside || cz >= pside) continue;
        const int c = ad_cover(A, l - 1, cx, cy, cz);
        if (c < 0) continue;
        // c: a leaf at level lc <= l - 1; in units of level-l cells it spans
        // [c * s, (c + 1) * s) per axis
        const int lc = ad_level(A, c);
        const uint64_t ck = A.keys[c];
        const int s = 1 << (l - lc);
        const int ax = (int)compact3(ck >> 2) * s, ay = (int)compact3(ck >> 1) * s, az = (int)compact3(ck) * s;
        const bool touch = x <= ax + s && x + 1 >= ax && y <= ay + s && y + 1 >= ay && z <= az + s && z + 1 >= az;
        if (touch) continue;
        bool seen = false;
        for (int k = 0; k < cnt; ++k) seen |= found[k] == c;
        if (!seen) found[cnt++] = c;
    }
    for (int k = 0; k < cnt; ++k) x27[(int64_t)t * 27 + k] = A.box2leaf[found[k]];
    xcnt[t] = (uint32_t)cnt;
}

__global__ void __launch_bounds__(256) k_ad_x_fill(const int32_t* __restrict__ x27, const uint32_t* __restrict__ xcnt,
                                                   const uint32_t* __restrict__ xoff, int nb,
                                                   const int32_t* __restrict__ leaf_begin, int g2,
                                                   int2* __restrict__ xrng, uint64_t* __restrict__ wk,
                                                   uint32_t* __restrict__ wv, uint32_t* __restrict__ wcnt,
                                                   int32_t* __restrict__ xlist, const uint32_t* __restrict__ lex,
                                                   const int32_t* __restrict__ box2leaf)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nb) return;
    const uint32_t c = xcnt[t];
    if (c == 0) return;
    if (box2leaf[g2 + t] < 0) xlist[lex[t]] = t;  // P2L for boxes with children only
    uint32_t w = xoff[t];
    for (uint32_t k = 0; k < c; ++k) {
        const int j = x27[(int64_t)t * 27 + k];
        xrng[w] = make_int2(leaf_begin[j], leaf_begin[j + 1]);
        wk[w] = (uint64_t)j;
        wv[w] = (uint32_t)(g2 + t);
        atomicAdd(&wcnt[j], 1u);
        ++w;
    }
}

// boxes whose X list runs as P2L: those with children (a leaf takes its X
// sources in its near list instead)
__global__ void __launch_bounds__(256) k_ad_p2l_flag(const uint32_t* __restrict__ c, int64_t n,
                                                     const int32_t* __restrict__ box2leaf, int g2,
                                                     uint32_t* __restrict__ f)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) f[i] = (c[i] && box2leaf[g2 + i] < 0) ? 1u : 0u;
}

// Near lists of the pair kernel: the U list, then the leaf's X sources (a
// leaf target: direct sums, cheaper and exact), then the W boxes of at most
// direct_max particles (direct sums instead of M2P).
__device__ inline uint32_t ad_small_w(const int32_t* __restrict__ w_gid, uint32_t b, uint32_t e,
                                      const int2* __restrict__ range, int direct_max)
{
    uint32_t c = 0;
    for (uint32_t k = b; k < e; ++k) {
        const int2 r = range[w_gid[k]];
        c += (r.y - r.x <= direct_max) ? 1u : 0u;
    }
    return c;
}

__global__ void __launch_bounds__(256) k_ad_near_count(const int32_t* __restrict__ leaf_gid, int nleaf,
                                                       const uint32_t* __restrict__ u_off,
                                                       const uint32_t* __restrict__ x_off, int g2,
                                                       const uint32_t* __restrict__ w_off,
                                                       const int32_t* __restrict__ w_gid, const int2* __restrict__ range,
                                                       int direct_max, uint32_t* __restrict__ ncnt)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > nleaf) return;
    if (i == nleaf) {
        ncnt[i] = 0u;
        return;
    }
    const int g = leaf_gid[i];
    uint32_t c = u_off[i + 1] - u_off[i];
    if (x_off && g >= g2) c += x_off[g - g2 + 1] - x_off[g - g2];
    if (w_gid) c += ad_small_w(w_gid, w_off[i], w_off[i + 1], range, direct_max);
    ncnt[i] = c;
}

__global__ void __launch_bounds__(256) k_ad_near_fill(const int32_t* __restrict__ leaf_gid, int nleaf,
                                                      const uint32_t* __restrict__ u_off, const int2* __restrict__ u_rng,
                                                      const uint32_t* __restrict__ x_off, const int2* __restrict__ x_rng,
                                                      int g2, const uint32_t* __restrict__ w_off,
                                                      const int32_t* __restrict__ w_gid, const int2* __restrict__ range,
                                                      int direct_max, const uint32_t* __restrict__ n_off,
                                                      int2* __restrict__ n_rng)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nleaf) return;
    const int g = leaf_gid[i];
    uint32_t w = n_off[i];
    for (uint32_t k = u_off[i]; k < u_off[i + 1]; ++k) n_rng[w++] = u_rng[k];
    if (x_off && g >= g2)
        for (uint32_t k = x_off[g - g2]; k < x_off[g - g2 + 1]; ++k) n_rng[w++] = x_rng[k];
    if (w_gid)
        for (uint32_t k = w_off[i]; k < w_off[i + 1]; ++k) {
            const int2 r = range[w_gid[k]];
            if (r.y - r.x <= direct_max) n_rng[w++] = r;
        }
}

// sum over leaves of (particles) x (particles in its U list)
__global__ void __launch_bounds__(256) k_ad_pairs(const int32_t* __restrict__ leaf_begin, const uint32_t* __restrict__ u_off,
                                                  const int2* __restrict__ u_rng, int64_t nleaf,
                                                  unsigned long long* __restrict__ out)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    unsigned long long c = 0, rd = 0;
    if (i < nleaf) {
        unsigned long long t = 0;
        for (uint32_t k = u_off[i]; k < u_off[i + 1]; ++k) t += (unsigned long long)(u_rng[k].y - u_rng[k].x);
        c = t * (unsigned long long)(leaf_begin[i + 1] - leaf_begin[i]);
        rd = t;
    }
    for (int off = 32; off > 0; off >>= 1) {
        c += (unsigned long long)__shfl_xor((long long)c, off);
        rd += (unsigned long long)__shfl_xor((long long)rd, off);
    }
    if ((threadIdx.x & 63) == 0 && c) atomicAdd(out, c);
    if ((threadIdx.x & 63) == 0 && rd) atomicAdd(out + 1, rd);
}

int bits_for(int64_t v)
{
    int b = 1;
    while (b < 63 && ((int64_t)1 << b) <= v) ++b;
    return b;
}

// A forest leaf's system: its ancestor cell (2a, 2b, 2c) at the system level
// holds system (a side + b) side + c; out = that system's length scale.
__global__ void __launch_bounds__(256) k_forest_leaf_scale(const uint64_t* __restrict__ leaf_key,
                                                           const int32_t* __restrict__ leaf_lvl, int64_t nleaf,
                                                           int level, int side, const double* __restrict__ sys_scale,
                                                           double* __restrict__ out)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= nleaf) return;
    const int lv = leaf_lvl[i];
    const uint64_t k = lv >= level ? leaf_key[i] >> (3 * (lv - level)) : 0;
    const int a = (int)(compact3(k >> 2) >> 1), b = (int)(compact3(k >> 1) >> 1), c = (int)(compact3(k) >> 1);
    out[i] = (lv >= level && a < side && b < side && c < side) ? sys_scale[(a * side + b) * side + c] : 1.0;
}

}  // namespace

}  // namespace lfmm

namespace lfmm::host {

AdLevels Tree::view() const
{
    AdLevels A{};
    A.depth = depth;
    for (int l = 0; l <= depth; ++l) A.off[l] = loff[l];
    A.keys = keys_all;
    for (int l = 0; l < depth; ++l) A.lut[l] = lut[l];
    A.box2leaf = box2leaf;
    A.range = range_all;
    return A;
}

// exclusive scan of n + 1 values whose last input is 0: out[n] = the total
static int scan_total(Arena& A, const uint32_t* in, uint32_t* out, int64_t n, hipStream_t st)
{
    uint32_t* ws = take<uint32_t>(A, scan_ws_elems(n + 1));
    if (!ws) return LFMM_ERR_NOMEM;
    LFMM_HIP(exclusive_scan(in, out, n + 1, ws, st));
    return LFMM_OK;
}

static int read_u32(const uint32_t* p, uint32_t* out, hipStream_t st)
{
    LFMM_HIP(hipMemcpyAsync(out, p, 4, hipMemcpyDeviceToHost, st));
    LFMM_HIP(hipStreamSynchronize(st));
    return LFMM_OK;
}

int build_adaptive(lfmm_ctx* ctx, int leaf_max, Tree& T, int min_leaf_level)
{
    hipStream_t st = ctx->stream;
    Arena& A = ctx->arena;
    const int depth = T.depth;
    const int64_t n = T.n;
    if (depth > kAdMaxLevels || leaf_max < 1) return LFMM_ERR_INVALID;
    const int last = depth - 1;
    if ((int)T.uparent.size() != depth) return LFMM_ERR_INVALID;

    // ---- particle begins of every uniform box (bottom-up)
    std::vector<int32_t*> ubeg(depth, nullptr);
    ubeg[last] = T.lbegin[last];
    for (int l = last - 1; l >= 0; --l) {
        ubeg[l] = take<int32_t>(A, T.cnt[l] + 1);
        if (!ubeg[l]) return LFMM_ERR_NOMEM;
        hipLaunchKernelGGL(k_ad_ubeg, dim3(nblk(T.cnt[l] + 1, 256)), dim3(256), 0, st, T.lbegin[l], T.cnt[l],
                           ubeg[l + 1], (int32_t)n, ubeg[l]);
    }
    // ---- kept boxes per level and their numbering
    std::vector<uint32_t*> keep(depth, nullptr), kex(depth, nullptr);
    for (int l = 0; l < depth; ++l) {
        keep[l] = take<uint32_t>(A, T.cnt[l] + 1);
        kex[l] = take<uint32_t>(A, T.cnt[l] + 1);
        if (!keep[l] || !kex[l]) return LFMM_ERR_NOMEM;
        if (l == 0) {
            LFMM_HIP(fill_u32(keep[0], 1, 1u, st));
        } else {
            hipLaunchKernelGGL(k_ad_keep, dim3(nblk(T.cnt[l], 256)), dim3(256), 0, st, T.uparent[l], T.cnt[l],
                               ubeg[l - 1], leaf_max, l <= min_leaf_level ? 1 : 0, keep[l]);
        }
        LFMM_HIP(fill_u32(keep[l] + T.cnt[l], 1, 0u, st));
        int rc = scan_total(A, keep[l], kex[l], T.cnt[l], st);
        if (rc != LFMM_OK) return rc;
    }
    std::vector<uint32_t> kept(depth, 0);
    for (int l = 0; l < depth; ++l)
        LFMM_HIP(hipMemcpyAsync(&kept[l], kex[l] + T.cnt[l], 4, hipMemcpyDeviceToHost, st));
    LFMM_HIP(hipStreamSynchronize(st));
    T.loff.assign(depth + 1, 0);
    for (int l = 0; l < depth; ++l) {
        if (kept[l] == 0) return LFMM_ERR_INVALID;  // every level holds a box (the depth needs it)
        T.loff[l + 1] = T.loff[l] + (int)kept[l];
    }
    const int total = T.loff[depth];
    T.keys_all = take<uint64_t>(A, total);
    T.range_all = take<int2>(A, total);
    T.box2leaf = take<int32_t>(A, total);
    if (!T.keys_all || !T.range_all || !T.box2leaf) return LFMM_ERR_NOMEM;
    std::vector<int32_t*> child(depth, nullptr);
    for (int l = 0; l < depth; ++l) {
        child[l] = take<int32_t>(A, (int64_t)kept[l] + 1);
        if (!child[l]) return LFMM_ERR_NOMEM;
        const bool lst = l == last;
        hipLaunchKernelGGL(k_ad_fill, dim3(nblk(T.cnt[l], 256)), dim3(256), 0, st, lst ? 1 : 0,
                           l >= min_leaf_level ? 1 : 0, T.lkeys[l], ubeg[l],
                           T.lbegin[l], T.cnt[l], keep[l], kex[l], lst ? nullptr : kex[l + 1], leaf_max, T.loff[l],
                           T.keys_all, T.range_all, child[l], T.box2leaf);
        hipLaunchKernelGGL(k_ad_set1, dim3(1), dim3(1), 0, st, child[l] + kept[l], lst ? 0 : (int32_t)kept[l + 1]);
    }
    // the tree's level arrays now describe the kept boxes
    for (int l = 0; l < depth; ++l) {
        T.cnt[l] = kept[l];
        T.lkeys[l] = T.keys_all + T.loff[l];
        T.lbegin[l] = child[l];
        T.lut[l] = nullptr;
        if (l >= 2 && l <= lut_max_level(3)) {
            T.lut[l] = take<int32_t>(A, (int64_t)1 << (3 * l));
            if (!T.lut[l]) return LFMM_ERR_NOMEM;
            LFMM_HIP(build_lut(3, T.lkeys[l], T.cnt[l], l, T.lut[l], st));
        }
    }
    T.adaptive = true;

    // ---- global leaf list (particle order)
    {
        int32_t* tag = take<int32_t>(A, n);
        uint32_t* fl = take<uint32_t>(A, n);
        uint32_t* ex = take<uint32_t>(A, n);
        uint32_t* sw = take<uint32_t>(A, scan_ws_elems(n));
        if (!tag || !fl || !ex || !sw) return LFMM_ERR_NOMEM;
        LFMM_HIP(fill_u32((uint32_t*)tag, n, 0xffffffffu, st));
        hipLaunchKernelGGL(k_ad_leaf_tag, dim3(nblk(total, 256)), dim3(256), 0, st, T.range_all, T.box2leaf, total, tag);
        hipLaunchKernelGGL(k_ad_flags, dim3(nblk(n, 256)), dim3(256), 0, st, tag, n, fl);
        LFMM_HIP(exclusive_scan(fl, ex, n, sw, st));
        uint32_t tail[2];
        LFMM_HIP(hipMemcpyAsync(&tail[0], ex + (n - 1), 4, hipMemcpyDeviceToHost, st));
        LFMM_HIP(hipMemcpyAsync(&tail[1], fl + (n - 1), 4, hipMemcpyDeviceToHost, st));
        LFMM_HIP(hipStreamSynchronize(st));
        T.nleaf = (int64_t)tail[0] + tail[1];
        T.leaf_begin = take<int32_t>(A, T.nleaf + 1);
        T.leaf_gid = take<int32_t>(A, T.nleaf);
        T.leaf_key = take<uint64_t>(A, T.nleaf);
        T.leaf_lvl = take<int32_t>(A, T.nleaf);
        if (!T.leaf_begin || !T.leaf_gid || !T.leaf_key || !T.leaf_lvl) return LFMM_ERR_NOMEM;
        hipLaunchKernelGGL(k_ad_leaf_list, dim3(nblk(n, 256)), dim3(256), 0, st, T.view(), tag, ex, n, T.box2leaf,
                           T.leaf_begin, T.leaf_gid, T.leaf_key, T.leaf_lvl, T.leaf_of);
    }
    const AdLevels V = T.view();
    const int64_t nl = T.nleaf;

    // ---- U lists
    {
        int32_t* nbr27 = take<int32_t>(A, nl * 27);
        uint32_t* nu1 = take<uint32_t>(A, nl + 1);
        uint32_t* nco = take<uint32_t>(A, nl + 1);
        uint32_t* coff = take<uint32_t>(A, nl + 1);
        uint32_t* fcnt = take<uint32_t>(A, nl + 1);
        uint32_t* foff = take<uint32_t>(A, nl + 1);
        uint32_t* ucnt = take<uint32_t>(A, nl + 1);
        T.u_off = take<uint32_t>(A, nl + 1);
        if (!nbr27 || !nu1 || !nco || !coff || !fcnt || !foff || !ucnt || !T.u_off) return LFMM_ERR_NOMEM;
        hipLaunchKernelGGL(k_ad_u1, dim3(nblk(nl, 256)), dim3(256), 0, st, V, T.leaf_gid, T.leaf_lvl, (int)nl, nbr27,
                           nu1, nco);
        LFMM_HIP(fill_u32(nco + nl, 1, 0u, st));
        int rc = scan_total(A, nco, coff, nl, st);
        if (rc != LFMM_OK) return rc;
        uint32_t npairs = 0;
        rc = read_u32(coff + nl, &npairs, st);
        if (rc != LFMM_OK) return rc;
        LFMM_HIP(fill_u32(fcnt, nl + 1, 0u, st));
        uint32_t* fine_leaf = nullptr;
        if (npairs > 0) {
            uint64_t* pk = take<uint64_t>(A, npairs);
            uint32_t* pv = take<uint32_t>(A, npairs);
            void* sws = A.alloc((size_t)sort_ws_bytes(npairs));
            if (!pk || !pv || !sws) return LFMM_ERR_NOMEM;
            hipLaunchKernelGGL(k_ad_u_pairs, dim3(nblk(nl, 256)), dim3(256), 0, st, nbr27, T.leaf_lvl, (int)nl, coff, pk,
                               pv, fcnt);
            uint64_t* sk = pk;
            LFMM_HIP(radix_sort(pk, pv, npairs, bits_for(nl), sws, st, &sk, &fine_leaf));
        }
        rc = scan_total(A, fcnt, foff, nl, st);
        if (rc != LFMM_OK) return rc;
        hipLaunchKernelGGL(k_ad_add, dim3(nblk(nl + 1, 256)), dim3(256), 0, st, nu1, fcnt, nl, ucnt);
        rc = scan_total(A, ucnt, T.u_off, nl, st);
        if (rc != LFMM_OK) return rc;
        uint32_t ut = 0;
        rc = read_u32(T.u_off + nl, &ut, st);
        if (rc != LFMM_OK) return rc;
        T.u_total = ut;
        T.u_rng = take<int2>(A, ut);
        if (!T.u_rng) return LFMM_ERR_NOMEM;
        hipLaunchKernelGGL(k_ad_u_fill, dim3(nblk(nl, 256)), dim3(256), 0, st, nbr27, T.leaf_begin, (int)nl, T.u_off,
                           foff, fcnt, fine_leaf, T.u_rng);
    }

    // ---- X and W lists
    T.xseg.assign(depth + 1, 0);
    T.x_pairs = 0;
    T.w_off = take<uint32_t>(A, nl + 1);
    if (!T.w_off) return LFMM_ERR_NOMEM;
    if (depth > 2) {
        const int g2 = T.loff[2];
        const int nb = total - g2;
        int32_t* x27 = take<int32_t>(A, (int64_t)nb * 27);
        uint32_t* xcnt = take<uint32_t>(A, nb + 1);
        uint32_t* nzf = take<uint32_t>(A, nb + 1);
        uint32_t* lex = take<uint32_t>(A, nb + 1);
        uint32_t* wcnt = take<uint32_t>(A, nl + 1);
        T.x_off = take<uint32_t>(A, nb + 1);
        if (!x27 || !xcnt || !nzf || !lex || !wcnt || !T.x_off) return LFMM_ERR_NOMEM;
        hipLaunchKernelGGL(k_ad_x, dim3(nblk(nb, 256)), dim3(256), 0, st, V, g2, total, x27, xcnt);
        LFMM_HIP(fill_u32(xcnt + nb, 1, 0u, st));
        int rc = scan_total(A, xcnt, T.x_off, nb, st);
        if (rc != LFMM_OK) return rc;
        hipLaunchKernelGGL(k_ad_p2l_flag, dim3(nblk(nb, 256)), dim3(256), 0, st, xcnt, (int64_t)nb, T.box2leaf, g2, nzf);
        LFMM_HIP(fill_u32(nzf + nb, 1, 0u, st));
        rc = scan_total(A, nzf, lex, nb, st);
        if (rc != LFMM_OK) return rc;
        // pair count and the xlist segments of every level
        std::vector<uint32_t> segs(depth + 1, 0);
        uint32_t xp = 0;
        LFMM_HIP(hipMemcpyAsync(&xp, T.x_off + nb, 4, hipMemcpyDeviceToHost, st));
        for (int l = 2; l <= depth; ++l)
            LFMM_HIP(hipMemcpyAsync(&segs[l], lex + (T.loff[l] - g2), 4, hipMemcpyDeviceToHost, st));
        LFMM_HIP(hipStreamSynchronize(st));
        for (int l = 2; l <= depth; ++l) T.xseg[l] = segs[l];
        T.x_pairs = xp;
        T.x_rng = take<int2>(A, xp);
        T.xlist = take<int32_t>(A, (int64_t)segs[depth]);
        uint64_t* wk = take<uint64_t>(A, xp);
        uint32_t* wv = take<uint32_t>(A, xp);
        void* sws = A.alloc((size_t)sort_ws_bytes(std::max<int64_t>(xp, 1)));
        if (!T.x_rng || !T.xlist || !wk || !wv || !sws) return LFMM_ERR_NOMEM;
        LFMM_HIP(fill_u32(wcnt, nl + 1, 0u, st));
        hipLaunchKernelGGL(k_ad_x_fill, dim3(nblk(nb, 256)), dim3(256), 0, st, x27, xcnt, T.x_off, nb, T.leaf_begin, g2,
                           T.x_rng, wk, wv, wcnt, T.xlist, lex, T.box2leaf);
        rc = scan_total(A, wcnt, T.w_off, nl, st);
        if (rc != LFMM_OK) return rc;
        T.w_gid = (int32_t*)wv;
        if (xp > 0) {
            uint64_t* sk = wk;
            uint32_t* sv = wv;
            LFMM_HIP(radix_sort(wk, wv, xp, bits_for(nl), sws, st, &sk, &sv));
            T.w_gid = (int32_t*)sv;
        }
    } else {
        LFMM_HIP(fill_u32(T.w_off, nl + 1, 0u, st));
    }

    // ---- near lists of the pair kernel: U, plus the X and small W pairs
    // that run as direct sums
    {
        const int g2 = depth > 2 ? T.loff[2] : total;
        const uint32_t* xo = (depth > 2 && T.x_pairs > 0) ? T.x_off : nullptr;
        const int32_t* wg = T.x_pairs > 0 ? T.w_gid : nullptr;
        uint32_t* ncnt = take<uint32_t>(A, nl + 1);
        T.n_off = take<uint32_t>(A, nl + 1);
        if (!ncnt || !T.n_off) return LFMM_ERR_NOMEM;
        hipLaunchKernelGGL(k_ad_near_count, dim3(nblk(nl + 1, 256)), dim3(256), 0, st, T.leaf_gid, (int)nl, T.u_off, xo,
                           g2, T.w_off, wg, T.range_all, kDirectMax, ncnt);
        int rc = scan_total(A, ncnt, T.n_off, nl, st);
        if (rc != LFMM_OK) return rc;
        uint32_t nt = 0;
        if ((rc = read_u32(T.n_off + nl, &nt, st)) != LFMM_OK) return rc;
        T.n_total = nt;
        T.n_rng = take<int2>(A, nt);
        if (!T.n_rng) return LFMM_ERR_NOMEM;
        hipLaunchKernelGGL(k_ad_near_fill, dim3(nblk(nl, 256)), dim3(256), 0, st, T.leaf_gid, (int)nl, T.u_off, T.u_rng,
                           xo, T.x_rng, g2, T.w_off, wg, T.range_all, kDirectMax, T.n_off, T.n_rng);
    }
    return LFMM_OK;
}

int forest_leaf_scale(lfmm_ctx* ctx, const Tree& T, const ForestSpec& fs, double* out)
{
    if (T.nleaf > 0)
        hipLaunchKernelGGL(k_forest_leaf_scale, dim3(nblk(T.nleaf, 256)), dim3(256), 0, ctx->stream, T.leaf_key,
                           T.leaf_lvl, T.nleaf, fs.level, fs.side, fs.sys_scale, out);
    LFMM_HIP(hipGetLastError());
    return LFMM_OK;
}

int ad_count_pairs(const int32_t* leaf_begin, const uint32_t* u_off, const int2* u_rng, int64_t nleaf,
                   unsigned long long* ws, int64_t* pairs, int64_t* reads, hipStream_t st)
{
    LFMM_HIP(fill_u32((uint32_t*)ws, 4, 0u, st));
    hipLaunchKernelGGL(k_ad_pairs, dim3(nblk(nleaf, 256)), dim3(256), 0, st, leaf_begin, u_off, u_rng, nleaf, ws);
    unsigned long long h[2] = {0, 0};
    LFMM_HIP(hipMemcpyAsync(h, ws, 16, hipMemcpyDeviceToHost, st));
    LFMM_HIP(hipStreamSynchronize(st));
    *pairs = (int64_t)h[0];
    if (reads) *reads = (int64_t)h[1];
    return LFMM_OK;
}

}  // namespace lfmm::host
