// api_internal.h -- host-side internals shared by the 3D and 2D solves:
// context, workspace arena, device tables, particle views and the GPU tree.
#pragma once
#include "lfmm.h"
#include "lfmm_internal.h"
#include "ops3d.h"
#include "tree.h"
#include "adaptive.h"
#include "order.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

namespace lfmm::host {

constexpr int kMaxDepth = 16;   // 15 bits per axis at the leaves (3D)
// dense key -> box tables up to 2^24 entries (levels <= 8 in 3D, <= 12 in 2D)
constexpr int lut_max_level(int dim) { return 24 / dim; }

struct Arena {
    struct Chunk { char* p; size_t size; };
    std::vector<Chunk> chunks;
    size_t cur = 0, off = 0, used = 0, peak = 0;
    void reset()
    {
        if (chunks.size() > 1) {  // consolidate into one chunk sized for the last peak
            for (auto& c : chunks) (void)hipFree(c.p);
            chunks.clear();
            const size_t want = peak + peak / 8 + (1 << 20);
            char* p = nullptr;
            if (hipMalloc(&p, want) == hipSuccess) chunks.push_back({p, want});
        }
        cur = 0;
        off = 0;
        used = 0;
    }
    void* alloc(size_t bytes)
    {
        bytes = (bytes + 255) & ~(size_t)255;
        if (bytes == 0) bytes = 256;
        while (cur < chunks.size()) {
            if (off + bytes <= chunks[cur].size) {
                void* p = chunks[cur].p + off;
                off += bytes;
                used += bytes;
                peak = std::max(peak, used);
                return p;
            }
            ++cur;
            off = 0;
        }
        size_t sz = std::max<size_t>(bytes, (size_t)64 << 20);
        if (!chunks.empty()) sz = std::max(sz, chunks.back().size);
        char* p = nullptr;
        if (hipMalloc(&p, sz) != hipSuccess) return nullptr;
        chunks.push_back({p, sz});
        cur = chunks.size() - 1;
        off = bytes;
        used += bytes;
        peak = std::max(peak, used);
        return p;
    }
    void release()
    {
        for (auto& c : chunks) (void)hipFree(c.p);
        chunks.clear();
        cur = off = used = peak = 0;
    }
};

struct DevTables {
    int N = 0;
    void* block = nullptr;
    RtrTables rt{};
    const double* knm = nullptr;
    uint4* m16 = nullptr;      // split-fp16 dense M2L operators (built on first mixed solve)
    const int2* m16_kpre = nullptr;  // their block prefixes (same allocation)
    double m16_kinv = 0.0;
    const uint4* m2m16 = nullptr;      // the same for the 8 M2M operators (same allocation)
    const int2* m2m16_kpre = nullptr;
    double m2m16_kinv = 0.0;
};



struct DevTables2d;  // api2d.cpp

}  // namespace lfmm::host

struct lfmm_ctx {
    int device = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    lfmm::host::Arena arena;
    std::vector<hipEvent_t> events;  // timing events, created on first timed call
    std::map<int, lfmm::host::DevTables> tables;
    std::map<int, lfmm::host::DevTables2d*> tables2d;
    std::vector<lfmm_ctx*> workers;  // lfmm_fmm3d_batch worker contexts, created on first use
    hipEvent_t switch_event = nullptr;  // orders a stream switch (lfmm_set_stream)
    void* vec_tmp = nullptr;            // two field buffers of the vector-charge device path
    size_t vec_tmp_bytes = 0;
    float* order_model = nullptr;       // the order network (lfmm_set_order_model), on the device
    double order_conf = 0.95;           // cumulative class probability of the chosen order
    double order_tol = 1e-6;            // sampled rel-L2 above which the fixed order reruns the solve
    int order_samples = 256;            // targets of the sampled direct-sum check
    void* forest_buf = nullptr;         // lfmm_fmm3d_batch with LFMM_FLAG_ADAPTIVE: the mapped systems
    size_t forest_bytes = 0;
    hipStream_t side_stream = nullptr;  // the sampled check's direct sums, beside the solve
    hipEvent_t side_events[2] = {nullptr, nullptr};
    std::mutex mu;
};


namespace lfmm::host {

#define LFMM_HIP(call)                                   \
    do {                                                 \
        hipError_t e_ = (call);                          \
        if (e_ != hipSuccess) return LFMM_ERR_HIP;       \
    } while (0)

template <typename T>
T* take(Arena& a, int64_t count)
{
    return static_cast<T*>(a.alloc(sizeof(T) * (size_t)std::max<int64_t>(count, 1)));
}

// Device-resident Morton tree of one solve (all buffers in the context arena).
struct Tree {
    int64_t n = 0;
    int dim = 3;
    int depth = 0;
    const double* dpos = nullptr;  // device view of the inputs
    const double* dq = nullptr;
    int64_t ps = 0, qs = 0;
    double* geom = nullptr;        // {shift x, y, z, scale}
    double geom_host[4] = {};      // the same, read back with the level sizes
    uint32_t* umax = nullptr;
    uint32_t* order = nullptr;     // sorted -> input index
    double4* src = nullptr;        // sorted (x, y, z, q)
    int32_t* leaf_of = nullptr;
    std::vector<int64_t> cnt;
    std::vector<uint64_t*> lkeys;
    std::vector<int32_t*> lbegin;  // leaves: particle ranges; above: child ranges
    std::vector<int32_t*> lut;
    uint32_t max_leaf = 0;
    // uniform build with keep_parents: parent index of every box at levels >= 1
    std::vector<int32_t*> uparent;
    // ---- adaptive octree (build_adaptive).  cnt / lkeys / lbegin / lut then
    // describe the kept boxes of each level (lbegin: child ranges, empty for
    // a leaf); boxes are also numbered globally (gid), level by level.
    bool adaptive = false;
    std::vector<int> loff;            // gid of the first box of each level, depth + 1 entries
    uint64_t* keys_all = nullptr;     // [total] (lkeys[l] = keys_all + loff[l])
    int2* range_all = nullptr;        // [total] particle range of each box
    int32_t* box2leaf = nullptr;      // [total] leaf index or -1
    int64_t nleaf = 0;                // global leaf list, in particle order:
    int32_t* leaf_begin = nullptr;    //   [nleaf + 1] first particle
    int32_t* leaf_gid = nullptr;      //   [nleaf] box
    uint64_t* leaf_key = nullptr;     //   [nleaf] key at its level
    int32_t* leaf_lvl = nullptr;      //   [nleaf] level
    uint32_t* u_off = nullptr;        // U lists: [nleaf + 1] offsets into u_rng
    int2* u_rng = nullptr;            //   particle ranges of the touching leaves
    int64_t u_total = 0;
    uint32_t* x_off = nullptr;        // X lists: per box t = gid - loff[2], [loff[depth] - loff[2] + 1]
    int2* x_rng = nullptr;            //   particle ranges of the source leaves
    int32_t* xlist = nullptr;         //   boxes with a non-empty X list (t), gid order
    std::vector<int64_t> xseg;        //   xlist entries of level l: [xseg[l], xseg[l+1])
    int64_t x_pairs = 0;
    uint32_t* w_off = nullptr;        // W lists: [nleaf + 1] offsets into w_gid
    int32_t* w_gid = nullptr;         //   boxes whose multipoles act on the leaf
    uint32_t* n_off = nullptr;        // near lists of the pair kernel: [nleaf + 1] offsets into n_rng
    int2* n_rng = nullptr;            //   U, the leaf's X sources, W boxes of <= kDirectMax particles
    int64_t n_total = 0;
    int total() const { return loff.empty() ? 0 : loff.back(); }
    AdLevels view() const;
};

// Device view of the particle set a solve works on.
struct DevPoints {
    const double* pos = nullptr;
    int64_t ps = 3;
    const double* q = nullptr;
    int64_t qs = 1;
    int64_t n = 0;
    int dim = 3;
};


hipEvent_t ctx_event(lfmm_ctx* ctx, size_t i);
double ms_since(std::chrono::steady_clock::time_point t0);
// Copies a host view (dim coordinates per point) to the device, keeping its strides.
int upload_view(lfmm_ctx* ctx, const lfmm_points3d* S, int dim, DevPoints& P);
// The particle set of a solve: the sources alone, or [targets (q = 0); sources].
int stage_inputs(lfmm_ctx* ctx, const lfmm_points3d* S, const lfmm_points3d* Tg, int memory, int dim, DevPoints& P);
// fixed_geom: device (shift x, y, z, scale) of the tree instead of the
// particles' extent (no non-finite check then); min_depth: at least so many
// levels.
int build_tree(lfmm_ctx* ctx, const DevPoints& DP, int depth_in, int leaf_max, bool partition, Tree& T,
               bool keep_parents = false, const double* fixed_geom = nullptr, int min_depth = 0);

// Many independent systems as one adaptive octree: system s occupies the
// level-`level` cell (2a, 2b, 2c) of the unit cube, a = s / side^2,
// b = (s / side) % side, c = s % side, its particles mapped affinely into
// that cell (forest length = sys_scale[s] x the system's length).  The cells
// between systems stay empty, so from level + 1 on no U, V, W or X list
// reaches from one system into another; the solve keeps every box down to
// `level` (the system cells are never merged into one leaf), runs M2L from
// level + 2 (a system's level 2) and the matrix-core M2L from level + 4, and
// the leaf kernels clamp and scale each system's pairs and outputs back to
// its own units.
struct ForestSpec {
    int level = 0;
    int side = 0;
    const double* geom = nullptr;       // device (0, 0, 0, 1)
    const double* sys_scale = nullptr;  // device [side^3]
};
// Per leaf of the forest's adaptive octree: sys_scale of its system.
int forest_leaf_scale(lfmm_ctx* ctx, const Tree& T, const ForestSpec& fs, double* out);
// W-list boxes of at most this many particles act on a leaf by direct sums
// instead of M2P (a direct sum costs ~20 flops per source and target, M2P
// ~20 per coefficient, tri(13) = 91 at 13 terms); X lists into a leaf always
// run as direct sums, P2L serves boxes with children only.
constexpr int kDirectMax = 96;
// Turns the uniform tree T (built with keep_parents) into the adaptive octree
// of leaf size leaf_max and builds its U, X and W lists (adaptive.hip).
// min_leaf_level > 0: every box above that level is kept and has children
// (leaves sit on that level or deeper).
int build_adaptive(lfmm_ctx* ctx, int leaf_max, Tree& T, int min_leaf_level = 0);
// ---- per-box orders (order.hip)
struct BoxOrders {
    uint32_t* box_terms = nullptr;   // [total] terms of each box (a parent: the largest of its children)
    int32_t* plist = nullptr;        // per box of the bucketed levels, its parent (index at its level),
                                     // sorted by (parent level, the box's terms, the box's class)
    std::vector<int64_t> seg;        // plist entries of key (level*32 + terms)*8 + class: [seg[key], seg[key+1])
    int max_terms = 0;               // over the leaves
    int min_leaf_terms = 0;
};
// 8 features per leaf of an adaptive tree (feat: nleaf * 8 floats, device).
int order_features(lfmm_ctx* ctx, const Tree& T, const DevTables* DT, float* feat);
// Order network forward: probabilities (n x 13) and/or terms by the confidence rule.
int order_predict(hipStream_t st, const float* model, const float* feat, int64_t n, float conf, float* probs,
                  int32_t* terms);
// Cross-entropy gradient over samples idx[0..n) (or 0..n), scaled by
// inv_batch, and the summed loss; workspace gpart [order_grad_parts(n) *
// LFMM_ORDER_MODEL_SIZE], lpart [order_grad_parts(n)].
int64_t order_grad_parts(int64_t n);
int order_grad(hipStream_t st, const float* model, const float* feat, const int32_t* label, const int32_t* idx,
               int64_t n, float inv_batch, float* grad, float* loss, float* gpart, float* lpart);
int order_accum(hipStream_t st, const float* loss, float* acc);
int order_adam(hipStream_t st, float* model, const float* grad, float* m, float* v, float lr, int step);
// Leaf terms = the largest request of the leaf's particles (requests in input order), clamped to [lo, hi].
int order_leaf_terms_from_requests(hipStream_t st, const Tree& T, const int32_t* req, int64_t n_req, int lo, int hi,
                                   uint32_t* leaf_terms);
// Box terms and the (level, terms) launch buckets of the boxes of levels >= first_bucket_level.  promote:
// buckets of few boxes take the next larger order of their level (fewer launches).  Synchronises.
int order_box_terms(lfmm_ctx* ctx, const Tree& T, const uint32_t* leaf_terms, int first_bucket_level, BoxOrders& O,
                    bool promote = false);
// Algorithmic FP64 flops of all M2L operations at their orders (levels below
// first_bucket_level at N).  Synchronises.
int order_m2l_flops(lfmm_ctx* ctx, const Tree& T, const BoxOrders& O, int N, int first_bucket_level, double* flops);
// Sampled direct-sum check: rel-L2 of potential (self pair excluded) and
// field on `samples` targets.  order_check_start launches the FP64 direct
// sums of the sampled targets on the context's side stream once the staged
// inputs and the packed sources `src` are ready (after the main stream's
// work so far); order_check_finish waits for them on the main stream,
// compares with the solve's outputs and synchronises.
struct OrderCheck {
    int S = 0;
    bool started = false;
    double4* rec = nullptr;
    int64_t* at = nullptr;
    double* dp = nullptr;
    double* df = nullptr;
    double4* out = nullptr;
};
int order_check_start(lfmm_ctx* ctx, const DevPoints& DP, const double4* src, int64_t n, int64_t n_out, bool has_q,
                      int samples, OrderCheck& C);
int order_check_finish(lfmm_ctx* ctx, const OrderCheck& C, const double* pot, const double* field, bool fmm_has_self,
                       double* rel_pot, double* rel_field);

// Near-field pair count of an adaptive tree's U lists and the particles of
// the lists summed over the leaves (ws: 2 counters).  Synchronises.
int ad_count_pairs(const int32_t* leaf_begin, const uint32_t* u_off, const int2* u_rng, int64_t nleaf,
                   unsigned long long* ws, int64_t* pairs, int64_t* reads, hipStream_t st);
// The 3D solve behind lfmm_fmm3d (arguments checked, ctx->mu held, device
// set).  tables: operator tables already on the device, or null to take the
// context's cached ones (built on first use).
// req: per-box orders requested per particle (input order, memory as the
// outputs; n_out entries), or null.
// fs: a forest of independent systems solved as one adaptive octree (see
// ForestSpec), or null.
struct ForestSpec;
int fmm3d_solve(lfmm_ctx* ctx, const lfmm_points3d* S, const lfmm_points3d* Tg, double eps, int terms_in,
                int depth_in, int leaf_max, int memory, unsigned flags, double* pot_out, double* field_out,
                lfmm_report* rep, const DevTables* tables, const int32_t* req = nullptr,
                const ForestSpec* fs = nullptr);
// fmm3d_solve's status when a learned-order solve failed its sampled check
// (the caller reruns it with the fixed order)
constexpr int kNeedFallback = -2;
int get_tables(lfmm_ctx* ctx, int N, const DevTables** out);
int terms_3d(double eps, int terms);
constexpr int kNeedSolve = -1;
int fmm3d_prologue(lfmm_ctx* c, const lfmm_points3d* sources, const lfmm_points3d* targets, double eps, int terms,
                   int tree_depth, int memory, double* potential, double* field, lfmm_report* report,
                   const lfmm_points3d** tg_out);
bool valid_points(const lfmm_points3d* p, bool need_q, int dim);
void free_tables2d(lfmm_ctx* ctx);

}  // namespace lfmm::host
