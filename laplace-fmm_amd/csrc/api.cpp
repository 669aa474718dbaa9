// api.cpp -- C ABI of the MI355X Laplace FMM (see include/lfmm.h).
//
// Host side of the solve: argument checks, the reference's parameter rules
// (src/cpu/multipole3d.cpp:1321-1337), the workspace arena, per-N operator
// tables on the device, and the launch sequence
//   extent -> keys -> radix sort -> gather -> levels -> P2M -> M2M (up)
//   -> M2L+L2L (down) -> P2P+L2P (leaves, scattered to input order).
// Everything runs on one HIP stream; the only host synchronisations are the
// box counts of each level (needed to size the next launch).
#include "api_internal.h"
#include "direct.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <vector>

namespace lfmm::host {

using namespace lfmm;

// First M2L level on the matrix cores: the two coarsest levels stay FP64.
constexpr int kMixedFirstLevel = 4;
// Adaptive octree, fixed order: the M2L of a level whose parent level holds
// at least this many boxes runs over the parents with children only (a
// parent list), so no lane holds a leaf parent; smaller levels keep the
// direction-split launch over all parents.
constexpr int64_t kBucketMinParents = 4096;

int get_tables(lfmm_ctx* ctx, int N, const DevTables** out)
{
    auto it = ctx->tables.find(N);
    if (it != ctx->tables.end()) {
        *out = &it->second;
        return LFMM_OK;
    }
    OpTables T;
    build_op_tables(N, T);
    const size_t b_rot = T.rot.size() * 8, b_eph = T.ephi.size() * 8, b_tr = T.trans.size() * 8;
    const size_t b_dt = T.dir_theta.size() * 4, b_dtr = T.dir_trans.size() * 4, b_dv = T.dir_vec.size() * 4,
                 b_cd = T.class_dirs.size() * 4, b_k = T.knm.size() * 8;
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t total = al(b_rot) + al(b_eph) + al(b_tr) + al(b_dt) + al(b_dtr) + al(b_dv) + al(b_cd) + al(b_k);
    char* blk = nullptr;
    if (hipMalloc(&blk, total) != hipSuccess) return LFMM_ERR_NOMEM;
    size_t o = 0;
    bool copy_ok = true;
    auto put = [&](const void* src, size_t bytes) -> void* {
        void* dst = blk + o;
        copy_ok = copy_ok && hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess;
        o += al(bytes);
        return dst;
    };
    DevTables D;
    D.N = N;
    D.block = blk;
    D.rt.rot = (const double2*)put(T.rot.data(), b_rot);
    D.rt.ephi = (const double2*)put(T.ephi.data(), b_eph);
    D.rt.trans = (const double*)put(T.trans.data(), b_tr);
    D.rt.dir_theta = (const int*)put(T.dir_theta.data(), b_dt);
    D.rt.dir_trans = (const int*)put(T.dir_trans.data(), b_dtr);
    D.rt.dir_vec = (const int*)put(T.dir_vec.data(), b_dv);
    D.rt.class_dirs = (const int*)put(T.class_dirs.data(), b_cd);
    D.knm = (const double*)put(T.knm.data(), b_k);
    if (!copy_ok || hipDeviceSynchronize() != hipSuccess) {
        (void)hipFree(blk);
        return LFMM_ERR_HIP;
    }
    auto res = ctx->tables.emplace(N, D);
    *out = &res.first->second;
    return LFMM_OK;
}

// Split-fp16 dense M2L operators of N terms (m2l_dense.cpp), built once per context.
int get_m16(lfmm_ctx* ctx, int N, const DevTables** out)
{
    DevTables& D = ctx->tables.at(N);
    if (!D.m16) {
        OpTables T;
        build_op_tables(N, T);
        std::vector<uint16_t> pk, pm;
        std::vector<int> kp, km;
        double kinv = 0.0, kinvm = 0.0;
        pack_m2l16(T, pk, kinv, kp);
        pack_m2m16(T, pm, kinvm, km);
        auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
        const size_t bk = al(pk.size() * 2), bkp = al(kp.size() * 4), bm = al(pm.size() * 2);
        char* p = nullptr;
        if (hipMalloc(&p, bk + bkp + bm + km.size() * 4) != hipSuccess) return LFMM_ERR_NOMEM;
        if (hipMemcpy(p, pk.data(), pk.size() * 2, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(p + bk, kp.data(), kp.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(p + bk + bkp, pm.data(), pm.size() * 2, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(p + bk + bkp + bm, km.data(), km.size() * 4, hipMemcpyHostToDevice) != hipSuccess) {
            (void)hipFree(p);
            return LFMM_ERR_HIP;
        }
        D.m16 = (uint4*)p;
        D.m16_kpre = (const int2*)(p + bk);
        D.m16_kinv = kinv;
        D.m2m16 = (const uint4*)(p + bk + bkp);
        D.m2m16_kpre = (const int2*)(p + bk + bkp + bm);
        D.m2m16_kinv = kinvm;
    }
    *out = &D;
    return LFMM_OK;
}

// The reference's expansion-size rule, src/cpu/multipole3d.cpp:1326-1331.
int terms_3d(double eps, int terms)
{
    if (terms == LFMM_AUTO) {
        eps = std::max(eps, 1.e-11);
        const int a = (int)(54.2 - 10.73 * std::sqrt(std::log(2.4 * eps) + 25.5));
        return std::min(LFMM_MAX_TERMS_3D, a);
    }
    return std::min(LFMM_MAX_TERMS_3D, terms);
}

// Timing events on the solve's stream: ev(i) grows the pool on demand.
hipEvent_t ctx_event(lfmm_ctx* ctx, size_t i)
{
    while (ctx->events.size() <= i) {
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        ctx->events.push_back(e);
    }
    return ctx->events[i];
}

double ms_since(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// Copies a host view to the device (keeping its strides).
int upload_view(lfmm_ctx* ctx, const lfmm_points3d* S, int dim, DevPoints& P)
{
    const int64_t n = S->n;
    const int64_t ps = S->pos_stride, qs = S->q_stride;
    hipStream_t st = ctx->stream;
    Arena& A = ctx->arena;
    P.n = n;
    P.ps = ps;
    P.qs = qs;
    P.dim = dim;
    const size_t pbytes = sizeof(double) * (size_t)((n - 1) * ps + dim);
    const bool shared = (S->q > S->pos && S->q < S->pos + (n - 1) * ps + dim && ps == qs);
    if (shared) {  // one interleaved block, e.g. the reference's particle AoS
        const size_t span = sizeof(double) * (size_t)((n - 1) * ps + (S->q - S->pos) + 1);
        const size_t bytes = std::max(span, pbytes);
        double* d = take<double>(A, (int64_t)(bytes / 8) + 1);
        if (!d) return LFMM_ERR_NOMEM;
        LFMM_HIP(hipMemcpyAsync(d, S->pos, bytes, hipMemcpyHostToDevice, st));
        P.q = d + (S->q - S->pos);
        P.pos = d;
        return LFMM_OK;
    }
    double* d = take<double>(A, (int64_t)(pbytes / 8) + 1);
    if (!d) return LFMM_ERR_NOMEM;
    LFMM_HIP(hipMemcpyAsync(d, S->pos, pbytes, hipMemcpyHostToDevice, st));
    P.pos = d;
    P.q = nullptr;
    if (S->q) {
        double* e = take<double>(A, (n - 1) * qs + 1);
        if (!e) return LFMM_ERR_NOMEM;
        LFMM_HIP(hipMemcpyAsync(e, S->q, sizeof(double) * (size_t)((n - 1) * qs + 1), hipMemcpyHostToDevice, st));
        P.q = e;
    }
    return LFMM_OK;
}

// The particle set of a solve: the sources alone, or [targets (q = 0);
// sources] as the reference's source/target constructor builds it
// (src/cpu/multipole3d.cpp:1311-1319), packed as (x, y, z, q) records.
int stage_inputs(lfmm_ctx* ctx, const lfmm_points3d* S, const lfmm_points3d* Tg, int memory, int dim, DevPoints& P)
{
    DevPoints s, t;
    if (memory == LFMM_MEM_HOST) {
        int rc = upload_view(ctx, S, dim, s);
        if (rc != LFMM_OK) return rc;
        if (Tg) {
            rc = upload_view(ctx, Tg, dim, t);
            if (rc != LFMM_OK) return rc;
        }
    } else {
        s = DevPoints{S->pos, S->pos_stride, S->q, S->q_stride, S->n, dim};
        if (Tg) t = DevPoints{Tg->pos, Tg->pos_stride, nullptr, 1, Tg->n, dim};
    }
    if (!Tg) {
        P = s;
        return LFMM_OK;
    }
    const int64_t n = t.n + s.n;
    double4* buf = take<double4>(ctx->arena, n);
    if (!buf) return LFMM_ERR_NOMEM;
    LFMM_HIP(pack_points(dim, t.pos, t.ps, nullptr, 1, t.n, buf, ctx->stream));
    LFMM_HIP(pack_points(dim, s.pos, s.ps, s.q, s.qs, s.n, buf + t.n, ctx->stream));
    P = DevPoints{(const double*)buf, 4, (const double*)buf + 3, 4, n, dim};
    return LFMM_OK;
}

// MortonTree constructor (src/cpu/morton_tree.cpp:286-493) on the GPU.
// `partition` applies the q = 0 partition of every leaf (morton_tree.cpp:426-468);
// a single-set solve skips it, since it evaluates every particle either way.
int build_tree(lfmm_ctx* ctx, const DevPoints& DP, int depth_in, int leaf_max, bool partition, Tree& T,
               bool keep_parents, const double* fixed_geom, int min_depth)
{
    const int64_t n = DP.n;
    const int dim = DP.dim;
    hipStream_t st = ctx->stream;
    Arena& A = ctx->arena;
    T.n = n;
    T.dim = dim;
    const double* dpos = DP.pos;
    const double* dq = DP.q;
    const int64_t ps = DP.ps, qs = DP.qs;
    T.dpos = dpos;
    T.dq = dq;
    T.ps = ps;
    T.qs = qs;

    double* geom = take<double>(A, 4);
    double* part = take<double>(A, 6 * 1024);
    int* nonfinite = take<int>(A, 1);
    uint32_t* umax = take<uint32_t>(A, 64);
    if (!geom || !part || !nonfinite || !umax) return LFMM_ERR_NOMEM;
    T.geom = geom;
    T.umax = umax;
    if (fixed_geom) {
        LFMM_HIP(hipMemcpyAsync(geom, fixed_geom, 4 * sizeof(double), hipMemcpyDeviceToDevice, st));
    } else {
        LFMM_HIP(hipMemsetAsync(nonfinite, 0, sizeof(int), st));
        LFMM_HIP(extent(dim, dpos, ps, n, part, nonfinite, geom, st));
        int bad = 0;
        LFMM_HIP(hipMemcpyAsync(&bad, nonfinite, sizeof(int), hipMemcpyDeviceToHost, st));
        LFMM_HIP(hipStreamSynchronize(st));
        if (bad) return LFMM_ERR_NONFINITE;
    }

    // ---- tree depth (reference rule :1334-1337, or the leaf-size bound)
    int depth = depth_in;
    if (depth_in == LFMM_AUTO) {
        if (leaf_max > 0) {
            int maxl = 1;
            while (maxl < lut_max_level(dim) && ((int64_t)1 << (dim * maxl)) < n) ++maxl;
            int64_t cnt = 0;
            for (int l = 0; l <= maxl; ++l) cnt += (int64_t)1 << (dim * l);
            uint32_t* cws = take<uint32_t>(A, cnt);
            if (!cws) return LFMM_ERR_NOMEM;
            int lvl = 0;
            LFMM_HIP(occupancy_depth(dim, dpos, ps, n, geom, leaf_max, maxl, cws, umax, &lvl, st));
            if (lvl > maxl) {
                // clustered input: beyond the dense histograms, sort the keys
                // of the deepest level once and take the largest run of each
                // coarser key prefix
                const int deep = dim == 3 ? kMaxDepth - 1 : 2 * kMaxDepth - 2;
                uint64_t* dk = take<uint64_t>(A, n);
                uint32_t* dv = take<uint32_t>(A, n);
                void* dws = A.alloc((size_t)sort_ws_bytes(n));
                uint32_t* fl = take<uint32_t>(A, n);
                uint32_t* ex = take<uint32_t>(A, n);
                uint32_t* sw = take<uint32_t>(A, scan_ws_elems(n));
                uint64_t* uk = take<uint64_t>(A, n);
                int32_t* ub = take<int32_t>(A, n + 1);
                if (!dk || !dv || !dws || !fl || !ex || !sw || !uk || !ub) return LFMM_ERR_NOMEM;
                LFMM_HIP(keys(dim, dpos, ps, n, deep, geom, dk, dv, st));
                uint64_t* sk = dk;
                uint32_t* so = dv;
                LFMM_HIP(radix_sort(dk, dv, n, dim * deep, dws, st, &sk, &so));
                lvl = deep;
                for (int l = maxl + 1; l <= deep; ++l) {
                    int64_t cnt = 0;
                    LFMM_HIP(run_compact(sk, n, dim * (deep - l), fl, ex, sw, uk, ub, nullptr, &cnt, st));
                    uint32_t mx = 0;
                    LFMM_HIP(max_run(ub, cnt, umax, &mx, st));
                    if ((int64_t)mx <= leaf_max) {
                        lvl = l;
                        break;
                    }
                }
            }
            depth = lvl + 1;
        } else if (dim == 3) {  // src/cpu/multipole3d.cpp:1334-1337
            depth = (int)std::max(2.0, 1 + std::log((double)n / 100.0) / std::log(8.0));
        } else {                // src/cpu/multipole.cpp:483-487
            depth = (int)std::max(3.0, 2 + std::log((double)n / 100.0) / std::log(4.0));
        }
    }
    depth = std::max(depth, min_depth);
    if (depth < 1 || depth > (dim == 3 ? kMaxDepth : 2 * kMaxDepth - 1)) return LFMM_ERR_DEPTH;
    T.depth = depth;
    const int leaf_level = depth - 1;

    // ---- keys and stable sort
    uint64_t* k0 = take<uint64_t>(A, n);
    uint32_t* v0 = take<uint32_t>(A, n);
    void* sws = A.alloc((size_t)sort_ws_bytes(n));
    if (!k0 || !v0 || !sws) return LFMM_ERR_NOMEM;
    LFMM_HIP(keys(dim, dpos, ps, n, leaf_level, geom, k0, v0, st));
    uint64_t* skeys = k0;
    uint32_t* order = v0;
    if (leaf_level > 0) LFMM_HIP(radix_sort(k0, v0, n, dim * leaf_level, sws, st, &skeys, &order));
    T.order = order;

    // ---- levels
    T.cnt.assign(depth, 0);
    T.lkeys.assign(depth, nullptr);
    T.lbegin.assign(depth, nullptr);
    T.lut.assign(depth, nullptr);
    uint32_t* flags_ws = take<uint32_t>(A, n);
    uint32_t* excl_ws = take<uint32_t>(A, n);
    uint32_t* scan_ws = take<uint32_t>(A, scan_ws_elems(n));
    T.leaf_of = take<int32_t>(A, n);
    if (!flags_ws || !excl_ws || !scan_ws || !T.leaf_of) return LFMM_ERR_NOMEM;
    // Level sizes stay on the device while the levels are compacted: the leaf
    // level's count is read back once (it bounds the launches above it), then
    // every coarser level is sized by the bound min(boxes below, 2^(dim l)),
    // and all counts and the largest leaf come back in one read.
    uint32_t* cnt_dev = take<uint32_t>(A, depth);
    if (!cnt_dev) return LFMM_ERR_NOMEM;
    LFMM_HIP(hipMemsetAsync(cnt_dev, 0, sizeof(uint32_t) * depth, st));
    {
        const int64_t cap = std::min<int64_t>(n, (int64_t)1 << std::min(60, dim * leaf_level));
        T.lkeys[leaf_level] = take<uint64_t>(A, cap);
        T.lbegin[leaf_level] = take<int32_t>(A, cap + 1);
        if (!T.lkeys[leaf_level] || !T.lbegin[leaf_level]) return LFMM_ERR_NOMEM;
        LFMM_HIP(run_compact(skeys, n, 0, flags_ws, excl_ws, scan_ws, T.lkeys[leaf_level], T.lbegin[leaf_level],
                             T.leaf_of, &T.cnt[leaf_level], st));
    }
    if (keep_parents) T.uparent.assign(depth, nullptr);
    std::vector<int64_t> bound(depth, 0);
    bound[leaf_level] = T.cnt[leaf_level];
    for (int l = leaf_level - 1; l >= 0; --l) {
        const int64_t m = bound[l + 1];
        bound[l] = std::min<int64_t>(m, (int64_t)1 << (dim * l));
        T.lkeys[l] = take<uint64_t>(A, bound[l]);
        T.lbegin[l] = take<int32_t>(A, bound[l] + 1);
        if (!T.lkeys[l] || !T.lbegin[l]) return LFMM_ERR_NOMEM;
        if (keep_parents) {
            T.uparent[l + 1] = take<int32_t>(A, m);
            if (!T.uparent[l + 1]) return LFMM_ERR_NOMEM;
        }
        LFMM_HIP(run_compact_dev(T.lkeys[l + 1], m, l + 1 == leaf_level ? nullptr : cnt_dev + l + 1, dim, flags_ws,
                                 excl_ws, scan_ws, T.lkeys[l], T.lbegin[l], keep_parents ? T.uparent[l + 1] : nullptr,
                                 cnt_dev + l, st));
    }
    LFMM_HIP(max_run(T.lbegin[leaf_level], T.cnt[leaf_level], umax, nullptr, st));
    {
        std::vector<uint32_t> h(depth);
        LFMM_HIP(hipMemcpyAsync(h.data(), cnt_dev, sizeof(uint32_t) * depth, hipMemcpyDeviceToHost, st));
        LFMM_HIP(hipMemcpyAsync(&T.max_leaf, umax, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        LFMM_HIP(hipMemcpyAsync(T.geom_host, geom, 4 * sizeof(double), hipMemcpyDeviceToHost, st));
        LFMM_HIP(hipStreamSynchronize(st));
        for (int l = 0; l < leaf_level; ++l) T.cnt[l] = h[l];
    }
    if (dq) {
        // q != 0 first inside every leaf, then the sorted sources
        if (partition) LFMM_HIP(partition_leaves(T.lbegin[leaf_level], T.cnt[leaf_level], order, dq, qs, st));
        T.src = take<double4>(A, n);
        if (!T.src) return LFMM_ERR_NOMEM;
        LFMM_HIP(gather_sources(dim, dpos, ps, dq, qs, order, n, T.src, st));
    }
    for (int l = 2; l < depth; ++l) {
        if (l > lut_max_level(dim)) continue;
        T.lut[l] = take<int32_t>(A, (int64_t)1 << (dim * l));
        if (!T.lut[l]) return LFMM_ERR_NOMEM;
        LFMM_HIP(build_lut(dim, T.lkeys[l], T.cnt[l], l, T.lut[l], st));
    }
    return LFMM_OK;
}

int fmm3d_solve(lfmm_ctx* ctx, const lfmm_points3d* S, const lfmm_points3d* Tg, double eps, int terms_in,
                int depth_in, int leaf_max, int memory, unsigned flags, double* pot_out, double* field_out,
                lfmm_report* rep, const DevTables* tables, const int32_t* req, const ForestSpec* fs)
{
    const auto T0 = std::chrono::steady_clock::now();
    const bool timed = (flags & LFMM_FLAG_TIMINGS) != 0;
    const int64_t n_out = Tg ? Tg->n : S->n;
    // per-box orders: chosen by the order network, or requested per particle
    const bool learned = (flags & LFMM_FLAG_LEARNED_ORDER) != 0 && !req;
    const bool per_box = learned || req;
    if (learned && !ctx->order_model) return LFMM_ERR_INVALID;
    if (per_box && tables) return LFMM_ERR_UNSUPPORTED;
    if (fs && (per_box || Tg || tables)) return LFMM_ERR_UNSUPPORTED;
    // first level with local expansions (M2L) and first matrix-core level;
    // a forest's systems start at fs->level
    const int first = fs ? fs->level + 2 : 2;
    const int mfl = fs ? fs->level + kMixedFirstLevel : kMixedFirstLevel;
    int N = terms_3d(eps, terms_in);
    if (!per_box && N < 1) return LFMM_ERR_TERMS;
    hipStream_t st = ctx->stream;
    Arena& A = ctx->arena;
    A.reset();
    DevPoints DP;
    int rc = stage_inputs(ctx, S, Tg, memory, 3, DP);
    if (rc != LFMM_OK) return rc;
    const int64_t n = DP.n;
    Tree T;
    const bool adaptive = per_box || (flags & LFMM_FLAG_ADAPTIVE) != 0;
    // the adaptive octree splits boxes of more than leaf_max particles (64
    // when not given), down to the depth at which every box holds that many
    if (adaptive && leaf_max <= 0) leaf_max = 64;
    rc = build_tree(ctx, DP, fs ? LFMM_AUTO : depth_in, leaf_max, Tg != nullptr, T, adaptive, fs ? fs->geom : nullptr,
                    fs ? fs->level + 2 : 0);
    if (rc != LFMM_OK) return rc;
    if (adaptive) {
        rc = build_adaptive(ctx, leaf_max, T, fs ? fs->level : 0);
        if (rc != LFMM_OK) return rc;
        LFMM_HIP(max_run(T.leaf_begin, T.nleaf, T.umax, &T.max_leaf, st));
    }
    const int depth = T.depth;
    const int leaf_level = depth - 1;
    const double* geom = T.geom;
    const double4* src = T.src;
    const uint32_t* order = T.order;
    // learned orders: the sampled direct sums start now, beside the solve
    OrderCheck chk;
    struct SideDrain {  // no return path leaves the side stream running on this solve's workspace
        lfmm_ctx* c;
        const OrderCheck& k;
        ~SideDrain()
        {
            if (k.started) (void)hipStreamSynchronize(c->side_stream);
        }
    } drain{ctx, chk};
    if (learned) {
        rc = order_check_start(ctx, DP, src, n, n_out, DP.q != nullptr, ctx->order_samples, chk);
        if (rc != LFMM_OK) return rc;
    }
    const int32_t* leaf_of = T.leaf_of;
    std::vector<int64_t>& cnt = T.cnt;
    std::vector<uint64_t*>& lkeys = T.lkeys;
    std::vector<int32_t*>& lbegin = T.lbegin;
    std::vector<int32_t*>& lut = T.lut;
    const uint32_t max_leaf = T.max_leaf;
    const int64_t nleaf = adaptive ? T.nleaf : cnt[leaf_level];

    // ---- per-box orders: leaf terms, box terms (a parent: the largest of its
    // children) and the (level, terms) launch buckets of levels >= 4
    BoxOrders O;
    if (adaptive && !per_box && N >= 1 && N <= LFMM_MAX_TERMS_3D) {
        // one order: the buckets are the parent lists of the levels
        uint32_t* leaf_terms = take<uint32_t>(A, T.nleaf);
        if (!leaf_terms) return LFMM_ERR_NOMEM;
        LFMM_HIP(fill_u32(leaf_terms, T.nleaf, (uint32_t)N, st));
        if ((rc = order_box_terms(ctx, T, leaf_terms, mfl, O)) != LFMM_OK) return rc;
    }
    if (per_box) {
        const DevTables* DTf = nullptr;  // direction tables (the same for every N)
        rc = get_tables(ctx, kOrderMinTerms + kOrderClasses - 1, &DTf);
        if (rc != LFMM_OK) return rc;
        uint32_t* leaf_terms = take<uint32_t>(A, T.nleaf);
        if (!leaf_terms) return LFMM_ERR_NOMEM;
        if (learned) {
            float* feat = take<float>(A, T.nleaf * kOrderFeatures);
            if (!feat) return LFMM_ERR_NOMEM;
            if ((rc = order_features(ctx, T, DTf, feat)) != LFMM_OK) return rc;
            rc = order_predict(st, ctx->order_model, feat, T.nleaf, (float)ctx->order_conf, nullptr, (int32_t*)leaf_terms);
            if (rc != LFMM_OK) return rc;
        } else {
            const int32_t* dreq = req;
            if (memory == LFMM_MEM_HOST) {
                int32_t* r = take<int32_t>(A, n_out);
                if (!r) return LFMM_ERR_NOMEM;
                LFMM_HIP(hipMemcpyAsync(r, req, sizeof(int32_t) * n_out, hipMemcpyHostToDevice, st));
                dreq = r;
            }
            rc = order_leaf_terms_from_requests(st, T, dreq, n_out, 1, LFMM_MAX_TERMS_3D, leaf_terms);
            if (rc != LFMM_OK) return rc;
        }
        if ((rc = order_box_terms(ctx, T, leaf_terms, mfl, O, learned)) != LFMM_OK) return rc;
        N = O.max_terms;
        if (N < 1 || N > LFMM_MAX_TERMS_3D) return LFMM_ERR_TERMS;
        // the operator tables of every bucket's order, before any launch
        for (int l = mfl; l < depth; ++l)
            for (int t = 1; t <= LFMM_MAX_TERMS_3D; ++t) {
                const int key = ((l - 1) * 32 + t) * 8;
                if (O.seg[key + 8] == O.seg[key]) continue;
                const DevTables* DB = nullptr;
                if ((rc = get_tables(ctx, t, &DB)) != LFMM_OK) return rc;
                if ((flags & LFMM_FLAG_MIXED) && (rc = get_m16(ctx, t, &DB)) != LFMM_OK) return rc;
            }
    }
    const int NX = tri(N);
    const double t_tree = timed ? ms_since(T0) : 0.0;

    // ---- expansions
    const DevTables* DT = tables;
    if (!DT) {
        rc = get_tables(ctx, N, &DT);
        if (rc != LFMM_OK) return rc;
    }
    std::vector<double2*> outer(depth, nullptr), inner(depth, nullptr);
    const bool far = depth >= first + 1;
    size_t nev = 0;
    int m2l_launches = 0, mixed_launches = 0;
    // per-kernel timing (LFMM_FLAG_TIMINGS): event pairs on the solve's
    // stream around each launch, summed per kind
    enum { EV_M2L64, EV_M2L16, EV_L2L, EV_M2M, EV_P2M, EV_P2L, EV_L2P, EV_P2P, EV_KINDS };
    std::vector<int> ev_kind;  // per event pair
    auto ev_open = [&]() -> hipError_t { return timed ? hipEventRecord(ctx_event(ctx, nev++), st) : hipSuccess; };
    auto ev_close = [&](int kind) -> hipError_t {
        if (!timed) return hipSuccess;
        ev_kind.push_back(kind);
        return hipEventRecord(ctx_event(ctx, nev++), st);
    };
    const bool mixed = far && (flags & LFMM_FLAG_MIXED) && depth - 1 >= mfl;
    if (mixed && tables) return LFMM_ERR_UNSUPPORTED;  // caller-placed tables hold no matrix-core operators
    uint4* xp = nullptr;
    uint32_t* xmax = nullptr;
    std::vector<uint4*> lpack;  // per-level packed sources (uniform tree), else empty
    if (mixed) {
        rc = get_m16(ctx, N, &DT);
        if (rc != LFMM_OK) return rc;
        // packed sources of the largest mixed level: box-major (size + 1
        // boxes) or class-major (8 classes x (parents + 1) slots)
        int64_t big = 0;
        for (int l = mfl; l < depth; ++l)
            big = std::max(big, m16_rowfull(N) ? 8 * (cnt[l - 1] + 1) : cnt[l] + 1);
        const int64_t nr = (int64_t)N * N, pieces = (nr + 15) / 16 * 4;
        xp = take<uint4>(A, big * pieces);
        xmax = take<uint32_t>(A, depth);
        if (!xp || !xmax) return LFMM_ERR_NOMEM;
        LFMM_HIP(fill_u32(xmax, depth, 0u, st));
        if (!adaptive && m16_rowfull(N)) {
            // uniform tree: every matrix-core level packed once, when its
            // multipoles are final, for its M2M into the level above and its
            // own M2L
            lpack.assign(depth, nullptr);
            for (int l = mfl; l < depth; ++l) {
                lpack[l] = take<uint4>(A, 8 * (cnt[l - 1] + 1) * pieces);
                if (!lpack[l]) return LFMM_ERR_NOMEM;
            }
        }
    }
    auto pack_level = [&](int l) -> hipError_t {
        M2L16Args mp{};
        mp.outer = outer[l];
        mp.num_parents = (int)cnt[l - 1];
        mp.child_begin = lbegin[l - 1];
        mp.child_keys = lkeys[l];
        mp.xmax = xmax + l;
        mp.xp = lpack[l];
        mp.size = (int)cnt[l];
        return launch_m2l16_pack(N, mp, st);
    };
    // adaptive octree: the expansions of all boxes in one array each, by gid
    double2* outer_all = nullptr;
    double2* inner_all = nullptr;
    if (far) {
        if (adaptive) {
            outer_all = take<double2>(A, (int64_t)T.total() * NX);
            inner_all = take<double2>(A, (int64_t)T.total() * NX);
            if (!outer_all || !inner_all) return LFMM_ERR_NOMEM;
        }
        for (int l = 2; l < depth; ++l) {
            if (adaptive) {
                outer[l] = outer_all + (int64_t)T.loff[l] * NX;
                inner[l] = inner_all + (int64_t)T.loff[l] * NX;
                continue;
            }
            outer[l] = take<double2>(A, cnt[l] * NX);
            inner[l] = take<double2>(A, cnt[l] * NX);
            if (!outer[l] || !inner[l]) return LFMM_ERR_NOMEM;
        }
        P2MArgs pa{src, lbegin[leaf_level], lkeys[leaf_level], (int)cnt[leaf_level], leaf_level, geom, DT->knm,
                   outer[leaf_level]};
        // the matrix-core levels need their largest |coefficient|: the producers max-combine it
        if (mixed && leaf_level >= mfl) pa.absmax = xmax + leaf_level;
        if (adaptive) {  // every leaf of the global list, at its level
            pa.leaf_begin = T.leaf_begin;
            pa.leaf_key = T.leaf_key;
            pa.nleaf = (int)T.nleaf;
            pa.outer = outer_all;
            pa.leaf_gid = T.leaf_gid;
            pa.leaf_lvl = T.leaf_lvl;
            pa.absmax = mixed ? xmax : nullptr;
        }
        LFMM_HIP(ev_open());
        LFMM_HIP(launch_p2m(N, pa, st));
        LFMM_HIP(ev_close(EV_P2M));
        if (!lpack.empty() && leaf_level >= mfl) LFMM_HIP(pack_level(leaf_level));
        for (int l = leaf_level - 1; l >= first; --l) {
            if (!lpack.empty() && l >= mfl) {
                // M2M into a matrix-core level on the matrix cores: the 8
                // child classes as the directions, the children's packing
                M2L16Args mm{};
                mm.kmat = DT->m2m16;
                mm.kpre = DT->m2m16_kpre;
                mm.kinv = DT->m2m16_kinv;
                mm.xp = lpack[l + 1];
                mm.xmax = xmax + l + 1;
                mm.class_dirs = DT->rt.class_dirs;
                mm.dir_vec = DT->rt.dir_vec;
                mm.level = l;
                mm.num_parents = (int)cnt[l - 1];
                mm.child_begin = lbegin[l - 1];
                mm.child_keys = lkeys[l];
                mm.lut = lut[l];
                mm.keys = lkeys[l];
                mm.plut = lut[l - 1];
                mm.pkeys = lkeys[l - 1];
                mm.size = (int)cnt[l];
                mm.dst = outer[l];
                mm.tiles = (int)((cnt[l - 1] + m16_targets(N) - 1) / m16_targets(N));
                mm.m2m = 1;
                mm.src_np1 = cnt[l] + 1;
                mm.absmax = xmax + l;
                LFMM_HIP(ev_open());
                LFMM_HIP(launch_m2l16(N, mm, st));
                LFMM_HIP(ev_close(EV_M2M));
                LFMM_HIP(pack_level(l));
                continue;
            }
            RtrLevel L{};
            L.level = l;
            L.num_parents = (int)cnt[l];
            L.child_begin = lbegin[l];
            L.child_keys = lkeys[l + 1];
            L.src = outer[l + 1];
            L.dst = outer[l];
            if (mixed && l >= mfl) L.absmax = xmax + l;
            // small uniform levels: the 8 child classes over 8 parts, summed
            // after (a few tiles alone leave the chip idle)
            if (!adaptive && !L.absmax && (cnt[l] + rtr_tile(N) - 1) / rtr_tile(N) < 128) {
                L.dsplit = 8;
                L.size = (int)cnt[l];
                L.partial = take<double2>(A, 8 * cnt[l] * NX);
                if (!L.partial) return LFMM_ERR_NOMEM;
            }
            LFMM_HIP(ev_open());
            LFMM_HIP(launch_rtr(N, 0, DT->rt, L, st));
            if (L.dsplit > 1) LFMM_HIP(launch_sum_parts(L.partial, L.dsplit, cnt[l] * NX, outer[l], false, st));
            LFMM_HIP(ev_close(EV_M2M));
        }
    }
    double t_up = 0.0;
    if (timed) {
        LFMM_HIP(hipStreamSynchronize(st));
        t_up = ms_since(T0) - t_tree;
    }
    // direction parts: levels with fewer workgroups than the chip holds split
    // their 189 directions over more workgroups (deterministic partial sums)
    auto dsplit_for = [](int64_t wgs) { return (int)std::max<int64_t>(1, std::min<int64_t>(32, (1024 + wgs - 1) / wgs)); };
    double2* list_partial = nullptr;  // direction parts of the matrix-core M2L buckets
    int64_t list_partial_elems = 0;
    if (far) {
        for (int l = first; l < depth; ++l) {
            RtrLevel L{};
            L.level = l;
            L.num_parents = (int)cnt[l - 1];
            L.child_begin = lbegin[l - 1];
            L.child_keys = lkeys[l];
            L.lut = lut[l];
            L.keys = lkeys[l];
            L.size = (int)cnt[l];
            L.src = outer[l];
            L.parent_local = (l > first) ? inner[l - 1] : nullptr;
            L.dst = inner[l];
            L.tiles = (int)((cnt[l - 1] + rtr_tile(N) - 1) / rtr_tile(N));
            const bool bucketed = adaptive && l >= mfl && (per_box || cnt[l - 1] >= kBucketMinParents);
            if (bucketed) {
                // the level's sources packed once, at the largest order of its buckets
                int tpack = 0;
                for (int t = 1; t <= LFMM_MAX_TERMS_3D; ++t)
                    if (O.seg[((l - 1) * 32 + t) * 8 + 8] > O.seg[((l - 1) * 32 + t) * 8]) tpack = t;
                if (mixed && tpack > 0) {
                    M2L16Args mp{};
                    mp.outer = outer[l];
                    mp.num_parents = (int)cnt[l - 1];
                    mp.child_begin = L.child_begin;
                    mp.child_keys = L.child_keys;
                    mp.xmax = xmax + l;
                    mp.xp = xp;
                    mp.stride = NX;
                    mp.size = L.size;
                    LFMM_HIP(launch_m2l16_pack(tpack, mp, st));
                }
                // one launch per order among the level's boxes ((parent, class) lists)
                for (int t = 1; t <= LFMM_MAX_TERMS_3D; ++t) {
                    const int key = ((l - 1) * 32 + t) * 8;
                    const int64_t b0 = O.seg[key], b1 = O.seg[key + 8];
                    if (b1 == b0) continue;
                    const DevTables* DB = &ctx->tables.at(t);
                    // the bucket's parents, one per child, sorted by the child's class
                    RtrLevel Lb = L;
                    Lb.plist = O.plist + b0;
                    Lb.num_parents = (int)(b1 - b0);
                    Lb.stride = NX;
                    Lb.dsplit = 1;
                    Lb.csorted = 1;
                    for (int c = 0; c <= 8; ++c) Lb.coff[c] = (int)(O.seg[key + c] - b0);
                    Lb.ctile[0] = 0;
                    for (int c = 0; c < 8; ++c)
                        Lb.ctile[c + 1] = Lb.ctile[c] + (Lb.coff[c + 1] - Lb.coff[c] + rtr_tile(t) - 1) / rtr_tile(t);
                    Lb.tiles = Lb.ctile[8];
                    if (mixed) {
                        M2L16Args m{};
                        m.kmat = DB->m16;
                        m.kpre = DB->m16_kpre;
                        m.outer = outer[l];
                        m.xp = xp;
                        m.xmax = xmax + l;
                        m.class_dirs = DB->rt.class_dirs;
                        m.dir_vec = DB->rt.dir_vec;
                        m.level = l;
                        m.num_parents = (int)cnt[l - 1];
                        m.child_begin = L.child_begin;
                        m.child_keys = L.child_keys;
                        m.lut = L.lut;
                        m.keys = L.keys;
                        m.plut = lut[l - 1];
                        m.pkeys = lkeys[l - 1];
                        m.size = L.size;
                        m.kinv = DB->m16_kinv;
                        m.dst = inner[l];
                        m.plist = Lb.plist;
                        m.plist_count = Lb.num_parents;
                        m.stride = NX;
                        m.dsplit = 1;
                        m.csorted = 1;
                        for (int c = 0; c <= 8; ++c) m.coff[c] = Lb.coff[c];
                        m.ctile[0] = 0;
                        for (int c = 0; c < 8; ++c)
                            m.ctile[c + 1] = m.ctile[c] + (m.coff[c + 1] - m.coff[c] + m16_targets(t) - 1) / m16_targets(t);
                        m.tiles = m.ctile[8];
                        m.xpieces = 4 * ((tpack * tpack + 15) / 16);
                        // a workgroup runs all 189 directions of its targets:
                        // small buckets split the directions into parts so
                        // that about 512 workgroups share the work
                        m.dsplit = (int)std::max<int64_t>(1, std::min<int64_t>(32, (512 + m.tiles - 1) /
                                                                                      std::max(1, m.tiles)));
                        if (m.dsplit > 1) {
                            // one buffer serves every bucket (the launches are ordered on the stream)
                            const int64_t need = (int64_t)m.dsplit * m.plist_count * tri(t);
                            if (need > list_partial_elems) {
                                list_partial = take<double2>(A, need);
                                if (!list_partial) return LFMM_ERR_NOMEM;
                                list_partial_elems = need;
                            }
                            m.partial = list_partial;
                        }
                        LFMM_HIP(ev_open());
                        LFMM_HIP(launch_rtr(t, 2, DB->rt, Lb, st));
                        LFMM_HIP(ev_close(EV_L2L));
                        LFMM_HIP(ev_open());
                        LFMM_HIP(launch_m2l16(t, m, st));
                        LFMM_HIP(launch_sum_list_parts(t, m, st));
                        LFMM_HIP(ev_close(EV_M2L16));
                        ++mixed_launches;
                    } else {
                        LFMM_HIP(ev_open());
                        LFMM_HIP(launch_rtr(t, 1, DB->rt, Lb, st));
                        LFMM_HIP(ev_close(EV_M2L64));
                    }
                    ++m2l_launches;
                }
            } else if (mixed && l >= mfl) {
                // FP64 L2L into the level, then += the matrix-core M2L
                M2L16Args m{};
                m.kmat = DT->m16;
                m.kpre = DT->m16_kpre;
                m.outer = outer[l];
                m.xp = lpack.empty() ? xp : lpack[l];
                m.xmax = xmax + l;
                m.class_dirs = DT->rt.class_dirs;
                m.dir_vec = DT->rt.dir_vec;
                m.level = l;
                m.num_parents = L.num_parents;
                m.child_begin = L.child_begin;
                m.child_keys = L.child_keys;
                m.lut = L.lut;
                m.keys = L.keys;
                m.plut = lut[l - 1];
                m.pkeys = lkeys[l - 1];
                m.size = L.size;
                m.kinv = DT->m16_kinv;
                m.dst = inner[l];
                m.tiles = (int)((cnt[l - 1] + m16_targets(N) - 1) / m16_targets(N));
                // the row-complete kernel holds one workgroup per CU: about
                // 256 workgroups, one wave of them (measured 0.13 ms faster
                // over levels 4-5 at 13 terms than 512: fewer parts to sum)
                m.dsplit = m16_rowfull(N) ? (int)std::max<int64_t>(1, std::min<int64_t>(32, (256 + m.tiles * 8 - 1) /
                                                                                                 (m.tiles * 8)))
                                          : dsplit_for((int64_t)m.tiles * 8);
                if (m.dsplit > 1) {
                    // every part writes its whole slice (no clearing needed)
                    m.partial = take<double2>(A, (int64_t)m.dsplit * cnt[l] * NX);
                    if (!m.partial) return LFMM_ERR_NOMEM;
                }
                if (lpack.empty()) LFMM_HIP(launch_m2l16_pack(N, m, st));  // else packed in the upward pass
                LFMM_HIP(ev_open());
                LFMM_HIP(launch_rtr(N, 2, DT->rt, L, st));
                LFMM_HIP(ev_close(EV_L2L));
                LFMM_HIP(ev_open());
                LFMM_HIP(launch_m2l16(N, m, st));
                if (m.dsplit > 1) LFMM_HIP(launch_sum_parts(m.partial, m.dsplit, cnt[l] * NX, inner[l], true, st));
                LFMM_HIP(ev_close(EV_M2L16));
                ++mixed_launches;
            } else {
                L.dsplit = dsplit_for((int64_t)L.tiles * 8);
                if (L.dsplit > 1) {
                    L.partial = take<double2>(A, (int64_t)L.dsplit * cnt[l] * NX);
                    if (!L.partial) return LFMM_ERR_NOMEM;
                }
                LFMM_HIP(ev_open());
                LFMM_HIP(launch_rtr(N, 1, DT->rt, L, st));
                if (L.dsplit > 1) LFMM_HIP(launch_sum_parts(L.partial, L.dsplit, cnt[l] * NX, inner[l], false, st));
                LFMM_HIP(ev_close(EV_M2L64));
            }
            if (!bucketed) ++m2l_launches;
            if (adaptive && T.xseg[l + 1] > T.xseg[l]) {
                // X lists of the level: source particles into the locals (P2L)
                P2MArgs px{src, nullptr, T.keys_all, (int)(T.xseg[l + 1] - T.xseg[l]), l, geom, DT->knm, inner_all};
                px.unit_box = T.xlist + T.xseg[l];
                px.x_off = T.x_off;
                px.x_rng = T.x_rng;
                px.g2 = T.loff[2];
                LFMM_HIP(ev_open());
                LFMM_HIP(launch_p2l(N, px, st));
                LFMM_HIP(ev_close(EV_P2L));
            }
        }
    }
    double t_down = 0.0;
    if (timed) {
        LFMM_HIP(hipStreamSynchronize(st));
        t_down = ms_since(T0) - t_tree - t_up;
    }

    // ---- leaves
    double* dpot = pot_out;
    double* dfield = field_out;
    if (memory == LFMM_MEM_HOST) {
        dpot = pot_out ? take<double>(A, n_out) : nullptr;
        dfield = field_out ? take<double>(A, 3 * n_out) : nullptr;
        if ((pot_out && !dpot) || (field_out && !dfield)) return LFMM_ERR_NOMEM;
    }
    if (learned) {  // the sampled check reads both outputs
        if (!dpot && !(dpot = take<double>(A, n_out))) return LFMM_ERR_NOMEM;
        if (!dfield && !(dfield = take<double>(A, 3 * n_out))) return LFMM_ERR_NOMEM;
    }
    LeafArgs la{};
    la.src = src;
    la.leaf_of = leaf_of;
    la.leaf_key = lkeys[leaf_level];
    la.leaf_begin = lbegin[leaf_level];
    la.nleaf = (int)nleaf;
    la.level = leaf_level;
    la.geom = geom;
    la.lut = leaf_level <= lut_max_level(3) ? (leaf_level >= 2 ? lut[leaf_level] : nullptr) : nullptr;
    la.local = far ? inner[leaf_level] : nullptr;
    la.knm = DT->knm;
    la.perm = order;
    if (adaptive) {
        la.leaf_key = T.leaf_key;
        la.leaf_begin = T.leaf_begin;
        la.local = inner_all;
        la.nbr = T.n_rng;
        la.nbr_off = T.n_off;
        la.box_range = T.range_all;
        la.direct_max = kDirectMax;
        la.leaf_lvl = T.leaf_lvl;
        la.leaf_gid = T.leaf_gid;
        la.w_off = T.w_off;
        la.w_gid = T.w_gid;
        la.keys_all = T.keys_all;
        la.outer_all = outer_all;
        la.depth = depth;
        for (int l = 0; l <= depth; ++l) la.loff[l] = T.loff[l];
        la.local_min = first;
        if (fs) {
            double* ls = take<double>(A, T.nleaf);
            if (!ls) return LFMM_ERR_NOMEM;
            if ((rc = forest_leaf_scale(ctx, T, *fs, ls)) != LFMM_OK) return rc;
            la.leaf_scale = ls;
        }
    } else {
        int2* nbr = take<int2>(A, cnt[leaf_level] * 27);
        if (!nbr) return LFMM_ERR_NOMEM;
        LFMM_HIP(launch_leaf_ranges(lkeys[leaf_level], lbegin[leaf_level], (int)cnt[leaf_level], leaf_level, la.lut, nbr, st));
        la.nbr = nbr;
    }
    la.n = n;
    la.n_out = n_out;
    la.exclude_self = (flags & LFMM_FLAG_EXCLUDE_SELF) ? 1 : 0;
    la.pot = dpot;
    la.field = dfield;
    if (far) {  // L2P and the pair loop as two kernels (see k_l2p)
        la.far = take<double4>(A, n);
        if (!la.far) return LFMM_ERR_NOMEM;
    }
    int32_t* p2p_ws = take<int32_t>(A, p2p_ws_elems(n, nleaf));
    uint32_t* p2p_scan = take<uint32_t>(A, scan_ws_elems(nleaf));
    if (!p2p_ws || !p2p_scan) return LFMM_ERR_NOMEM;
    if (far) {
        LFMM_HIP(ev_open());
        LFMM_HIP(launch_l2p(N, la, st));
        LFMM_HIP(ev_close(EV_L2P));
    }
    LFMM_HIP(ev_open());
    LFMM_HIP(launch_p2p(la, p2p_ws, p2p_scan, st));  // near field, several targets per lane
    LFMM_HIP(ev_close(EV_P2P));
    // learned orders: the sampled direct-sum check judges the run
    double chk_p = 0.0, chk_f = 0.0;
    if (learned) {
        rc = order_check_finish(ctx, chk, dpot, dfield, !(flags & LFMM_FLAG_EXCLUDE_SELF), &chk_p, &chk_f);
        if (rc != LFMM_OK) return rc;
    }
    const bool fail_check = learned && !(chk_p <= ctx->order_tol);
    if (memory == LFMM_MEM_HOST && !fail_check) {
        if (pot_out) LFMM_HIP(hipMemcpyAsync(pot_out, dpot, sizeof(double) * n_out, hipMemcpyDeviceToHost, st));
        if (field_out) LFMM_HIP(hipMemcpyAsync(field_out, dfield, sizeof(double) * 3 * n_out, hipMemcpyDeviceToHost, st));
    }
    if (timed || memory == LFMM_MEM_HOST) LFMM_HIP(hipStreamSynchronize(st));
    LFMM_HIP(hipGetLastError());
    if (rep) {
        rep->terms = N;
        rep->tree_depth = depth;
        rep->num_leaves = nleaf;
        rep->num_boxes = 0;
        for (int l = 0; l < depth; ++l) rep->num_boxes += cnt[l];
        rep->xw_pairs = adaptive ? T.x_pairs : 0;
        rep->max_leaf = max_leaf;
        rep->m2l_ops = -1;
        rep->m2l_ops_mixed = -1;
        rep->p2p_pairs = -1;
        rep->m2l_launches = m2l_launches;
        rep->m2l_mixed_launches = mixed_launches;
        rep->terms_min = per_box ? O.min_leaf_terms : N;
        rep->fallback = 0;
        rep->check_rel_l2 = chk_p;
        rep->check_rel_l2_field = chk_f;
        rep->m2l_flops = -1.0;
        if (timed) {
            double tk[EV_KINDS] = {};
            for (size_t k = 0; k < ev_kind.size(); ++k) {
                float ms = 0.f;
                LFMM_HIP(hipEventElapsedTime(&ms, ctx->events[2 * k], ctx->events[2 * k + 1]));
                tk[ev_kind[k]] += ms;
            }
            rep->t_m2l_mixed_ms = tk[EV_M2L16];
            rep->t_m2l_kernels_ms = tk[EV_M2L64] + tk[EV_M2L16];
            rep->t_leaf_kernel_ms = tk[EV_L2P] + tk[EV_P2P];
            rep->t_p2m_ms = tk[EV_P2M];
            rep->t_m2m_ms = tk[EV_M2M];
            rep->t_l2l_ms = tk[EV_L2L];
            rep->t_p2l_ms = tk[EV_P2L];
            rep->t_l2p_ms = tk[EV_L2P];
            rep->t_p2p_ms = tk[EV_P2P];
        }
        // algorithmic HBM bytes of the M2M and standalone L2L launches: every
        // expansion read once and written once (16 bytes per coefficient)
        rep->m2m_bytes = 0.0;
        rep->l2l_bytes = 0.0;
        if (far) {
            for (int l = first; l + 1 < depth; ++l) rep->m2m_bytes += 16.0 * NX * (double)(cnt[l + 1] + cnt[l]);
            if (mixed)
                for (int l = mfl; l < depth; ++l) rep->l2l_bytes += 16.0 * NX * (double)(cnt[l - 1] + cnt[l]);
        }
        rep->t_tree_ms = t_tree;
        rep->t_upward_ms = t_up;
        rep->t_downward_ms = t_down;
        rep->t_total_ms = timed ? ms_since(T0) : 0.0;
        rep->t_leaf_ms = timed ? rep->t_total_ms - t_tree - t_up - t_down : 0.0;
        if (flags & LFMM_FLAG_COUNT_OPS) {  // exact op counts, after the solve
            unsigned long long* cws = take<unsigned long long>(A, 2 + depth);
            if (!cws) return LFMM_ERR_NOMEM;
            std::vector<int64_t> per(depth, 0);
            LFMM_HIP(count_ops(3, lkeys, lbegin, lut, cnt, cws, &rep->m2l_ops, &rep->p2p_pairs, per.data(), st,
                               &rep->p2p_source_reads));
            if (adaptive) {
                rc = ad_count_pairs(T.leaf_begin, T.n_off, T.n_rng, T.nleaf, cws, &rep->p2p_pairs,
                                    &rep->p2p_source_reads, st);
                if (rc != LFMM_OK) return rc;
            }
            rep->m2l_ops_mixed = 0;
            if (mixed)
                for (int l = mfl; l < depth; ++l) rep->m2l_ops_mixed += per[l];
            if (fs) {  // a forest's solve translates within its systems only
                rep->m2l_ops = 0;
                for (int l = first; l < depth; ++l) rep->m2l_ops += per[l];
            }
            rep->m2l_flops = (double)rep->m2l_ops * (double)m2l_flops_per_op(N);
            if (per_box) {
                rc = order_m2l_flops(ctx, T, O, N, mfl, &rep->m2l_flops);
                if (rc != LFMM_OK) return rc;
            }
        }
    }
    return fail_check ? kNeedFallback : LFMM_OK;
}

bool valid_points(const lfmm_points3d* p, bool need_q, int dim)
{
    if (!p || p->n < 0) return false;
    if (p->n == 0) return true;
    if (!p->pos || p->pos_stride < dim) return false;
    if (need_q && (!p->q || p->q_stride < 1)) return false;
    return p->n < ((int64_t)1 << 31);
}

// Argument checks and the trivial cases of lfmm_fmm3d (nothing to evaluate;
// no sources: zero outputs).  Returns kNeedSolve when a solve must run, with
// *tg the target view (null for a single-set solve), else the final status.
int fmm3d_prologue(lfmm_ctx* c, const lfmm_points3d* sources, const lfmm_points3d* targets, double eps, int terms,
                   int tree_depth, int memory, double* potential, double* field, lfmm_report* report,
                   const lfmm_points3d** tg_out)
{
    if (!c || !valid_points(sources, true, 3)) return LFMM_ERR_INVALID;
    if (memory != LFMM_MEM_HOST && memory != LFMM_MEM_DEVICE) return LFMM_ERR_INVALID;
    if (terms != LFMM_AUTO && terms < 1) return LFMM_ERR_TERMS;
    if (tree_depth != LFMM_AUTO && (tree_depth < 1 || tree_depth > kMaxDepth)) return LFMM_ERR_DEPTH;
    if (!(eps > 0.0) && terms == LFMM_AUTO) return LFMM_ERR_INVALID;
    if (targets && !valid_points(targets, false, 3)) return LFMM_ERR_INVALID;
    const lfmm_points3d* tg = (targets && targets->n > 0) ? targets : nullptr;
    if (tg && sources->n + tg->n >= ((int64_t)1 << 31)) return LFMM_ERR_INVALID;
    if (report) std::memset(report, 0, sizeof(*report));
    *tg_out = tg;
    if ((tg ? tg->n : sources->n) == 0) return LFMM_OK;
    if (sources->n == 0) {  // no sources: every output is zero
        std::lock_guard<std::mutex> g(c->mu);
        if (hipSetDevice(c->device) != hipSuccess) return LFMM_ERR_HIP;
        const size_t nt = (size_t)tg->n;
        if (memory == LFMM_MEM_HOST) {
            if (potential) std::memset(potential, 0, nt * sizeof(double));
            if (field) std::memset(field, 0, 3 * nt * sizeof(double));
        } else {
            if (potential && fill_u32((uint32_t*)potential, 2 * (int64_t)nt, 0u, c->stream) != hipSuccess)
                return LFMM_ERR_HIP;
            if (field && fill_u32((uint32_t*)field, 6 * (int64_t)nt, 0u, c->stream) != hipSuccess) return LFMM_ERR_HIP;
        }
        return LFMM_OK;
    }
    return kNeedSolve;
}

static int tree_export(lfmm_ctx* c, const lfmm_points3d* particles, int dim, int tree_depth, int leaf_max,
                       int32_t* depth_out, int64_t* level_sizes, double* geom, int64_t* pos_map, uint64_t* keys,
                       int64_t* begins)
{
    if (!c || !valid_points(particles, false, dim) || particles->n == 0) return LFMM_ERR_INVALID;
    if (tree_depth != LFMM_AUTO && (tree_depth < 1 || tree_depth > (dim == 3 ? kMaxDepth : 2 * kMaxDepth - 1)))
        return LFMM_ERR_DEPTH;
    std::lock_guard<std::mutex> g(c->mu);
    if (hipSetDevice(c->device) != hipSuccess) return LFMM_ERR_HIP;
    try {
        c->arena.reset();
        Tree T;
        DevPoints DP;
        int rc = upload_view(c, particles, dim, DP);
        if (rc != LFMM_OK) return rc;
        rc = build_tree(c, DP, tree_depth, leaf_max, true, T);
        if (rc != LFMM_OK) return rc;
        hipStream_t st = c->stream;
        const int64_t n = T.n;
        if (depth_out) *depth_out = T.depth;
        if (level_sizes)
            for (int l = 0; l < T.depth; ++l) level_sizes[l] = T.cnt[l];
        if (geom) LFMM_HIP(hipMemcpyAsync(geom, T.geom, 4 * sizeof(double), hipMemcpyDeviceToHost, st));
        if (pos_map) {
            std::vector<uint32_t> ord(n);
            LFMM_HIP(hipMemcpyAsync(ord.data(), T.order, n * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
            LFMM_HIP(hipStreamSynchronize(st));
            for (int64_t s = 0; s < n; ++s) pos_map[ord[s]] = s;
        }
        int64_t ko = 0, bo = 0;
        for (int l = 0; l < T.depth; ++l) {
            if (keys) LFMM_HIP(hipMemcpyAsync(keys + ko, T.lkeys[l], T.cnt[l] * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
            if (begins) {
                std::vector<int32_t> b(T.cnt[l] + 1);
                LFMM_HIP(hipMemcpyAsync(b.data(), T.lbegin[l], (T.cnt[l] + 1) * sizeof(int32_t), hipMemcpyDeviceToHost, st));
                LFMM_HIP(hipStreamSynchronize(st));
                for (int64_t i = 0; i <= T.cnt[l]; ++i) begins[bo + i] = b[i];
            }
            ko += T.cnt[l];
            bo += T.cnt[l] + 1;
        }
        LFMM_HIP(hipStreamSynchronize(st));
        return LFMM_OK;
    } catch (const std::bad_alloc&) {
        return LFMM_ERR_NOMEM;
    } catch (...) {
        return LFMM_ERR_INVALID;
    }
}

template <typename T>
static int fetch(std::vector<T>& h, const T* d, int64_t count, hipStream_t st)
{
    h.resize((size_t)std::max<int64_t>(count, 0));
    if (count > 0) LFMM_HIP(hipMemcpyAsync(h.data(), d, sizeof(T) * (size_t)count, hipMemcpyDeviceToHost, st));
    return LFMM_OK;
}

static int tree_export_adaptive(lfmm_ctx* c, const lfmm_points3d* particles, int tree_depth, int leaf_max,
                                int64_t* counts, int32_t* depth_out, double* geom, int32_t* box_level,
                                uint64_t* box_key, int32_t* box_range, int32_t* leaf_box, int64_t* u_off,
                                int32_t* u_leaf, int64_t* x_off, int32_t* x_leaf, int64_t* w_off, int32_t* w_box)
{
    if (!c || !valid_points(particles, false, 3) || particles->n == 0 || !counts) return LFMM_ERR_INVALID;
    if (tree_depth != LFMM_AUTO && (tree_depth < 1 || tree_depth > kMaxDepth)) return LFMM_ERR_DEPTH;
    if (leaf_max <= 0) leaf_max = 64;
    std::lock_guard<std::mutex> g(c->mu);
    if (hipSetDevice(c->device) != hipSuccess) return LFMM_ERR_HIP;
    try {
        c->arena.reset();
        Tree T;
        DevPoints DP;
        int rc = upload_view(c, particles, 3, DP);
        if (rc != LFMM_OK) return rc;
        rc = build_tree(c, DP, tree_depth, leaf_max, false, T, true);
        if (rc != LFMM_OK) return rc;
        rc = build_adaptive(c, leaf_max, T);
        if (rc != LFMM_OK) return rc;
        hipStream_t st = c->stream;
        const int total = T.total(), g2 = T.depth > 2 ? T.loff[2] : total;
        std::vector<uint64_t> keys;
        std::vector<int2> rng, urng, xrng;
        std::vector<int32_t> lb, lg, wg, b2l;
        std::vector<uint32_t> uo, xo, wo;
        if ((rc = fetch(keys, T.keys_all, total, st)) || (rc = fetch(rng, T.range_all, total, st)) ||
            (rc = fetch(lb, T.leaf_begin, T.nleaf + 1, st)) || (rc = fetch(lg, T.leaf_gid, T.nleaf, st)) ||
            (rc = fetch(uo, T.u_off, T.nleaf + 1, st)) || (rc = fetch(urng, T.u_rng, T.u_total, st)) ||
            (rc = fetch(wo, T.w_off, T.nleaf + 1, st)) || (rc = fetch(wg, T.w_gid, T.x_pairs, st)) ||
            (rc = fetch(b2l, T.box2leaf, total, st)))
            return rc;
        if (total > g2) {
            if ((rc = fetch(xo, T.x_off, total - g2 + 1, st)) || (rc = fetch(xrng, T.x_rng, T.x_pairs, st))) return rc;
        }
        LFMM_HIP(hipStreamSynchronize(st));
        counts[0] = total;
        counts[1] = T.nleaf;
        counts[2] = T.u_total;
        counts[3] = T.x_pairs;
        if (depth_out) *depth_out = T.depth;
        if (geom) LFMM_HIP(hipMemcpy(geom, T.geom, 4 * sizeof(double), hipMemcpyDeviceToHost));
        // a particle range of a leaf -> the leaf's number
        auto leaf_at = [&](int b) {
            return (int32_t)(std::upper_bound(lb.begin(), lb.end() - 1, b) - lb.begin() - 1);
        };
        for (int gi = 0; gi < total; ++gi) {
            int l = 0;
            while (l + 1 < T.depth && gi >= T.loff[l + 1]) ++l;
            if (box_level) box_level[gi] = l;
            if (box_key) box_key[gi] = keys[gi];
            if (box_range) {
                box_range[2 * gi] = rng[gi].x;
                box_range[2 * gi + 1] = rng[gi].y;
            }
        }
        for (int64_t i = 0; i < T.nleaf; ++i) {
            if (leaf_box) leaf_box[i] = lg[i];
            if (u_off) u_off[i] = uo[i];
            if (w_off) w_off[i] = wo[i];
        }
        if (u_off) u_off[T.nleaf] = uo[T.nleaf];
        if (w_off) w_off[T.nleaf] = wo[T.nleaf];
        if (u_leaf)
            for (int64_t k = 0; k < T.u_total; ++k) u_leaf[k] = leaf_at(urng[k].x);
        if (w_box)
            for (int64_t k = 0; k < T.x_pairs; ++k) w_box[k] = wg[k];
        if (x_off) {
            for (int gi = 0; gi <= total; ++gi) x_off[gi] = gi < g2 ? 0 : (int64_t)xo[gi - g2];
        }
        if (x_leaf)
            for (int64_t k = 0; k < T.x_pairs; ++k) x_leaf[k] = leaf_at(xrng[k].x);
        return LFMM_OK;
    } catch (const std::bad_alloc&) {
        return LFMM_ERR_NOMEM;
    } catch (...) {
        return LFMM_ERR_INVALID;
    }
}

}  // namespace lfmm::host

using namespace lfmm;
using namespace lfmm::host;

extern "C" {

int lfmm_create(lfmm_ctx** out, int device)
{
    if (!out) return LFMM_ERR_INVALID;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return LFMM_ERR_NODEVICE;
    if (device < 0 || device >= count) return LFMM_ERR_NODEVICE;
    if (hipSetDevice(device) != hipSuccess) return LFMM_ERR_HIP;
    lfmm_ctx* c = new (std::nothrow) lfmm_ctx();
    if (!c) return LFMM_ERR_NOMEM;
    c->device = device;
    // this device's address of the allocation-free entry point's scratch
    if (noalloc_scratch_init(device) != hipSuccess) {
        delete c;
        return LFMM_ERR_HIP;
    }
    if (hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return LFMM_ERR_HIP;
    }
    c->stream = c->own_stream;
    *out = c;
    return LFMM_OK;
}

void lfmm_destroy(lfmm_ctx* c)
{
    if (!c) return;
    for (lfmm_ctx* w : c->workers) lfmm_destroy(w);
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    c->arena.release();
    for (auto& kv : c->tables) {
        (void)hipFree(kv.second.block);
        (void)hipFree(kv.second.m16);
    }
    free_tables2d(c);
    (void)hipFree(c->vec_tmp);
    (void)hipFree(c->forest_buf);
    if (c->switch_event) (void)hipEventDestroy(c->switch_event);
    for (auto e : c->events) (void)hipEventDestroy(e);
    if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
    if (c->side_stream) (void)hipStreamSynchronize(c->side_stream);
    for (hipEvent_t e : c->side_events)
        if (e) (void)hipEventDestroy(e);
    if (c->side_stream) (void)hipStreamDestroy(c->side_stream);
    delete c;
}

const char* lfmm_strerror(int s)
{
    switch (s) {
    case LFMM_OK: return "ok";
    case LFMM_ERR_INVALID: return "invalid argument";
    case LFMM_ERR_TERMS: return "expansion terms out of range";
    case LFMM_ERR_DEPTH: return "tree depth out of range";
    case LFMM_ERR_NOMEM: return "out of device memory";
    case LFMM_ERR_HIP: return "HIP runtime error";
    case LFMM_ERR_NONFINITE: return "non-finite input";
    case LFMM_ERR_NODEVICE: return "no usable GPU";
    case LFMM_ERR_UNSUPPORTED: return "unsupported option";
    case LFMM_ERR_TOO_LARGE: return "input larger than the fixed scratch of an allocation-free entry point";
    default: return "unknown status";
    }
}

int lfmm_set_stream(lfmm_ctx* c, void* s)
{
    if (!c) return LFMM_ERR_INVALID;
    std::lock_guard<std::mutex> g(c->mu);
    // LFMM_STREAM_DEFAULT: the null stream (handle 0), which a NULL argument cannot name
    hipStream_t ns = s == LFMM_STREAM_DEFAULT ? (hipStream_t)0 : (s ? (hipStream_t)s : c->own_stream);
    if (ns == c->stream) return LFMM_OK;
    // work already queued on the old stream (a device-memory solve returns
    // before its kernels finish) comes before anything on the new one, which
    // reuses the same workspace
    if (hipSetDevice(c->device) != hipSuccess) return LFMM_ERR_HIP;
    if (!c->switch_event && hipEventCreateWithFlags(&c->switch_event, hipEventDisableTiming) != hipSuccess)
        return LFMM_ERR_HIP;
    if (hipEventRecord(c->switch_event, c->stream) != hipSuccess) return LFMM_ERR_HIP;
    if (hipStreamWaitEvent(ns, c->switch_event, 0) != hipSuccess) return LFMM_ERR_HIP;
    c->stream = ns;
    return LFMM_OK;
}

int lfmm_trim(lfmm_ctx* c)
{
    if (!c) return LFMM_ERR_INVALID;
    std::lock_guard<std::mutex> g(c->mu);
    (void)hipSetDevice(c->device);
    if (hipStreamSynchronize(c->stream) != hipSuccess) return LFMM_ERR_HIP;
    c->arena.release();
    return LFMM_OK;
}

int lfmm_fmm3d(lfmm_ctx* c, const lfmm_points3d* sources, const lfmm_points3d* targets, double eps, int terms,
               int tree_depth, int leaf_max, int memory, unsigned flags, double* potential, double* field,
               lfmm_report* report)
{
    const lfmm_points3d* tg = nullptr;
    const int pre = fmm3d_prologue(c, sources, targets, eps, terms, tree_depth, memory, potential, field, report, &tg);
    if (pre != kNeedSolve) return pre;
    std::lock_guard<std::mutex> g(c->mu);
    if (hipSetDevice(c->device) != hipSuccess) return LFMM_ERR_HIP;
    try {
        int rc = fmm3d_solve(c, sources, tg, eps, terms, tree_depth, leaf_max, memory, flags, potential, field, report,
                             nullptr);
        if (rc == kNeedFallback) {
            // the learned orders missed the sampled check: the fixed order
            // (p = 16) on the same adaptive tree
            lfmm_report keep{};
            if (report) keep = *report;
            rc = fmm3d_solve(c, sources, tg, eps, kOrderMinTerms + kOrderClasses - 1, tree_depth, leaf_max, memory,
                             (flags & ~LFMM_FLAG_LEARNED_ORDER) | LFMM_FLAG_ADAPTIVE, potential, field, report, nullptr);
            if (report) {
                report->fallback = 1;
                report->check_rel_l2 = keep.check_rel_l2;
                report->check_rel_l2_field = keep.check_rel_l2_field;
                report->terms_min = keep.terms_min;
            }
        }
        return rc;
    } catch (const std::bad_alloc&) {
        return LFMM_ERR_NOMEM;
    } catch (...) {
        return LFMM_ERR_INVALID;
    }
}

int lfmm_tree3d(lfmm_ctx* c, const lfmm_points3d* particles, int tree_depth, int leaf_max, int32_t* depth_out,
                int64_t* level_sizes, double* geom, int64_t* pos_map, uint64_t* keys, int64_t* begins)
{
    return tree_export(c, particles, 3, tree_depth, leaf_max, depth_out, level_sizes, geom, pos_map, keys, begins);
}

int lfmm_tree3d_adaptive(lfmm_ctx* c, const lfmm_points3d* particles, int tree_depth, int leaf_max, int64_t* counts,
                         int32_t* depth_out, double* geom, int32_t* box_level, uint64_t* box_key, int32_t* box_range,
                         int32_t* leaf_box, int64_t* u_off, int32_t* u_leaf, int64_t* x_off, int32_t* x_leaf,
                         int64_t* w_off, int32_t* w_box)
{
    return tree_export_adaptive(c, particles, tree_depth, leaf_max, counts, depth_out, geom, box_level, box_key,
                                box_range, leaf_box, u_off, u_leaf, x_off, x_leaf, w_off, w_box);
}

int lfmm_tree2d(lfmm_ctx* c, const lfmm_points2d* particles, int tree_depth, int leaf_max, int32_t* depth_out,
                int64_t* level_sizes, double* geom, int64_t* pos_map, uint64_t* keys, int64_t* begins)
{
    return tree_export(c, particles, 2, tree_depth, leaf_max, depth_out, level_sizes, geom, pos_map, keys, begins);
}

int lfmm_host_wigner_d(int j, double beta, double* out)
{
    if (j < 0 || j > 64 || !out) return LFMM_ERR_INVALID;
    host_wigner_small_d(j, beta, out);
    return LFMM_OK;
}

int64_t lfmm_host_op_tables(int N, int which, void* out, int64_t capacity)
{
    if (N < 1 || N > LFMM_MAX_TERMS_3D) return -(int64_t)LFMM_ERR_TERMS;
    try {
        OpTables T;
        build_op_tables(N, T);
        const void* p = nullptr;
        int64_t count = 0;
        size_t esz = 8;
        switch (which) {
        case 0: p = T.rot.data(); count = (int64_t)T.rot.size(); break;
        case 1: p = T.ephi.data(); count = (int64_t)T.ephi.size(); break;
        case 2: p = T.trans.data(); count = (int64_t)T.trans.size(); break;
        case 3: p = T.dir_theta.data(); count = (int64_t)T.dir_theta.size(); esz = 4; break;
        case 4: p = T.dir_trans.data(); count = (int64_t)T.dir_trans.size(); esz = 4; break;
        case 5: p = T.dir_vec.data(); count = (int64_t)T.dir_vec.size(); esz = 4; break;
        case 6: p = T.class_dirs.data(); count = (int64_t)T.class_dirs.size(); esz = 4; break;
        case 7: p = T.knm.data(); count = (int64_t)T.knm.size(); break;
        default: return -(int64_t)LFMM_ERR_INVALID;
        }
        if (out && capacity >= count) std::memcpy(out, p, esz * (size_t)count);
        return count;
    } catch (...) {
        return -(int64_t)LFMM_ERR_NOMEM;
    }
}

int lfmm_host_m2l_dense(int N, int dir, double* out)
{
    if (N < 1 || N > LFMM_MAX_TERMS_3D) return LFMM_ERR_TERMS;
    if (dir < 0 || dir >= kNumM2LDirs || !out) return LFMM_ERR_INVALID;
    try {
        OpTables T;
        build_op_tables(N, T);
        m2l_dense_matrix(T, dir, out);
        return LFMM_OK;
    } catch (...) {
        return LFMM_ERR_NOMEM;
    }
}

int lfmm_host_m2m_dense(int N, int cls, double* out)
{
    if (N < 1 || N > LFMM_MAX_TERMS_3D) return LFMM_ERR_TERMS;
    if (cls < 0 || cls >= 8 || !out) return LFMM_ERR_INVALID;
    try {
        OpTables T;
        build_op_tables(N, T);
        m2m_dense_matrix(T, cls, out);
        return LFMM_OK;
    } catch (...) {
        return LFMM_ERR_NOMEM;
    }
}

int64_t lfmm_host_m2l16_pack(int N, uint16_t* out, int64_t capacity, double* kinv, int32_t* kpre)
{
    if (N < 1 || N > LFMM_MAX_TERMS_3D) return -(int64_t)LFMM_ERR_TERMS;
    try {
        OpTables T;
        build_op_tables(N, T);
        std::vector<uint16_t> pk;
        std::vector<int> kp;
        double ki = 0.0;
        pack_m2l16(T, pk, ki, kp);
        if (kinv) *kinv = ki;
        if (kpre) std::memcpy(kpre, kp.data(), kp.size() * sizeof(int32_t));
        if (out && capacity >= (int64_t)pk.size()) std::memcpy(out, pk.data(), pk.size() * 2);
        return (int64_t)pk.size();
    } catch (...) {
        return -(int64_t)LFMM_ERR_NOMEM;
    }
}

const char* lfmm_version(void) { return "lfmm-mi355x 0.1.0 (gfx950)"; }

}  // extern "C"
