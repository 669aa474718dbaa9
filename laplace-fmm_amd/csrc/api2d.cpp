// api2d.cpp -- the 2D (complex-series) FMM through the C ABI: lfmm_fmm2d and
// lfmm_tree2d.  Same pipeline as the 3D solve on a quadtree:
//   extent -> keys -> radix sort -> levels -> P2M -> M2M (up)
//   -> L2L+M2L per level (down) -> L2P+P2P at the leaves, scattered to input order.
// Reference: src/cpu/multipole.cpp (Solve :473-520, parameter rules :478-487).
#include "api_internal.h"
#include "ops2d.h"

#include <cmath>
#include <new>

namespace lfmm::host {

using namespace lfmm;

struct DevTables2d {
    int N = 0;
    void* block = nullptr;
    Tables2d t{};
};

static int get_tables2d(lfmm_ctx* ctx, int N, const DevTables2d** out)
{
    auto it = ctx->tables2d.find(N);
    if (it != ctx->tables2d.end()) {
        *out = it->second;
        return LFMM_OK;
    }
    OpTables2d T;
    build_op_tables_2d(N, T);
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const std::vector<float> m2l32(T.m2l.begin(), T.m2l.end());  // round to nearest
    const size_t b1 = T.m2l.size() * 8, b2 = T.m2m.size() * 8, b3 = T.l2l.size() * 8, b4 = T.dir_vec.size() * 4,
                 b5 = T.class_dirs.size() * 4, b6 = m2l32.size() * 4;
    char* blk = nullptr;
    if (hipMalloc(&blk, al(b1) + al(b2) + al(b3) + al(b4) + al(b5) + al(b6)) != hipSuccess) return LFMM_ERR_NOMEM;
    size_t o = 0;
    bool ok = true;
    auto put = [&](const void* src, size_t bytes) -> void* {
        void* dst = blk + o;
        ok = ok && hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess;
        o += al(bytes);
        return dst;
    };
    DevTables2d* D = new (std::nothrow) DevTables2d();
    if (!D) {
        (void)hipFree(blk);
        return LFMM_ERR_NOMEM;
    }
    D->N = N;
    D->block = blk;
    D->t.m2l = (const double2*)put(T.m2l.data(), b1);
    D->t.m2m = (const double2*)put(T.m2m.data(), b2);
    D->t.l2l = (const double2*)put(T.l2l.data(), b3);
    D->t.dir_vec = (const int*)put(T.dir_vec.data(), b4);
    D->t.class_dirs = (const int*)put(T.class_dirs.data(), b5);
    D->t.m2l32 = (const float2*)put(m2l32.data(), b6);
    if (!ok || hipDeviceSynchronize() != hipSuccess) {
        (void)hipFree(blk);
        delete D;
        return LFMM_ERR_HIP;
    }
    ctx->tables2d[N] = D;
    *out = D;
    return LFMM_OK;
}

void free_tables2d(lfmm_ctx* ctx)
{
    for (auto& kv : ctx->tables2d) {
        (void)hipFree(kv.second->block);
        delete kv.second;
    }
    ctx->tables2d.clear();
}

// src/cpu/multipole.cpp:478-481
static int terms_2d(double eps, int terms)
{
    if (terms == LFMM_AUTO) return std::min(LFMM_MAX_TERMS_2D, (int)(-1.18 * std::log(1.65 * eps)));
    return std::min(LFMM_MAX_TERMS_2D, terms);
}

// With LFMM_FLAG_MIXED the M2L of levels >= 4 forms each direction's sum in
// FP32 (the two coarsest M2L levels, 2 and 3, stay FP64, as in 3D).
constexpr int kMixed2dFirstLevel = 4;

static int fmm2d_solve(lfmm_ctx* ctx, const lfmm_points3d* S, const lfmm_points3d* Tg, double eps, int terms_in,
                       int depth_in, int leaf_max, int memory, unsigned flags, double* pot_out, double* field_out,
                       lfmm_report* rep)
{
    const auto T0 = std::chrono::steady_clock::now();
    const bool timed = (flags & LFMM_FLAG_TIMINGS) != 0;
    const bool mixed = (flags & LFMM_FLAG_MIXED) != 0;
    const int64_t n_out = Tg ? Tg->n : S->n;
    const int N = terms_2d(eps, terms_in);
    if (N < 1) return LFMM_ERR_TERMS;
    hipStream_t st = ctx->stream;
    Arena& A = ctx->arena;
    A.reset();
    DevPoints DP;
    int rc = stage_inputs(ctx, S, Tg, memory, 2, DP);
    if (rc != LFMM_OK) return rc;
    const int64_t n = DP.n;
    Tree T;
    rc = build_tree(ctx, DP, depth_in, leaf_max, Tg != nullptr, T);
    if (rc != LFMM_OK) return rc;
    const int depth = T.depth, leaf_level = depth - 1;
    const double t_tree = timed ? ms_since(T0) : 0.0;
    const DevTables2d* DT = nullptr;
    rc = get_tables2d(ctx, N, &DT);
    if (rc != LFMM_OK) return rc;
    const double scale = T.geom_host[3];  // box edge at level l = scale / 2^l (for the M2L level term)

    std::vector<double2*> outer(depth, nullptr), inner(depth, nullptr);
    const bool far = depth >= 3;
    if (far) {
        for (int l = 2; l < depth; ++l) {
            outer[l] = take<double2>(A, T.cnt[l] * N);
            inner[l] = take<double2>(A, T.cnt[l] * N);
            if (!outer[l] || !inner[l]) return LFMM_ERR_NOMEM;
        }
        P2M2dArgs pa{T.src, T.lbegin[leaf_level], T.lkeys[leaf_level], (int)T.cnt[leaf_level], leaf_level, T.geom, N,
                     outer[leaf_level]};
        LFMM_HIP(launch_p2m2d(pa, st));
        for (int l = leaf_level - 1; l >= 2; --l) {
            M2M2dArgs ma{N, (int)T.cnt[l], T.lbegin[l], T.lkeys[l + 1], outer[l + 1], outer[l]};
            LFMM_HIP(launch_m2m2d(ma, DT->t, st));
        }
    }
    double t_up = 0.0;
    if (timed) {
        LFMM_HIP(hipStreamSynchronize(st));
        t_up = ms_since(T0) - t_tree;
    }
    size_t nev = 0;
    int launches = 0, mixed_launches = 0;
    if (far) {
        for (int l = 2; l < depth; ++l) {
            M2L2dArgs ma{};
            ma.N = N;
            ma.level = l;
            ma.num_parents = (int)T.cnt[l - 1];
            ma.child_begin = T.lbegin[l - 1];
            ma.child_keys = T.lkeys[l];
            ma.lut = T.lut[l];
            ma.keys = T.lkeys[l];
            ma.size = (int)T.cnt[l];
            ma.log_h = std::log(std::ldexp(scale, -l));
            ma.src = outer[l];
            ma.parent_local = (l > 2) ? inner[l - 1] : nullptr;
            ma.dst = inner[l];
            ma.f32 = mixed && l >= kMixed2dFirstLevel;
            mixed_launches += ma.f32 ? 1 : 0;
            if (timed) LFMM_HIP(hipEventRecord(ctx_event(ctx, nev++), st));
            LFMM_HIP(launch_m2l2d(ma, DT->t, st));
            if (timed) LFMM_HIP(hipEventRecord(ctx_event(ctx, nev++), st));
            ++launches;
        }
    }
    double t_down = 0.0;
    if (timed) {
        LFMM_HIP(hipStreamSynchronize(st));
        t_down = ms_since(T0) - t_tree - t_up;
    }
    double* dpot = pot_out;
    double* dfield = field_out;
    if (memory == LFMM_MEM_HOST) {
        dpot = pot_out ? take<double>(A, n_out) : nullptr;
        dfield = field_out ? take<double>(A, 2 * n_out) : nullptr;
        if ((pot_out && !dpot) || (field_out && !dfield)) return LFMM_ERR_NOMEM;
    }
    Leaf2dArgs la{};
    la.src = T.src;
    la.leaf_of = T.leaf_of;
    la.leaf_key = T.lkeys[leaf_level];
    la.leaf_begin = T.lbegin[leaf_level];
    la.nleaf = (int)T.cnt[leaf_level];
    la.level = leaf_level;
    la.geom = T.geom;
    la.lut = (leaf_level >= 2 && leaf_level <= lut_max_level(2)) ? T.lut[leaf_level] : nullptr;
    la.local = far ? inner[leaf_level] : nullptr;
    la.N = N;
    la.perm = T.order;
    la.n = n;
    la.n_out = n_out;
    la.exclude_self = (flags & LFMM_FLAG_EXCLUDE_SELF) ? 1 : 0;
    la.pot = dpot;
    la.field = dfield;
    const size_t ev_leaf = nev;
    if (timed) LFMM_HIP(hipEventRecord(ctx_event(ctx, nev++), st));
    LFMM_HIP(launch_leaf2d(la, st));
    if (timed) LFMM_HIP(hipEventRecord(ctx_event(ctx, nev++), st));
    if (memory == LFMM_MEM_HOST) {
        if (pot_out) LFMM_HIP(hipMemcpyAsync(pot_out, dpot, sizeof(double) * n_out, hipMemcpyDeviceToHost, st));
        if (field_out) LFMM_HIP(hipMemcpyAsync(field_out, dfield, sizeof(double) * 2 * n_out, hipMemcpyDeviceToHost, st));
    }
    if (timed || memory == LFMM_MEM_HOST) LFMM_HIP(hipStreamSynchronize(st));
    LFMM_HIP(hipGetLastError());
    if (rep) {
        rep->terms = N;
        rep->tree_depth = depth;
        rep->num_leaves = T.cnt[leaf_level];
        rep->max_leaf = T.max_leaf;
        rep->m2l_ops = -1;
        rep->p2p_pairs = -1;
        rep->m2l_ops_mixed = -1;
        rep->m2l_launches = launches;
        rep->m2l_mixed_launches = mixed_launches;
        if (timed) {
            double tm = 0.0;
            for (size_t i = 0; i + 1 < ev_leaf; i += 2) {
                float ms = 0.f;
                LFMM_HIP(hipEventElapsedTime(&ms, ctx->events[i], ctx->events[i + 1]));
                tm += ms;
            }
            float ml = 0.f;
            LFMM_HIP(hipEventElapsedTime(&ml, ctx->events[ev_leaf], ctx->events[ev_leaf + 1]));
            rep->t_m2l_kernels_ms = tm;
            rep->t_leaf_kernel_ms = ml;
        }
        rep->t_tree_ms = t_tree;
        rep->t_upward_ms = t_up;
        rep->t_downward_ms = t_down;
        rep->t_total_ms = timed ? ms_since(T0) : 0.0;
        rep->t_leaf_ms = timed ? rep->t_total_ms - t_tree - t_up - t_down : 0.0;
        if (flags & LFMM_FLAG_COUNT_OPS) {
            unsigned long long* cws = take<unsigned long long>(A, 2 + depth);
            if (!cws) return LFMM_ERR_NOMEM;
            LFMM_HIP(count_ops(2, T.lkeys, T.lbegin, T.lut, T.cnt, cws, &rep->m2l_ops, &rep->p2p_pairs, nullptr, st));
        }
    }
    return LFMM_OK;
}

}  // namespace lfmm::host

using namespace lfmm;
using namespace lfmm::host;

extern "C" {

int lfmm_fmm2d(lfmm_ctx* c, const lfmm_points2d* sources, const lfmm_points2d* targets, double eps, int terms,
               int tree_depth, int leaf_max, int memory, unsigned flags, double* potential, double* field,
               lfmm_report* report)
{
    if (!c || !valid_points(sources, true, 2)) return LFMM_ERR_INVALID;
    if (memory != LFMM_MEM_HOST && memory != LFMM_MEM_DEVICE) return LFMM_ERR_INVALID;
    if (terms != LFMM_AUTO && terms < 1) return LFMM_ERR_TERMS;
    if (tree_depth != LFMM_AUTO && (tree_depth < 1 || tree_depth > 2 * kMaxDepth - 1)) return LFMM_ERR_DEPTH;
    if (!(eps > 0.0) && terms == LFMM_AUTO) return LFMM_ERR_INVALID;
    if (targets && !valid_points(targets, false, 2)) return LFMM_ERR_INVALID;
    const lfmm_points2d* tg = (targets && targets->n > 0) ? targets : nullptr;
    if (tg && sources->n + tg->n >= ((int64_t)1 << 31)) return LFMM_ERR_INVALID;
    if (report) std::memset(report, 0, sizeof(*report));
    if ((tg ? tg->n : sources->n) == 0) return LFMM_OK;
    std::lock_guard<std::mutex> g(c->mu);
    if (hipSetDevice(c->device) != hipSuccess) return LFMM_ERR_HIP;
    if (sources->n == 0) {  // no sources: every output is zero
        const size_t nt = (size_t)tg->n;
        if (memory == LFMM_MEM_HOST) {
            if (potential) std::memset(potential, 0, nt * sizeof(double));
            if (field) std::memset(field, 0, 2 * nt * sizeof(double));
        } else {
            if (potential && hipMemsetAsync(potential, 0, nt * sizeof(double), c->stream) != hipSuccess) return LFMM_ERR_HIP;
            if (field && hipMemsetAsync(field, 0, 2 * nt * sizeof(double), c->stream) != hipSuccess) return LFMM_ERR_HIP;
        }
        return LFMM_OK;
    }
    try {
        return fmm2d_solve(c, sources, tg, eps, terms, tree_depth, leaf_max, memory, flags, potential, field, report);
    } catch (const std::bad_alloc&) {
        return LFMM_ERR_NOMEM;
    } catch (...) {
        return LFMM_ERR_INVALID;
    }
}

}  // extern "C"
