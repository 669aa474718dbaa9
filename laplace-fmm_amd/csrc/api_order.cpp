// api_order.cpp -- C ABI of per-box expansion orders and of the order
// network (training and inference on the GPU, order.hip).  No reference
// counterpart: the reference picks one expansion size per solve
// (src/cpu/multipole3d.cpp:1328-1331).
#include "api_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <numeric>
#include <random>
#include <vector>

using namespace lfmm;
using namespace lfmm::host;

namespace {

template <typename T>
T* upload(lfmm_ctx* c, const T* h, int64_t count)
{
    T* d = take<T>(c->arena, count);
    if (d && count > 0 && hipMemcpyAsync(d, h, sizeof(T) * (size_t)count, hipMemcpyHostToDevice, c->stream) != hipSuccess)
        return nullptr;
    return d;
}

template <typename T>
int download(T* h, const T* d, int64_t count, hipStream_t st)
{
    if (count > 0) LFMM_HIP(hipMemcpyAsync(h, d, sizeof(T) * (size_t)count, hipMemcpyDeviceToHost, st));
    return LFMM_OK;
}

bool valid_labels(const int32_t* labels, int64_t n)
{
    for (int64_t i = 0; i < n; ++i)
        if (labels[i] < 0 || labels[i] >= kOrderClasses) return false;
    return true;
}

}  // namespace

extern "C" {

int lfmm_fmm3d_orders(lfmm_ctx* c, const lfmm_points3d* sources, const lfmm_points3d* targets,
                      const int32_t* particle_terms, int tree_depth, int leaf_max, int memory, unsigned flags,
                      double* potential, double* field, lfmm_report* report)
{
    const lfmm_points3d* tg = nullptr;
    const int pre = fmm3d_prologue(c, sources, targets, 1e-8, 1, tree_depth, memory, potential, field, report, &tg);
    if (pre != kNeedSolve) return pre;
    if (!particle_terms) return LFMM_ERR_INVALID;
    std::lock_guard<std::mutex> g(c->mu);
    if (hipSetDevice(c->device) != hipSuccess) return LFMM_ERR_HIP;
    try {
        return fmm3d_solve(c, sources, tg, 1e-8, 1, tree_depth, leaf_max, memory,
                           (flags & ~LFMM_FLAG_LEARNED_ORDER) | LFMM_FLAG_ADAPTIVE, potential, field, report, nullptr,
                           particle_terms);
    } catch (const std::bad_alloc&) {
        return LFMM_ERR_NOMEM;
    } catch (...) {
        return LFMM_ERR_INVALID;
    }
}

int lfmm_set_order_model(lfmm_ctx* c, const float* model, double confidence, double check_tol, int check_samples)
{
    if (!c) return LFMM_ERR_INVALID;
    if (model && (!(confidence > 0.0 && confidence <= 1.0) || !(check_tol > 0.0) || check_samples < 1))
        return LFMM_ERR_INVALID;
    std::lock_guard<std::mutex> g(c->mu);
    if (hipSetDevice(c->device) != hipSuccess) return LFMM_ERR_HIP;
    if (!model) {
        if (c->order_model) {
            (void)hipStreamSynchronize(c->stream);
            (void)hipFree(c->order_model);
        }
        c->order_model = nullptr;
        return LFMM_OK;
    }
    for (int i = 0; i < kOrderModelSize; ++i)
        if (!std::isfinite(model[i])) return LFMM_ERR_NONFINITE;
    for (int k = 0; k < kOrderFeatures; ++k)
        if (!(model[kOffSd + k] > 0.0f)) return LFMM_ERR_INVALID;
    if (!c->order_model && hipMalloc(&c->order_model, sizeof(float) * kOrderModelSize) != hipSuccess) {
        c->order_model = nullptr;
        return LFMM_ERR_NOMEM;
    }
    // ordered after the context's queued work (a solve may still read the old model)
    LFMM_HIP(hipMemcpyAsync(c->order_model, model, sizeof(float) * kOrderModelSize, hipMemcpyHostToDevice, c->stream));
    LFMM_HIP(hipStreamSynchronize(c->stream));
    c->order_conf = confidence;
    c->order_tol = check_tol;
    c->order_samples = check_samples;
    return LFMM_OK;
}

int lfmm_order_features(lfmm_ctx* c, const lfmm_points3d* particles, int tree_depth, int leaf_max, int64_t* nleaf,
                        float* features, int32_t* particle_leaf)
{
    if (!c || !valid_points(particles, true, 3) || particles->n == 0 || !nleaf) return LFMM_ERR_INVALID;
    if (tree_depth != LFMM_AUTO && (tree_depth < 1 || tree_depth > 16)) return LFMM_ERR_DEPTH;
    if (leaf_max <= 0) leaf_max = 64;
    std::lock_guard<std::mutex> g(c->mu);
    if (hipSetDevice(c->device) != hipSuccess) return LFMM_ERR_HIP;
    try {
        c->arena.reset();
        hipStream_t st = c->stream;
        DevPoints DP;
        int rc = upload_view(c, particles, 3, DP);
        if (rc != LFMM_OK) return rc;
        Tree T;
        if ((rc = build_tree(c, DP, tree_depth, leaf_max, false, T, true)) != LFMM_OK) return rc;
        if ((rc = build_adaptive(c, leaf_max, T)) != LFMM_OK) return rc;
        *nleaf = T.nleaf;
        if (features) {
            const DevTables* DT = nullptr;
            if ((rc = get_tables(c, kOrderMinTerms + kOrderClasses - 1, &DT)) != LFMM_OK) return rc;
            float* f = take<float>(c->arena, T.nleaf * kOrderFeatures);
            if (!f) return LFMM_ERR_NOMEM;
            if ((rc = order_features(c, T, DT, f)) != LFMM_OK) return rc;
            if ((rc = download(features, f, T.nleaf * kOrderFeatures, st)) != LFMM_OK) return rc;
        }
        if (particle_leaf) {
            std::vector<int32_t> lo(T.n);
            std::vector<uint32_t> ord(T.n);
            if ((rc = download(lo.data(), T.leaf_of, T.n, st)) != LFMM_OK) return rc;
            if ((rc = download(ord.data(), T.order, T.n, st)) != LFMM_OK) return rc;
            LFMM_HIP(hipStreamSynchronize(st));
            for (int64_t s = 0; s < T.n; ++s) particle_leaf[ord[s]] = lo[s];
        }
        LFMM_HIP(hipStreamSynchronize(st));
        return LFMM_OK;
    } catch (const std::bad_alloc&) {
        return LFMM_ERR_NOMEM;
    } catch (...) {
        return LFMM_ERR_INVALID;
    }
}

int lfmm_order_predict(lfmm_ctx* c, const float* model, const float* features, int64_t n, double confidence,
                       float* probs, int32_t* terms)
{
    if (!c || !model || n < 0 || (n > 0 && !features) || !(confidence > 0.0 && confidence <= 1.0))
        return LFMM_ERR_INVALID;
    if (n == 0) return LFMM_OK;
    std::lock_guard<std::mutex> g(c->mu);
    if (hipSetDevice(c->device) != hipSuccess) return LFMM_ERR_HIP;
    try {
        c->arena.reset();
        hipStream_t st = c->stream;
        const float* dm = upload(c, model, kOrderModelSize);
        const float* df = upload(c, features, n * kOrderFeatures);
        float* dp = probs ? take<float>(c->arena, n * kOrderClasses) : nullptr;
        int32_t* dt = terms ? take<int32_t>(c->arena, n) : nullptr;
        if (!dm || !df || (probs && !dp) || (terms && !dt)) return LFMM_ERR_NOMEM;
        int rc = order_predict(st, dm, df, n, (float)confidence, dp, dt);
        if (rc != LFMM_OK) return rc;
        if (probs && (rc = download(probs, dp, n * kOrderClasses, st)) != LFMM_OK) return rc;
        if (terms && (rc = download(terms, dt, n, st)) != LFMM_OK) return rc;
        LFMM_HIP(hipStreamSynchronize(st));
        return LFMM_OK;
    } catch (const std::bad_alloc&) {
        return LFMM_ERR_NOMEM;
    }
}

int lfmm_order_loss_grad(lfmm_ctx* c, const float* model, const float* features, const int32_t* labels, int64_t n,
                         float* loss, float* grad)
{
    if (!c || !model || n <= 0 || !features || !labels || !valid_labels(labels, n)) return LFMM_ERR_INVALID;
    std::lock_guard<std::mutex> g(c->mu);
    if (hipSetDevice(c->device) != hipSuccess) return LFMM_ERR_HIP;
    try {
        c->arena.reset();
        hipStream_t st = c->stream;
        const float* dm = upload(c, model, kOrderModelSize);
        const float* df = upload(c, features, n * kOrderFeatures);
        const int32_t* dl = upload(c, labels, n);
        float* dg = take<float>(c->arena, kOrderModelSize);
        float* dloss = take<float>(c->arena, 1);
        const int64_t parts = order_grad_parts(n);
        float* gp = take<float>(c->arena, parts * kOrderModelSize);
        float* lp = take<float>(c->arena, parts);
        if (!dm || !df || !dl || !dg || !dloss || !gp || !lp) return LFMM_ERR_NOMEM;
        int rc = order_grad(st, dm, df, dl, nullptr, n, 1.0f / (float)n, dg, dloss, gp, lp);
        if (rc != LFMM_OK) return rc;
        float ls = 0.0f;
        if (grad && (rc = download(grad, dg, kOrderModelSize, st)) != LFMM_OK) return rc;
        if ((rc = download(&ls, dloss, 1, st)) != LFMM_OK) return rc;
        LFMM_HIP(hipStreamSynchronize(st));
        if (loss) *loss = ls / (float)n;
        return LFMM_OK;
    } catch (const std::bad_alloc&) {
        return LFMM_ERR_NOMEM;
    }
}

int lfmm_order_train(lfmm_ctx* c, float* model, const float* features, const int32_t* labels, int64_t n, int epochs,
                     int batch, double lr, uint64_t seed, float* loss_history)
{
    if (!c || !model || n <= 0 || !features || !labels || epochs < 1 || batch < 1 || !(lr > 0.0) ||
        !valid_labels(labels, n))
        return LFMM_ERR_INVALID;
    std::lock_guard<std::mutex> g(c->mu);
    if (hipSetDevice(c->device) != hipSuccess) return LFMM_ERR_HIP;
    try {
        c->arena.reset();
        hipStream_t st = c->stream;
        const int64_t bs = std::min<int64_t>(batch, n);
        float* dm = upload(c, model, kOrderModelSize);
        const float* df = upload(c, features, n * kOrderFeatures);
        const int32_t* dl = upload(c, labels, n);
        float* dg = take<float>(c->arena, kOrderModelSize);
        float* mv = take<float>(c->arena, 2 * kOrderModelSize);
        float* dloss = take<float>(c->arena, 1);
        float* hist = take<float>(c->arena, epochs);
        int32_t* didx = take<int32_t>(c->arena, n);
        const int64_t parts = order_grad_parts(bs);
        float* gp = take<float>(c->arena, parts * kOrderModelSize);
        float* lp = take<float>(c->arena, parts);
        if (!dm || !df || !dl || !dg || !mv || !dloss || !hist || !didx || !gp || !lp) return LFMM_ERR_NOMEM;
        LFMM_HIP(fill_u32((uint32_t*)mv, 2 * kOrderModelSize, 0u, st));
        LFMM_HIP(fill_u32((uint32_t*)hist, epochs, 0u, st));
        std::vector<int32_t> idx(n);
        std::mt19937_64 rng(seed);
        int step = 0;
        for (int e = 0; e < epochs; ++e) {
            std::iota(idx.begin(), idx.end(), 0);
            std::shuffle(idx.begin(), idx.end(), rng);
            // the previous epoch's batches have read didx: the copy is ordered after them on the stream
            LFMM_HIP(hipMemcpyAsync(didx, idx.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, st));
            LFMM_HIP(hipStreamSynchronize(st));  // idx is reused by the next epoch
            for (int64_t b = 0; b < n; b += bs) {
                const int64_t nb = std::min<int64_t>(bs, n - b);
                int rc = order_grad(st, dm, df, dl, didx + b, nb, 1.0f / (float)nb, dg, dloss, gp, lp);
                if (rc != LFMM_OK) return rc;
                if ((rc = order_accum(st, dloss, hist + e)) != LFMM_OK) return rc;
                if ((rc = order_adam(st, dm, dg, mv, mv + kOrderModelSize, (float)lr, ++step)) != LFMM_OK) return rc;
            }
        }
        std::vector<float> hh(epochs);
        int rc = download(model, dm, kOrderModelSize, st);
        if (rc == LFMM_OK) rc = download(hh.data(), hist, epochs, st);
        if (rc != LFMM_OK) return rc;
        LFMM_HIP(hipStreamSynchronize(st));
        if (loss_history)
            for (int e = 0; e < epochs; ++e) loss_history[e] = hh[e] / (float)n;
        return LFMM_OK;
    } catch (const std::bad_alloc&) {
        return LFMM_ERR_NOMEM;
    }
}

}  // extern "C"
