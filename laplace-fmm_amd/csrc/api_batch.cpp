// api_batch.cpp -- many independent 3D systems in one call (lfmm_fmm3d_batch).
//
// The reference solves one system per FastMultipole3d construction
// (src/cuda/multipole3d.h:16) on the legacy default stream, so a caller with
// many systems loops and every solve serialises behind the previous one.
// Here the systems are dealt, largest first, to a pool of worker contexts.
// Each worker has its own stream, workspace and operator tables and is driven
// by its own host thread, so the host-side steps of one solve (tree level
// sizes read back, launch set-up) overlap with the kernels of the others and
// small systems fill the GPU together.
//
// With LFMM_FLAG_ADAPTIVE the systems are instead mapped into cells of one
// unit cube, with empty cells between them, and solved as one adaptive
// octree (ForestSpec): one tree build and one pass of every kernel for all
// systems, no host work per system.
#include "api_internal.h"
#include "forest.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <numeric>
#include <thread>
#include <vector>

namespace {

using namespace lfmm;
using namespace lfmm::host;

// Grows the context's forest buffer (outside the arena, which the solve resets).
int forest_reserve(lfmm_ctx* c, size_t bytes)
{
    if (c->forest_bytes >= bytes) return LFMM_OK;
    if (c->forest_buf) {
        LFMM_HIP(hipStreamSynchronize(c->stream));
        LFMM_HIP(hipFree(c->forest_buf));
        c->forest_buf = nullptr;
        c->forest_bytes = 0;
    }
    if (hipMalloc(&c->forest_buf, bytes) != hipSuccess) {
        c->forest_buf = nullptr;
        return LFMM_ERR_NOMEM;
    }
    c->forest_bytes = bytes;
    return LFMM_OK;
}

int forest_batch(lfmm_ctx* c, int n_sys, const int64_t* offsets, const double* pos, const double* q, double eps,
                 int terms, int leaf_max, int memory, unsigned flags, double* potential, double* field,
                 lfmm_report* reports)
{
    if (flags & LFMM_FLAG_LEARNED_ORDER) return LFMM_ERR_UNSUPPORTED;
    const int64_t o0 = offsets[0], n = offsets[n_sys] - o0;
    int side = 1;
    while ((int64_t)side * side * side < n_sys) ++side;
    int level = 1;
    while ((1 << level) < 2 * side) ++level;
    const int64_t ncell = (int64_t)side * side * side;
    // buffer: points, sorted offsets, extents, geometries, scales, geometry, flag,
    // and with host memory the staged inputs and outputs
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    size_t at = 0;
    const size_t p_pts = at;  at += al(sizeof(double) * 4 * n);
    const size_t p_off = at;  at += al(sizeof(int64_t) * (n_sys + 1));
    const size_t p_ext = at;  at += al(sizeof(double) * 6 * n_sys);
    const size_t p_geo = at;  at += al(sizeof(double) * 4 * n_sys);
    const size_t p_scl = at;  at += al(sizeof(double) * ncell);
    const size_t p_fix = at;  at += al(sizeof(double) * 4);
    const size_t p_bad = at;  at += al(sizeof(int));
    const bool host = memory == LFMM_MEM_HOST;
    const size_t p_in = at;   if (host) at += al(sizeof(double) * 4 * n);
    const size_t p_out = at;  if (host) at += al(sizeof(double) * 4 * n);
    int rc = forest_reserve(c, at);
    if (rc != LFMM_OK) return rc;
    char* base = (char*)c->forest_buf;
    hipStream_t st = c->stream;
    double4* pts = (double4*)(base + p_pts);
    int64_t* doff = (int64_t*)(base + p_off);
    double* ext = (double*)(base + p_ext);
    double* geo = (double*)(base + p_geo);
    double* scl = (double*)(base + p_scl);
    double* fix = (double*)(base + p_fix);
    int* bad = (int*)(base + p_bad);
    const double* dpos = pos + 3 * o0;
    const double* dq = q + o0;
    if (host) {
        double* in = (double*)(base + p_in);
        LFMM_HIP(hipMemcpyAsync(in, dpos, sizeof(double) * 3 * n, hipMemcpyHostToDevice, st));
        LFMM_HIP(hipMemcpyAsync(in + 3 * n, dq, sizeof(double) * n, hipMemcpyHostToDevice, st));
        dpos = in;
        dq = in + 3 * n;
    }
    std::vector<int64_t> hoff(n_sys + 1);
    for (int s = 0; s <= n_sys; ++s) hoff[s] = offsets[s] - o0;
    LFMM_HIP(hipMemcpyAsync(doff, hoff.data(), sizeof(int64_t) * (n_sys + 1), hipMemcpyHostToDevice, st));
    int zero = 0;
    LFMM_HIP(hipMemcpyAsync(bad, &zero, sizeof(int), hipMemcpyHostToDevice, st));
    LFMM_HIP(launch_forest_extent(dpos, dq, doff, n_sys, ext, bad, st));
    std::vector<double> hext(6 * (size_t)n_sys);
    int hbad = 0;
    LFMM_HIP(hipMemcpyAsync(hext.data(), ext, sizeof(double) * 6 * n_sys, hipMemcpyDeviceToHost, st));
    LFMM_HIP(hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, st));
    LFMM_HIP(hipStreamSynchronize(st));
    if (hbad) return LFMM_ERR_NONFINITE;
    // every system's own tree geometry (the reference's extent rule,
    // src/cpu/morton_tree.cpp:293-301) and its length scale in the forest
    std::vector<double> hgeo(4 * (size_t)n_sys, 0.0), hscl((size_t)ncell, 1.0);
    const double cell = std::ldexp(1.0, -level);
    for (int s = 0; s < n_sys; ++s) {
        const double* e = &hext[6 * (size_t)s];
        if (hoff[s + 1] == hoff[s]) {
            hgeo[4 * (size_t)s + 3] = 1.0;
            continue;
        }
        const double dw = std::max(std::max(e[3] - e[0], e[4] - e[1]), e[5] - e[2]);
        hgeo[4 * (size_t)s] = e[0] - 1.e-7;
        hgeo[4 * (size_t)s + 1] = e[1] - 1.e-7;
        hgeo[4 * (size_t)s + 2] = e[2] - 1.e-7;
        hgeo[4 * (size_t)s + 3] = dw + 1.e-6;
        hscl[s] = cell / (dw + 1.e-6);
    }
    const double hfix[4] = {0.0, 0.0, 0.0, 1.0};
    LFMM_HIP(hipMemcpyAsync(geo, hgeo.data(), sizeof(double) * 4 * n_sys, hipMemcpyHostToDevice, st));
    LFMM_HIP(hipMemcpyAsync(scl, hscl.data(), sizeof(double) * ncell, hipMemcpyHostToDevice, st));
    LFMM_HIP(hipMemcpyAsync(fix, hfix, sizeof(hfix), hipMemcpyHostToDevice, st));
    LFMM_HIP(launch_forest_map(dpos, dq, doff, n_sys, n, geo, side, level, pts, st));
    // the host copies above read pageable vectors: done before they go out of scope
    LFMM_HIP(hipStreamSynchronize(st));

    ForestSpec fs;
    fs.level = level;
    fs.side = side;
    fs.geom = fix;
    fs.sys_scale = scl;
    const lfmm_points3d fp{n, (const double*)pts, 4, (const double*)pts + 3, 4};
    double* opot = potential ? potential + o0 : nullptr;
    double* ofield = field ? field + 3 * o0 : nullptr;
    if (host) {
        opot = potential ? (double*)(base + p_out) : nullptr;
        ofield = field ? (double*)(base + p_out) + n : nullptr;
    }
    lfmm_report r{};
    rc = fmm3d_solve(c, &fp, nullptr, eps, terms, LFMM_AUTO, leaf_max > 0 ? leaf_max : 64, LFMM_MEM_DEVICE,
                     flags | LFMM_FLAG_ADAPTIVE, opot, ofield, reports ? &r : nullptr, nullptr, nullptr, &fs);
    if (rc != LFMM_OK) return rc;
    if (host) {
        if (potential)
            LFMM_HIP(hipMemcpyAsync(potential + o0, opot, sizeof(double) * n, hipMemcpyDeviceToHost, st));
        if (field)
            LFMM_HIP(hipMemcpyAsync(field + 3 * o0, ofield, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, st));
        LFMM_HIP(hipStreamSynchronize(st));
    }
    if (reports)
        for (int s = 0; s < n_sys; ++s) reports[s] = r;
    return LFMM_OK;
}

}  // namespace

extern "C" {

int lfmm_fmm3d_batch(lfmm_ctx* c, int n_sys, const int64_t* offsets, const double* pos, const double* q,
                     double eps, int terms, int tree_depth, int leaf_max, int memory, unsigned flags, int workers,
                     double* potential, double* field, lfmm_report* reports)
{
    if (!c || n_sys < 0 || (n_sys > 0 && !offsets)) return LFMM_ERR_INVALID;
    if (memory != LFMM_MEM_HOST && memory != LFMM_MEM_DEVICE) return LFMM_ERR_INVALID;
    if (workers < 0 || workers > LFMM_BATCH_MAX_WORKERS) return LFMM_ERR_INVALID;
    for (int s = 0; s < n_sys; ++s)
        if (offsets[s] < 0 || offsets[s + 1] < offsets[s]) return LFMM_ERR_INVALID;
    if (n_sys == 0 || offsets[n_sys] == offsets[0]) return LFMM_OK;
    if (!pos || !q) return LFMM_ERR_INVALID;
    if (reports) std::memset(reports, 0, sizeof(lfmm_report) * (size_t)n_sys);
    if (flags & LFMM_FLAG_ADAPTIVE) {
        if (offsets[n_sys] - offsets[0] >= ((int64_t)1 << 31)) return LFMM_ERR_INVALID;
        std::lock_guard<std::mutex> g(c->mu);
        if (hipSetDevice(c->device) != hipSuccess) return LFMM_ERR_HIP;
        try {
            return forest_batch(c, n_sys, offsets, pos, q, eps, terms, leaf_max, memory, flags, potential, field,
                                reports);
        } catch (const std::bad_alloc&) {
            return LFMM_ERR_NOMEM;
        }
    }
    const int nw = std::min(n_sys, workers > 0 ? workers : LFMM_BATCH_DEFAULT_WORKERS);

    hipEvent_t ready = nullptr;
    {
        std::lock_guard<std::mutex> g(c->mu);
        if (hipSetDevice(c->device) != hipSuccess) return LFMM_ERR_HIP;
        while ((int)c->workers.size() < nw) {
            lfmm_ctx* w = nullptr;
            const int st = lfmm_create(&w, c->device);
            if (st != LFMM_OK) return st;
            c->workers.push_back(w);
        }
        // device inputs may still be in flight on the caller's stream
        if (memory == LFMM_MEM_DEVICE) {
            if (hipEventCreateWithFlags(&ready, hipEventDisableTiming) != hipSuccess) return LFMM_ERR_HIP;
            if (hipEventRecord(ready, c->stream) != hipSuccess) {
                (void)hipEventDestroy(ready);
                return LFMM_ERR_HIP;
            }
        }
    }

    // largest systems first, so that the last ones to start are the short ones
    std::vector<int> order(n_sys);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        return offsets[a + 1] - offsets[a] > offsets[b + 1] - offsets[b];
    });
    std::atomic<int> next{0};
    std::atomic<int> status{LFMM_OK};
    auto run = [&](lfmm_ctx* w) {
        if (hipSetDevice(c->device) != hipSuccess) {
            status.store(LFMM_ERR_HIP);
            return;
        }
        if (ready && hipStreamWaitEvent(w->stream, ready, 0) != hipSuccess) {
            status.store(LFMM_ERR_HIP);
            return;
        }
        while (status.load() == LFMM_OK) {
            const int i = next.fetch_add(1);
            if (i >= n_sys) break;
            const int s = order[i];
            const int64_t o = offsets[s], n = offsets[s + 1] - o;
            if (n == 0) continue;
            const lfmm_points3d src{n, pos + 3 * o, 3, q + o, 1};
            const int st = lfmm_fmm3d(w, &src, nullptr, eps, terms, tree_depth, leaf_max, memory, flags,
                                      potential ? potential + o : nullptr, field ? field + 3 * o : nullptr,
                                      reports ? reports + s : nullptr);
            if (st != LFMM_OK) {
                int ok = LFMM_OK;
                status.compare_exchange_strong(ok, st);
            }
        }
        // also after a failure: nothing of this worker may still run when the call returns
        if (hipStreamSynchronize(w->stream) != hipSuccess) {
            int ok = LFMM_OK;
            status.compare_exchange_strong(ok, LFMM_ERR_HIP);
        }
    };
    std::vector<std::thread> pool;
    try {
        pool.reserve(nw);
        for (int k = 0; k < nw; ++k) pool.emplace_back(run, c->workers[k]);
    } catch (...) {
        // fewer threads than asked: those started take every system; with
        // none started, this thread runs them
        if (pool.empty()) run(c->workers[0]);
    }
    for (auto& t : pool) t.join();
    if (ready) {
        for (int k = 0; k < nw; ++k) (void)hipStreamSynchronize(c->workers[k]->stream);
        (void)hipEventDestroy(ready);
    }
    return status.load();
}

}  // extern "C"
