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
#include "api_internal.h"

#include <algorithm>
#include <atomic>
#include <numeric>
#include <thread>
#include <vector>

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
