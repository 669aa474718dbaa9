// api_vec.cpp -- vector charges: lfmm_fmm3d_vec and lfmm_direct3d_vec.
//
// The reference's FastMultipole3d<Vector3d> (src/cuda/multipole3d.cu:633-839,
// CPU src/cpu/multipole3d.cpp:414-853) evaluates F_i = sum_j q_j x (r_i - r_j)
// / |r_i - r_j|^3 for vector charges q_j: it runs the scalar operators on the
// three charge components and combines the three gradient fields with a cross
// product (forces_cross, src/cuda/multipole3d.cu:765-776).  Here each
// component is one scalar solve through lfmm_fmm3d (the same tree, since the
// tree depends on the positions only) and k_cross_fields combines the fields.
#include <cstring>
#include <mutex>
#include <vector>

#include "api_internal.h"
#include "direct.h"

using namespace lfmm;
using namespace lfmm::host;

namespace {

// scalar view of charge component c of a vector-charge view
lfmm_points3d component(const lfmm_points3d* p, int c)
{
    lfmm_points3d v = *p;
    if (v.q) v.q += c;
    return v;
}

// out (n x 3) <- cross combination of host fields e1, e2, e3
void cross_host(const double* e1, const double* e2, const double* e3, double* out, int64_t n)
{
    for (int64_t i = 0; i < n; ++i) {
        const double f0 = e2[3 * i + 2] - e3[3 * i + 1];
        const double f1 = e3[3 * i] - e1[3 * i + 2];
        const double f2 = e1[3 * i + 1] - e2[3 * i];
        out[3 * i] = f0;
        out[3 * i + 1] = f1;
        out[3 * i + 2] = f2;
    }
}

template <typename Fn>
int vec_solve(lfmm_ctx* c, const lfmm_points3d* sources, const lfmm_points3d* targets, int memory, double* force,
              Fn scalar)
{
    if (!c || !sources || sources->n < 0 || (sources->n > 0 && (!sources->q || sources->q_stride < 3)))
        return LFMM_ERR_INVALID;
    if (memory != LFMM_MEM_HOST && memory != LFMM_MEM_DEVICE) return LFMM_ERR_INVALID;
    const lfmm_points3d* tg = (targets && targets->n > 0) ? targets : nullptr;
    const int64_t n_out = tg ? tg->n : sources->n;
    if (n_out == 0 || !force) return LFMM_OK;
    const size_t bytes = sizeof(double) * 3 * (size_t)n_out;
    if (memory == LFMM_MEM_HOST) {
        std::vector<double> e1(3 * (size_t)n_out), e2(3 * (size_t)n_out);
        const lfmm_points3d s0 = component(sources, 0), s1 = component(sources, 1), s2 = component(sources, 2);
        int rc = scalar(&s0, tg, e1.data());
        if (rc == LFMM_OK) rc = scalar(&s1, tg, e2.data());
        if (rc == LFMM_OK) rc = scalar(&s2, tg, force);
        if (rc != LFMM_OK) return rc;
        cross_host(e1.data(), e2.data(), force, force, n_out);
        return LFMM_OK;
    }
    double* tmp = nullptr;
    {
        // the two component fields live in a per-context buffer that only
        // grows (kept outside the workspace the scalar solves reset)
        std::lock_guard<std::mutex> g(c->mu);
        if (hipSetDevice(c->device) != hipSuccess) return LFMM_ERR_HIP;
        if (c->vec_tmp_bytes < 2 * bytes) {
            // earlier work on the stream may still read the old buffer
            if (c->vec_tmp && hipStreamSynchronize(c->stream) != hipSuccess) return LFMM_ERR_HIP;
            (void)hipFree(c->vec_tmp);
            c->vec_tmp = nullptr;
            c->vec_tmp_bytes = 0;
            if (hipMalloc(&c->vec_tmp, 2 * bytes) != hipSuccess) return LFMM_ERR_NOMEM;
            c->vec_tmp_bytes = 2 * bytes;
        }
        tmp = (double*)c->vec_tmp;
    }
    const lfmm_points3d s0 = component(sources, 0), s1 = component(sources, 1), s2 = component(sources, 2);
    int rc = scalar(&s0, tg, tmp);
    if (rc == LFMM_OK) rc = scalar(&s1, tg, tmp + 3 * n_out);
    if (rc == LFMM_OK) rc = scalar(&s2, tg, force);
    std::lock_guard<std::mutex> g(c->mu);
    if (hipSetDevice(c->device) != hipSuccess) rc = LFMM_ERR_HIP;
    if (rc == LFMM_OK && launch_cross_fields(tmp, tmp + 3 * n_out, force, n_out, c->stream) != hipSuccess)
        rc = LFMM_ERR_HIP;
    // device outputs follow the stream order of the context (no host wait)
    return rc;
}

}  // namespace

extern "C" {

int lfmm_fmm3d_vec(lfmm_ctx* c, const lfmm_points3d* sources, const lfmm_points3d* targets, double eps, int terms,
                   int tree_depth, int leaf_max, int memory, unsigned flags, double* force, lfmm_report* report)
{
    if (report) std::memset(report, 0, sizeof(*report));
    lfmm_report part{};
    int calls = 0;
    auto scalar = [&](const lfmm_points3d* s, const lfmm_points3d* t, double* field) {
        lfmm_report r{};
        const int rc = lfmm_fmm3d(c, s, t, eps, terms, tree_depth, leaf_max, memory, flags, nullptr, field, &r);
        if (rc == LFMM_OK) {
            if (calls++ == 0) part = r;
            else {  // timings add up over the three components
                part.t_tree_ms += r.t_tree_ms;
                part.t_upward_ms += r.t_upward_ms;
                part.t_downward_ms += r.t_downward_ms;
                part.t_leaf_ms += r.t_leaf_ms;
                part.t_total_ms += r.t_total_ms;
                part.t_m2l_kernels_ms += r.t_m2l_kernels_ms;
                part.t_leaf_kernel_ms += r.t_leaf_kernel_ms;
                part.t_m2l_mixed_ms += r.t_m2l_mixed_ms;
                part.t_p2m_ms += r.t_p2m_ms;
                part.t_m2m_ms += r.t_m2m_ms;
                part.t_l2l_ms += r.t_l2l_ms;
                part.t_p2l_ms += r.t_p2l_ms;
                part.t_l2p_ms += r.t_l2p_ms;
                part.t_p2p_ms += r.t_p2p_ms;
            }
        }
        return rc;
    };
    const int rc = vec_solve(c, sources, targets, memory, force, scalar);
    if (report && rc == LFMM_OK) *report = part;
    return rc;
}

int lfmm_direct3d_vec(lfmm_ctx* c, const lfmm_points3d* sources, const lfmm_points3d* targets, int memory,
                      double* force)
{
    auto scalar = [&](const lfmm_points3d* s, const lfmm_points3d* t, double* field) {
        return lfmm_direct3d(c, s, t, memory, 0u, nullptr, field);
    };
    return vec_solve(c, sources, targets, memory, force, scalar);
}

}  // extern "C"
