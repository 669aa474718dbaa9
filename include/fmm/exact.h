// fmm/exact.h -- the reference's checking path: direct sums written to
// "<kind>_exact.txt" and the error report that reads them back.
//
// Mirrors src/common/simple_math.h:218-402 and src/common/simple_math.cu:24-130:
//   fmm::InteractionType                        (src/common/utils.h:19-24)
//   fmm::gpu::ComputeExact<it>(sources, targets), fmm::gpu::ComputeExact<it>(particles)
//   fmm::ComputeExact<it>(...)                  the reference's CPU twin; runs on the GPU here
//   fmm::ReadError<it>(results)                 prints max / l2 / relative error
// The sums run on the GPU in FP64 through lfmm_direct3d / lfmm_direct2d; the
// files use the reference's text format (20 significant digits, complex as
// "(re,im)", vectors as "x y z") in the working directory.  Failures throw
// std::runtime_error (the reference has no error path).
#pragma once

#include <complex>
#include <iostream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "multipole3d.h"

namespace fmm {

enum class InteractionType { Potential2d, Force2d, Potential3d, Force3d };

namespace detail {

inline void exact_check(int rc, const char* what)
{
    if (rc != LFMM_OK) throw std::runtime_error(std::string(what) + ": " + lfmm_strerror(rc));
}

template <typename point_type, typename value_type>
lfmm_points3d exact_view(const std::vector<particle<point_type, value_type>>& p)
{
    static_assert(sizeof(particle<point_type, value_type>) % sizeof(double) == 0, "particle layout");
    lfmm_points3d v{};
    v.n = (int64_t)p.size();
    v.pos = p.empty() ? nullptr : reinterpret_cast<const double*>(&p[0].center);
    v.pos_stride = (int64_t)(sizeof(particle<point_type, value_type>) / sizeof(double));
    v.q = p.empty() ? nullptr : reinterpret_cast<const double*>(&p[0].q);
    v.q_stride = v.pos_stride;
    return v;
}

}  // namespace detail

namespace gpu {

template <InteractionType it, typename point_type, typename value_type>
void ComputeExact(const std::vector<particle<point_type, value_type>>& source_particles,
                  const std::vector<particle<point_type, value_type>>& target_particles)
{
    constexpr bool two_d = it == InteractionType::Potential2d || it == InteractionType::Force2d;
    constexpr bool potential = it == InteractionType::Potential2d || it == InteractionType::Potential3d;
    static_assert(two_d == std::is_same_v<point_type, std::complex<double>>, "2D kinds take complex positions");
    const lfmm_points3d s = detail::exact_view(source_particles);
    const lfmm_points3d t = detail::exact_view(target_particles);
    const bool same = &source_particles == &target_particles;
    const size_t n = target_particles.size();
    if constexpr (std::is_same_v<value_type, Vector3d>) {
        // vector charges: the force q x r / |r|^3 only (simple_math.h:148-152)
        static_assert(it == InteractionType::Force3d, "vector charges have forces only");
        std::vector<double> f(3 * n);
        detail::exact_check(lfmm_direct3d_vec(detail::thread_context(), &s, same ? nullptr : &t, LFMM_MEM_HOST,
                                              f.data()),
                            "lfmm_direct3d_vec");
        const int kind = (int)it;
        detail::exact_check(lfmm_write_exact(lfmm_exact_filename(kind), kind, (int64_t)n, f.data()), "lfmm_write_exact");
        return;
    }
    static_assert(std::is_same_v<value_type, double> || std::is_same_v<value_type, Vector3d>, "charges");
    const int width = potential ? 1 : (two_d ? 2 : 3);
    std::vector<double> values(n * width);
    double* pot = potential ? values.data() : nullptr;
    double* field = potential ? nullptr : values.data();
    const int rc = two_d ? lfmm_direct2d(detail::thread_context(), &s, same ? nullptr : &t, LFMM_MEM_HOST, 0u, pot, field)
                         : lfmm_direct3d(detail::thread_context(), &s, same ? nullptr : &t, LFMM_MEM_HOST, 0u, pot, field);
    detail::exact_check(rc, two_d ? "lfmm_direct2d" : "lfmm_direct3d");
    const int kind = (int)it;
    detail::exact_check(lfmm_write_exact(lfmm_exact_filename(kind), kind, (int64_t)n, values.data()), "lfmm_write_exact");
}

template <InteractionType it, typename point_type, typename value_type>
void ComputeExact(const std::vector<particle<point_type, value_type>>& particles)
{
    fmm::gpu::ComputeExact<it, point_type, value_type>(particles, particles);
}

}  // namespace gpu

template <InteractionType it, typename point_type, typename value_type>
void ComputeExact(const std::vector<particle<point_type, value_type>>& source_particles,
                  const std::vector<particle<point_type, value_type>>& target_particles)
{
    fmm::gpu::ComputeExact<it, point_type, value_type>(source_particles, target_particles);
}

template <InteractionType it, typename point_type, typename value_type>
void ComputeExact(const std::vector<particle<point_type, value_type>>& particles)
{
    fmm::gpu::ComputeExact<it, point_type, value_type>(particles, particles);
}

template <InteractionType it, typename T>
void ReadError(const std::vector<T>& res)
{
    static_assert(sizeof(T) % sizeof(double) == 0, "result layout");
    constexpr int width = (int)(sizeof(T) / sizeof(double));
    constexpr bool two_d = it == InteractionType::Potential2d || it == InteractionType::Force2d;
    constexpr bool potential = it == InteractionType::Potential2d || it == InteractionType::Potential3d;
    static_assert(width == (potential ? 1 : (two_d ? 2 : 3)), "result type of this kind: double, complex or Vector3d");
    const int kind = (int)it;
    const int64_t n = (int64_t)res.size();
    std::vector<double> exact((size_t)n * width, 0.0);
    const int64_t got = lfmm_read_exact(lfmm_exact_filename(kind), kind, n, exact.data());
    if (got < 0) detail::exact_check((int)-got, "lfmm_read_exact");
    double max_error = 0.0, l2 = 0.0, rel = 0.0;
    detail::exact_check(lfmm_error_stats(kind, n, exact.data(), reinterpret_cast<const double*>(res.data()), &max_error,
                                         &l2, &rel),
                        "lfmm_error_stats");
    std::cout << "--------------------------------" << std::endl;
    std::cout << "max error = " << max_error << std::endl;
    std::cout << "l2 error = " << l2 << std::endl;
    std::cout << "relative error = " << rel << std::endl;
    std::cout << "--------------------------------" << std::endl;
}

}  // namespace fmm
