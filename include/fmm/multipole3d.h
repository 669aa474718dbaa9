// fmm/multipole3d.h -- drop-in C++ front end of the MI355X Laplace FMM.
//
// Mirrors the reference's public 3D API so existing callers recompile
// against this header and link liblfmm.so instead of the CUDA objects:
//   fmm::FMM_AUTO, fmm::Vector3d, fmm::point3d, fmm::particle<point, value>,
//   fmm::particle3d                           (src/common/defs.h:22, utils.h:26-112)
//   fmm::gpu::FastMultipole3d<double>         (src/cuda/multipole3d.h:7-42)
//   fmm::gpu::FastMultipole3d<Vector3d>, fmm::particle3d3   vector charges (forces only)
//     FastMultipole3d(particles, eps = 1e-8, N = FMM_AUTO, tree_depth = FMM_AUTO)
//     FastMultipole3d(sources, targets, eps, N, tree_depth)
//     std::vector<Vector3d> forces;  std::vector<double> potentials;
//   fmm::detail::InitMathConstants / cudaCopyMathConstants
//     (src/common/special_functions.h:155-157): accepted and ignored; the
//     library builds its own tables on first use.
// The computation runs inside the constructor, as in the reference.  Where
// the reference has no error path (unchecked CUDA calls), this header throws
// std::runtime_error carrying the library status.  The (sources, targets)
// constructor ignores the targets' charges (the reference requires q = 0).
#pragma once

#include <array>
#include <cmath>
#include <cstddef>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#include "../lfmm.h"

namespace fmm {

inline constexpr int FMM_AUTO = std::numeric_limits<int>::max();

template <typename T>
class Vector3 : public std::array<T, 3> {
public:
    constexpr Vector3() : std::array<T, 3>{} {}
    constexpr Vector3(T x, T y, T z) : std::array<T, 3>{x, y, z} {}
    constexpr Vector3 operator+(const Vector3& o) const { return {(*this)[0] + o[0], (*this)[1] + o[1], (*this)[2] + o[2]}; }
    constexpr Vector3 operator-(const Vector3& o) const { return {(*this)[0] - o[0], (*this)[1] - o[1], (*this)[2] - o[2]}; }
    constexpr Vector3 operator*(double s) const { return {(*this)[0] * s, (*this)[1] * s, (*this)[2] * s}; }
    constexpr Vector3 operator/(double s) const { return {(*this)[0] / s, (*this)[1] / s, (*this)[2] / s}; }
    constexpr void operator+=(const Vector3& o) { (*this)[0] += o[0]; (*this)[1] += o[1]; (*this)[2] += o[2]; }
    constexpr void operator-=(const Vector3& o) { (*this)[0] -= o[0]; (*this)[1] -= o[1]; (*this)[2] -= o[2]; }
    constexpr void operator*=(double s) { (*this)[0] *= s; (*this)[1] *= s; (*this)[2] *= s; }
    constexpr void operator/=(double s) { (*this)[0] /= s; (*this)[1] /= s; (*this)[2] /= s; }
};

template <typename T>
constexpr Vector3<T> operator*(double s, const Vector3<T>& v) { return v * s; }
template <typename T>
inline double norm(const Vector3<T>& v) { return v[0] * v[0] + v[1] * v[1] + v[2] * v[2]; }
template <typename T>
inline double abs(const Vector3<T>& v) { return std::sqrt(norm(v)); }

using Vector3d = Vector3<double>;
using point3d = Vector3d;

template <typename point, typename value>
struct particle {
    point center;
    value q;
    size_t morton_code;  // scratch in the reference; ignored here
};

using particle3d = particle<point3d, double>;

namespace detail {
inline void InitMathConstants(int) {}
inline void cudaCopyMathConstants() {}
inline void cudaClearMathConstants() {}

// One library context per host thread (device 0).
inline lfmm_ctx* thread_context()
{
    struct Holder {
        lfmm_ctx* ctx = nullptr;
        ~Holder() { lfmm_destroy(ctx); }
    };
    thread_local Holder h;
    if (!h.ctx) {
        const int rc = lfmm_create(&h.ctx, 0);
        if (rc != LFMM_OK) throw std::runtime_error(std::string("lfmm_create: ") + lfmm_strerror(rc));
    }
    return h.ctx;
}
}  // namespace detail

namespace gpu {

template <typename value>
class FastMultipole3d;

template <>
class FastMultipole3d<double> {
public:
    using particle_t = fmm::particle<point3d, double>;

    FastMultipole3d(const std::vector<particle_t>& particles, double eps = 1.e-8, int N = FMM_AUTO,
                    int tree_depth = FMM_AUTO)
    {
        Solve(particles, nullptr, eps, N, tree_depth);
    }

    FastMultipole3d(const std::vector<particle_t>& source_particles, const std::vector<particle_t>& target_particles,
                    double eps = 1.e-8, int N = FMM_AUTO, int tree_depth = FMM_AUTO)
    {
        Solve(source_particles, &target_particles, eps, N, tree_depth);
    }

    std::vector<Vector3d> forces;
    std::vector<double> potentials;

private:
    static_assert(sizeof(particle_t) == 5 * sizeof(double), "reference particle layout: 40 bytes");
    static_assert(sizeof(Vector3d) == 3 * sizeof(double), "Vector3d must be three packed doubles");

    static lfmm_points3d view(const std::vector<particle_t>& p)
    {
        lfmm_points3d v{};
        v.n = (int64_t)p.size();
        v.pos = p.empty() ? nullptr : &p[0].center[0];
        v.pos_stride = 5;
        v.q = p.empty() ? nullptr : &p[0].q;
        v.q_stride = 5;
        return v;
    }

    void Solve(const std::vector<particle_t>& src, const std::vector<particle_t>* tgt, double eps, int N, int depth)
    {
        const lfmm_points3d s = view(src);
        lfmm_points3d t{};
        if (tgt) t = view(*tgt);
        const size_t n_out = tgt ? tgt->size() : src.size();
        potentials.assign(n_out, 0.0);
        forces.assign(n_out, Vector3d());
        const int rc = lfmm_fmm3d(detail::thread_context(), &s, tgt ? &t : nullptr, eps, N, depth, 0, LFMM_MEM_HOST, 0u,
                                  potentials.data(), forces.empty() ? nullptr : &forces[0][0], nullptr);
        if (rc != LFMM_OK) throw std::runtime_error(std::string("lfmm_fmm3d: ") + lfmm_strerror(rc));
    }
};

// Vector charges (the reference's FastMultipole3d<Vector3d>): forces
// F_i = sum_j q_j x (r_i - r_j) / |r_i - r_j|^3; `potentials` stays empty, as
// in the reference.  Particles are the reference's 56-byte particle3d3.
template <>
class FastMultipole3d<Vector3d> {
public:
    using particle_t = fmm::particle<point3d, Vector3d>;

    FastMultipole3d(const std::vector<particle_t>& particles, double eps = 1.e-8, int N = FMM_AUTO,
                    int tree_depth = FMM_AUTO)
    {
        Solve(particles, nullptr, eps, N, tree_depth);
    }

    FastMultipole3d(const std::vector<particle_t>& source_particles, const std::vector<particle_t>& target_particles,
                    double eps = 1.e-8, int N = FMM_AUTO, int tree_depth = FMM_AUTO)
    {
        Solve(source_particles, &target_particles, eps, N, tree_depth);
    }

    std::vector<Vector3d> forces;
    std::vector<double> potentials;

private:
    static_assert(sizeof(particle_t) == 7 * sizeof(double), "reference particle3d3 layout: 56 bytes");

    static lfmm_points3d view(const std::vector<particle_t>& p)
    {
        lfmm_points3d v{};
        v.n = (int64_t)p.size();
        v.pos = p.empty() ? nullptr : &p[0].center[0];
        v.pos_stride = 7;
        v.q = p.empty() ? nullptr : &p[0].q[0];
        v.q_stride = 7;
        return v;
    }

    void Solve(const std::vector<particle_t>& src, const std::vector<particle_t>* tgt, double eps, int N, int depth)
    {
        const lfmm_points3d s = view(src);
        lfmm_points3d t{};
        if (tgt) t = view(*tgt);
        const size_t n_out = tgt ? tgt->size() : src.size();
        forces.assign(n_out, Vector3d());
        const int rc = lfmm_fmm3d_vec(detail::thread_context(), &s, tgt ? &t : nullptr, eps, N, depth, 0,
                                      LFMM_MEM_HOST, 0u, forces.empty() ? nullptr : &forces[0][0], nullptr);
        if (rc != LFMM_OK) throw std::runtime_error(std::string("lfmm_fmm3d_vec: ") + lfmm_strerror(rc));
    }
};

// deduction from the particle type, as the reference's drivers construct
// (src/examples/main.cpp:190, 241)
template <typename value>
FastMultipole3d(const std::vector<fmm::particle<point3d, value>>&, double = 1.e-8, int = FMM_AUTO, int = FMM_AUTO)
    -> FastMultipole3d<value>;
template <typename value>
FastMultipole3d(const std::vector<fmm::particle<point3d, value>>&, const std::vector<fmm::particle<point3d, value>>&,
                double = 1.e-8, int = FMM_AUTO, int = FMM_AUTO) -> FastMultipole3d<value>;

}  // namespace gpu

using particle3d3 = particle<point3d, Vector3d>;

}  // namespace fmm

// the checking path (direct sums, exact-result files, ReadError)
#include "exact.h"
