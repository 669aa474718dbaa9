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
// Beyond the reference: constructors taking fmm::Options, i.e. the order p
// and the leaf bound leaf_max directly ({p, leaf_max}), a self-term switch
// (reference_conventions) and the adaptive-octree and matrix-core options.
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

// Parameters of the (particles, Options) constructors.  The reference's
// constructors take (eps, N, tree_depth) and pick the depth from the particle
// count (src/cpu/multipole3d.cpp:1328-1337); these name the order and the
// leaf bound instead: FastMultipole3d<double> f(particles, {p, leaf_max}).
struct Options {
    int p = FMM_AUTO;            // expansion order, N = p + 1 terms; FMM_AUTO: from eps, as the reference
    int leaf_max = 64;           // the uniform depth (or adaptive split) keeps every leaf at <= leaf_max
                                 // particles; <= 0: tree_depth or the reference's depth formula
    bool reference_conventions = true;  // true: every potential includes the i == j pair at the clamped
                                        // distance 1e-4, as the reference (src/common/simple_math.h:130-133);
                                        // false: the self pair is skipped (single-set solves)
    int tree_depth = FMM_AUTO;   // used when leaf_max <= 0
    double eps = 1.e-8;          // used when p == FMM_AUTO
    bool adaptive = false;       // 3D: adaptive octree (leaves on any level, U/V/W/X lists)
    bool mixed = false;          // M2L of levels >= 4 on the matrix cores (split FP16 with FP32 sums)

    Options() = default;
    Options(int p_, int leaf_max_, bool reference_conventions_ = true)
        : p(p_), leaf_max(leaf_max_), reference_conventions(reference_conventions_) {}

    int terms() const { return p == FMM_AUTO ? FMM_AUTO : p + 1; }
    int depth() const { return leaf_max > 0 ? FMM_AUTO : tree_depth; }
    unsigned flags(bool single_set) const
    {
        unsigned f = 0;
        if (!reference_conventions && single_set) f |= LFMM_FLAG_EXCLUDE_SELF;
        if (adaptive) f |= LFMM_FLAG_ADAPTIVE;
        if (mixed) f |= LFMM_FLAG_MIXED;
        return f;
    }
};

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

    FastMultipole3d(const std::vector<particle_t>& particles, const Options& o)
    {
        Solve(particles, nullptr, o.eps, o.terms(), o.depth(), o.leaf_max, o.flags(true));
    }

    FastMultipole3d(const std::vector<particle_t>& source_particles, const std::vector<particle_t>& target_particles,
                    const Options& o)
    {
        Solve(source_particles, &target_particles, o.eps, o.terms(), o.depth(), o.leaf_max, o.flags(false));
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

    void Solve(const std::vector<particle_t>& src, const std::vector<particle_t>* tgt, double eps, int N, int depth,
               int leaf_max = 0, unsigned flags = 0u)
    {
        const lfmm_points3d s = view(src);
        lfmm_points3d t{};
        if (tgt) t = view(*tgt);
        const size_t n_out = tgt ? tgt->size() : src.size();
        potentials.assign(n_out, 0.0);
        forces.assign(n_out, Vector3d());
        const int rc = lfmm_fmm3d(detail::thread_context(), &s, tgt ? &t : nullptr, eps, N, depth, leaf_max,
                                  LFMM_MEM_HOST, flags, potentials.data(), forces.empty() ? nullptr : &forces[0][0],
                                  nullptr);
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

    FastMultipole3d(const std::vector<particle_t>& particles, const Options& o)
    {
        Solve(particles, nullptr, o.eps, o.terms(), o.depth(), o.leaf_max, o.flags(true));
    }

    FastMultipole3d(const std::vector<particle_t>& source_particles, const std::vector<particle_t>& target_particles,
                    const Options& o)
    {
        Solve(source_particles, &target_particles, o.eps, o.terms(), o.depth(), o.leaf_max, o.flags(false));
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

    void Solve(const std::vector<particle_t>& src, const std::vector<particle_t>* tgt, double eps, int N, int depth,
               int leaf_max = 0, unsigned flags = 0u)
    {
        const lfmm_points3d s = view(src);
        lfmm_points3d t{};
        if (tgt) t = view(*tgt);
        const size_t n_out = tgt ? tgt->size() : src.size();
        forces.assign(n_out, Vector3d());
        const int rc = lfmm_fmm3d_vec(detail::thread_context(), &s, tgt ? &t : nullptr, eps, N, depth, leaf_max,
                                      LFMM_MEM_HOST, flags, forces.empty() ? nullptr : &forces[0][0], nullptr);
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
template <typename value>
FastMultipole3d(const std::vector<fmm::particle<point3d, value>>&, const Options&) -> FastMultipole3d<value>;
template <typename value>
FastMultipole3d(const std::vector<fmm::particle<point3d, value>>&, const std::vector<fmm::particle<point3d, value>>&,
                const Options&) -> FastMultipole3d<value>;

}  // namespace gpu

using particle3d3 = particle<point3d, Vector3d>;

}  // namespace fmm

// the checking path (direct sums, exact-result files, ReadError)
#include "exact.h"
