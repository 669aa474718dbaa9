// fmm/multipole.h -- drop-in C++ front end of the MI355X 2D Laplace FMM.
//
// Mirrors the reference's 2D API (src/cuda/multipole.h:7-35, CPU twin
// src/cpu/multipole.h:10-40):
//   fmm::point2d = std::complex<double>, fmm::particle2d   (src/common/utils.h:149-151)
//   fmm::gpu::FastMultipole
//     FastMultipole(particles, eps = 1e-8, N = FMM_AUTO, tree_depth = FMM_AUTO)
//     FastMultipole(sources, targets, eps, N, tree_depth)
//     std::vector<std::complex<double>> forces;  std::vector<double> potentials;
//   FastMultipole(particles, fmm::Options{p, leaf_max}) and the (sources,
//     targets, Options) twin: order, leaf bound, self-term switch and the
//     FP32 M2L option (Options::adaptive does not apply in 2D).
// The computation runs inside the constructor; failures throw
// std::runtime_error with the library status.  The (sources, targets)
// constructor ignores the targets' charges (the reference requires q = 0).
#pragma once

#include <complex>

#include "multipole3d.h"

namespace fmm {

using point2d = std::complex<double>;
using particle2d = particle<point2d, double>;

namespace gpu {

class FastMultipole {
public:
    FastMultipole(const std::vector<particle2d>& particles, double eps = 1.e-8, int N = FMM_AUTO,
                  int tree_depth = FMM_AUTO)
    {
        Solve(particles, nullptr, eps, N, tree_depth);
    }

    FastMultipole(const std::vector<particle2d>& source_particles, const std::vector<particle2d>& target_particles,
                  double eps = 1.e-8, int N = FMM_AUTO, int tree_depth = FMM_AUTO)
    {
        Solve(source_particles, &target_particles, eps, N, tree_depth);
    }

    FastMultipole(const std::vector<particle2d>& particles, const Options& o)
    {
        Solve(particles, nullptr, o.eps, o.terms(), o.depth(), o.leaf_max, o.flags(true) & ~LFMM_FLAG_ADAPTIVE);
    }

    FastMultipole(const std::vector<particle2d>& source_particles, const std::vector<particle2d>& target_particles,
                  const Options& o)
    {
        Solve(source_particles, &target_particles, o.eps, o.terms(), o.depth(), o.leaf_max,
              o.flags(false) & ~LFMM_FLAG_ADAPTIVE);
    }

    std::vector<std::complex<double>> forces;
    std::vector<double> potentials;

private:
    static_assert(sizeof(particle2d) == 4 * sizeof(double), "reference particle2d layout: 32 bytes");
    static_assert(sizeof(std::complex<double>) == 2 * sizeof(double), "complex must be two packed doubles");

    static lfmm_points2d view(const std::vector<particle2d>& p)
    {
        lfmm_points2d v{};
        v.n = (int64_t)p.size();
        v.pos = p.empty() ? nullptr : reinterpret_cast<const double*>(&p[0].center);
        v.pos_stride = 4;
        v.q = p.empty() ? nullptr : &p[0].q;
        v.q_stride = 4;
        return v;
    }

    void Solve(const std::vector<particle2d>& src, const std::vector<particle2d>* tgt, double eps, int N, int depth,
               int leaf_max = 0, unsigned flags = 0u)
    {
        const lfmm_points2d s = view(src);
        lfmm_points2d t{};
        if (tgt) t = view(*tgt);
        const size_t n_out = tgt ? tgt->size() : src.size();
        potentials.assign(n_out, 0.0);
        forces.assign(n_out, std::complex<double>());
        const int rc = lfmm_fmm2d(detail::thread_context(), &s, tgt ? &t : nullptr, eps, N, depth, leaf_max,
                                  LFMM_MEM_HOST, flags,
                                  potentials.data(), forces.empty() ? nullptr : reinterpret_cast<double*>(forces.data()),
                                  nullptr);
        if (rc != LFMM_OK) throw std::runtime_error(std::string("lfmm_fmm2d: ") + lfmm_strerror(rc));
    }
};

}  // namespace gpu
}  // namespace fmm

// the checking path (direct sums, exact-result files, ReadError)
#include "exact.h"
