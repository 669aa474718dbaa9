"""GPU: vector charges (lfmm_fmm3d_vec, lfmm_direct3d_vec; the reference's
FastMultipole3d<Vector3d> and ComputeExact on particle3d3).

Pinned by:
  the pair formula  F_i = sum_j q_j x (r_i - r_j) / max(|r_i - r_j|, 1e-4)^3
                    (src/common/simple_math.h:148-152) written out in numpy on a
                    small set: direct3d_vec within 1e-13 rel-L2
  the combination   three scalar fields combined by the cross product of the
                    reference's forces_cross (cuda/multipole3d.cu:765-776):
                    lfmm_fmm3d_vec equals that combination of three lfmm_fmm3d
                    calls bit for bit (host and device paths)
  accuracy          FMM vs direct sums at the scalar solver's error level
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, rel_l2, uniform_cube

pytestmark = pytest.mark.gpu


def case_name(request):
    print("\n[case] %s" % request.node.name, flush=True)


def brute(src, qv, tgt):
    d = tgt[:, None, :] - src[None, :, :]
    r = np.maximum(np.sqrt((d * d).sum(-1)), 1e-4)
    return (np.cross(qv[None, :, :], d) / r[..., None] ** 3).sum(1)


def vec_charges(n, seed):
    x, _ = uniform_cube(n, seed)
    qv = np.random.default_rng(seed + 1).uniform(-1, 1, (n, 3))
    return x, qv


def test_direct_vector_matches_formula(lfmm, request):
    case_name(request)
    x, qv = vec_charges(1500, 5)
    x[10:20] = x[:10]  # coincident pairs: the clamp
    f = lfmm.direct3d_vec(x, qv)
    assert rel_l2(f, brute(x, qv, x)) <= 1e-13
    t, _ = uniform_cube(400, 6)
    t[:5] = x[:5]
    f = lfmm.direct3d_vec(x, qv, targets=t * 3 - 1)
    assert rel_l2(f, brute(x, qv, t * 3 - 1)) <= 1e-13


@pytest.mark.parametrize("mixed", [False, True])
def test_fmm_vector_is_the_cross_of_three_scalar_solves(lfmm, request, mixed):
    case_name(request)
    import torch
    x, qv = vec_charges(30000, 11)
    f, rep = lfmm.fmm3d_vec(x, qv, terms=11, tree_depth=6, mixed=mixed)
    e = [lfmm.fmm3d(x, qv[:, c].copy(), terms=11, tree_depth=6, mixed=mixed)[1] for c in range(3)]
    want = np.stack([e[1][:, 2] - e[2][:, 1], e[2][:, 0] - e[0][:, 2], e[0][:, 1] - e[1][:, 0]], axis=1)
    assert np.array_equal(f, want)
    assert rep["terms"] == 11 and rep["tree_depth"] == 6
    fd, _ = lfmm.fmm3d_vec(torch.from_numpy(x).cuda(), torch.from_numpy(qv).cuda(), terms=11, tree_depth=6,
                           mixed=mixed)
    torch.cuda.synchronize()
    assert np.array_equal(fd.cpu().numpy(), f)


def test_fmm_vector_accuracy_and_targets(lfmm, request):
    case_name(request)
    x, qv = vec_charges(40000, 21)
    d = lfmm.direct3d_vec(x, qv)
    f9, _ = lfmm.fmm3d_vec(x, qv, terms=9, tree_depth=6)
    f14, _ = lfmm.fmm3d_vec(x, qv, terms=14, tree_depth=6)
    e9, e14 = rel_l2(f9, d), rel_l2(f14, d)
    assert e14 < 1e-5 and e14 < e9 / 10
    t = np.random.default_rng(22).uniform(-0.5, 1.5, (3000, 3))
    ft, _ = lfmm.fmm3d_vec(x, qv, terms=12, tree_depth=6, targets=t)
    assert rel_l2(ft, lfmm.direct3d_vec(x, qv, targets=t)) < 1e-4


def test_cpp_vector_dropin_and_checking_flow(lfmm, tmp_path, request):
    """src/examples/main.cpp:251-267 with this library: particle3d3 in place,
    FastMultipole3d deduced from the particles, ComputeExact / ReadError."""
    case_name(request)
    src = tmp_path / "vec.cpp"
    src.write_text(r'''
#include "fmm/multipole3d.h"
#include <cstdio>
#include <random>
int main() {
    std::mt19937_64 g(9);
    std::uniform_real_distribution<double> u(0.0, 1.0), c(-1.0, 1.0);
    std::vector<fmm::particle3d3> p(6000);
    for (auto& x : p) x.center = fmm::point3d(u(g), u(g), u(g));
    for (auto& x : p) x.q = fmm::Vector3d(c(g), c(g), c(g));
    fmm::gpu::ComputeExact<fmm::InteractionType::Force3d>(p);
    fmm::gpu::FastMultipole3d fmm(p, 1.e-8, 10, 4);
    fmm::ReadError<fmm::InteractionType::Force3d>(fmm.forces);
    for (size_t i = 0; i < p.size(); ++i)
        std::fprintf(stderr, "%.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n", p[i].center[0],
                     p[i].center[1], p[i].center[2], p[i].q[0], p[i].q[1], p[i].q[2], fmm.forces[i][0],
                     fmm.forces[i][1], fmm.forces[i][2]);
    return fmm.potentials.empty() ? 0 : 3;
}
''')
    exe = tmp_path / "vec"
    r = subprocess.run(["g++", "-O1", "-std=c++17", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe),
                        lfmm.LIB_PATH, "-Wl,-rpath," + os.path.dirname(lfmm.LIB_PATH)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120, cwd=str(tmp_path))
    assert out.returncode == 0, out.stderr[-2000:]
    a = np.loadtxt(out.stderr.splitlines())
    x, qv, f = a[:, :3], a[:, 3:6], a[:, 6:9]
    want, _ = lfmm.fmm3d_vec(x, qv, terms=10, tree_depth=4)
    assert np.array_equal(f, want)
    n, exact = lfmm.read_exact(tmp_path / "force3d_exact.txt", lfmm.FORCE3D, len(x))
    assert n == len(x) and np.array_equal(exact, lfmm.direct3d_vec(x, qv))
    nums = [float(l.split("=")[1]) for l in out.stdout.splitlines() if "=" in l]
    assert np.allclose(nums, lfmm.error_stats(lfmm.FORCE3D, exact, f), rtol=1e-5)
