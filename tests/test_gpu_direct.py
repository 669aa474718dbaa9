"""GPU: FP64 direct sums (lfmm_direct3d / lfmm_direct2d, the reference's
fmm::gpu::ComputeExact) against the CPU oracle's direct sums, and the
reference's checking flow through the C++ front end.

Bars: rel-L2 <= 1e-13 and max relative <= 1e-11 per target against the
oracle (both FP64 in the same summation order; the GPU's 1/r is v_rsq_f64
refined to full precision, the oracle's 1/sqrt, so terms differ by ~1 ulp);
host and device paths bit-identical; the self pair q * 1e4 exactly
removable.
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, rel_l2, uniform_cube

pytestmark = pytest.mark.gpu


def case_name(request):
    print("\n[case] %s" % request.node.name, flush=True)


def close(a, b):
    a, b = np.asarray(a), np.asarray(b)
    scale = np.abs(b).max()
    if scale == 0.0:  # e.g. the field of a lone particle
        return np.abs(a).max() == 0.0
    return rel_l2(a, b) <= 1e-13 and np.abs(a - b).max() <= 1e-11 * scale


@pytest.mark.parametrize("n,self_", [(3000, True), (3000, False), (1, True), (257, False)])
def test_direct3d_matches_oracle(lfmm, orc, request, n, self_):
    case_name(request)
    x, q = uniform_cube(n, 50 + n)
    pot, field = lfmm.direct3d(x, q, exclude_self=not self_)
    op, of = orc.direct3d(x, q, tgt_index=np.arange(n), include_self=self_)
    assert close(pot, op) and close(field, of)


def test_direct3d_targets_duplicates_and_device_path(lfmm, orc, request):
    case_name(request)
    import torch
    x, q = uniform_cube(4000, 61)
    x[100:200] = x[:100]  # coincident points: the 1e-4 clamp
    t, _ = uniform_cube(1500, 62)
    t[:10] = x[:10]       # targets on sources
    pot, field = lfmm.direct3d(x, q, targets=t)
    op, of = orc.direct3d(x, q, tgt_xyz=t)
    assert close(pot, op) and close(field, of)
    pd, fd = lfmm.direct3d(torch.from_numpy(x).cuda(), torch.from_numpy(q).cuda(), targets=torch.from_numpy(t).cuda())
    torch.cuda.synchronize()
    assert np.array_equal(pd.cpu().numpy(), pot) and np.array_equal(fd.cpu().numpy(), field)
    # self pair: exactly q_i / 1e-4 on the diagonal of a single-set call
    p1, f1 = lfmm.direct3d(x[:500], q[:500])
    p0, f0 = lfmm.direct3d(x[:500], q[:500], exclude_self=True)
    assert np.allclose(p1 - p0, q[:500] * 1e4, rtol=1e-12, atol=1e-8)


@pytest.mark.parametrize("self_", [True, False])
def test_direct2d_matches_oracle(lfmm, orc, request, self_):
    case_name(request)
    rng = np.random.default_rng(71)
    x = rng.random((2500, 2))
    q = rng.uniform(-1, 1, 2500)
    pot, field = lfmm.direct2d(x, q, exclude_self=not self_)
    op, of = orc.direct2d(x, q, tgt_index=np.arange(2500), include_self=self_)
    assert close(pot, op) and close(field, of)
    t = rng.random((700, 2))
    pot, field = lfmm.direct2d(x, q, targets=t)
    op, of = orc.direct2d(x, q, tgt_xy=t)
    assert close(pot, op) and close(field, of)


def test_fmm_error_against_gpu_direct_matches_oracle(lfmm, orc, request):
    case_name(request)
    x, q = uniform_cube(20000, 81)
    pot, field, _ = lfmm.fmm3d(x, q, terms=9, tree_depth=4)
    dp, df = lfmm.direct3d(x, q)
    idx = np.arange(0, 20000, 10)
    op, of = orc.direct3d(x, q, tgt_index=idx)
    e_gpu = lfmm.error_stats(lfmm.POTENTIAL3D, dp[idx], pot[idx])
    e_orc = orc.read_error(op, pot[idx])
    assert np.allclose(e_gpu, e_orc, rtol=1e-9)
    assert np.allclose(lfmm.error_stats(lfmm.FORCE3D, df[idx], field[idx]), orc.read_error(of, field[idx]), rtol=1e-9)


def test_cpp_reference_checking_flow(lfmm, tmp_path, request):
    """src/examples/main.cpp:204-223 as a user would run it on this library:
    ComputeExact writes the exact files, the FMM runs, ReadError reports."""
    case_name(request)
    src = tmp_path / "check.cpp"
    src.write_text(r'''
#include "fmm/multipole3d.h"
#include <cstdio>
#include <random>
int main() {
    std::mt19937_64 g(3);
    std::uniform_real_distribution<double> u(0.0, 1.0), c(-1.0, 1.0);
    std::vector<fmm::particle3d> p(5000);
    for (auto& x : p) x.center = fmm::point3d(u(g), u(g), u(g));
    for (auto& x : p) x.q = c(g);
    fmm::gpu::ComputeExact<fmm::InteractionType::Force3d>(p);
    fmm::gpu::ComputeExact<fmm::InteractionType::Potential3d>(p);
    fmm::gpu::FastMultipole3d<double> f(p, 1e-8, 9, 4);
    fmm::ReadError<fmm::InteractionType::Force3d>(f.forces);
    fmm::ReadError<fmm::InteractionType::Potential3d>(f.potentials);
    for (size_t i = 0; i < p.size(); ++i)
        std::fprintf(stderr, "%.17g %.17g %.17g %.17g\n", p[i].center[0], p[i].center[1], p[i].center[2], p[i].q);
    return 0;
}
''')
    exe = tmp_path / "check"
    r = subprocess.run(["g++", "-O1", "-std=c++17", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe),
                        lfmm.LIB_PATH, "-Wl,-rpath," + os.path.dirname(lfmm.LIB_PATH)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120, cwd=str(tmp_path))
    assert out.returncode == 0, out.stderr
    lines = out.stdout.splitlines()
    assert lines[0] == "-" * 32 and lines[4] == "-" * 32 and len(lines) == 10
    nums = [float(l.split("=")[1]) for l in lines if "=" in l]
    a = np.loadtxt(out.stderr.splitlines())
    x, q = a[:, :3], a[:, 3]
    n, exact_f = lfmm.read_exact(tmp_path / "force3d_exact.txt", lfmm.FORCE3D, len(q))
    m, exact_p = lfmm.read_exact(tmp_path / "potential3d_exact.txt", lfmm.POTENTIAL3D, len(q))
    assert n == m == len(q)
    dp, df = lfmm.direct3d(x, q)
    assert np.array_equal(exact_p, dp) and np.array_equal(exact_f, df)
    pot, field, _ = lfmm.fmm3d(x, q, terms=9, tree_depth=4)
    want = lfmm.error_stats(lfmm.FORCE3D, df, field) + lfmm.error_stats(lfmm.POTENTIAL3D, dp, pot)
    assert np.allclose(nums, want, rtol=1e-5)  # printed with 6 significant digits
