"""GPU parity of the 2D (complex-series) FMM, through the C ABI, against the
CPU oracle's restatement of src/cpu/multipole.cpp.

Bars:
  quadtree             bit-exact (geometry, stable permutation, keys, ranges)
  potential / force    rel-L2 vs the oracle <= 1e-11 (the GPU keeps the
                       expansions normalised per level; the oracle follows the
                       reference's unnormalised series, so results agree to
                       round-off); accuracy vs the FP64 direct sum equal to the
                       oracle's own within 10 %.
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, rel_l2

pytestmark = pytest.mark.gpu


def case_name(request):
    print("\n[case] %s" % request.node.name, flush=True)


def square(n, seed):
    rng = np.random.default_rng(seed)
    return rng.random((n, 2)), rng.uniform(-1, 1, n)


def blobs2(n, seed):
    rng = np.random.default_rng(seed)
    c = rng.uniform(0.2, 0.8, (4, 2))
    x = c[rng.integers(0, 4, n)] + 0.05 * rng.standard_normal((n, 2))
    return x, rng.uniform(-1, 1, n)


@pytest.mark.parametrize("kind,n,depth,seed", [("uniform", 20000, 6, 1), ("uniform", 3000, 1, 2), ("uniform", 3000, 2, 3),
                                               ("blobs", 12000, 9, 4), ("uniform", 1, 4, 5), ("dups", 5000, 5, 6)])
def test_quadtree_bit_exact(lfmm, orc, request, kind, n, depth, seed):
    case_name(request)
    x, q = blobs2(n, seed) if kind == "blobs" else square(n, seed)
    if kind == "dups":
        x[n // 2:] = x[: n - n // 2]
    g = lfmm.tree2d(x, q, depth)
    t = orc.Tree2d(x, q, depth)
    assert g["depth"] == depth
    assert np.array_equal(g["level_sizes"], t.level_sizes)
    assert np.array_equal(g["shift"], t.shift) and g["scale"] == t.scale
    assert np.array_equal(g["pos_map"], t.pos_map)
    for l in range(depth):
        assert np.array_equal(g["keys"][l], t.levels[l]["keys"])
        src = t.levels[l]["src"]
        assert np.array_equal(g["begin"][l][:-1], src[:, 0]) and g["begin"][l][-1] == src[-1, 1]


@pytest.mark.parametrize("N,depth", [(1, 4), (5, 5), (12, 5), (21, 6), (30, 4), (21, 2)])
def test_fmm2d_parity_with_oracle(lfmm, orc, request, N, depth):
    case_name(request)
    x, q = square(8000, 10 + N)
    pot, field, rep = lfmm.fmm2d(x, q, terms=N, tree_depth=depth)
    assert rep["terms"] == N and rep["tree_depth"] == depth
    op, of = orc.fmm2d(x, q, N, depth)
    assert rel_l2(pot, op) <= 1e-11
    assert rel_l2(field, of) <= 1e-11


@pytest.mark.parametrize("dist", ["uniform", "blobs"])
def test_fmm2d_accuracy_vs_direct_matches_oracle(lfmm, orc, request, dist):
    case_name(request)
    n = 30000
    x, q = square(n, 21) if dist == "uniform" else blobs2(n, 22)
    idx = np.random.default_rng(23).choice(n, 1000, replace=False)
    dp, df = orc.direct2d(x, q, tgt_index=idx, include_self=True)
    for N in (8, 21):
        pot, field, _ = lfmm.fmm2d(x, q, terms=N, tree_depth=7)
        op, of = orc.fmm2d(x, q, N, 7)
        assert rel_l2(pot[idx], dp) <= 1.1 * rel_l2(op[idx], dp) + 1e-12
        assert rel_l2(field[idx], df) <= 1.1 * rel_l2(of[idx], df) + 1e-12
        if N == 21:
            assert rel_l2(pot[idx], dp) < 1e-6


def test_fmm2d_source_target_exclude_self_and_layouts(lfmm, orc, request):
    case_name(request)
    import ctypes
    import torch
    src, q = square(9000, 31)
    tgt, _ = blobs2(2500, 32)
    pot, field, _ = lfmm.fmm2d(src, q, terms=16, tree_depth=5, targets=tgt)
    op, of = orc.fmm2d(src, q, 16, 5, tgt_xy=tgt)
    assert rel_l2(pot, op) <= 1e-11 and rel_l2(field, of) <= 1e-11
    # exclude_self removes exactly q_i * 0.5 log(1e-8) from each potential
    p1, f1, _ = lfmm.fmm2d(src, q, terms=16, tree_depth=5)
    p0, f0, _ = lfmm.fmm2d(src, q, terms=16, tree_depth=5, exclude_self=True)
    assert np.allclose(p1 - p0, q * 0.5 * np.log(1e-8), rtol=1e-12, atol=1e-9)
    assert np.array_equal(f1, f0)
    # reference particle2d AoS {complex center; q; morton}: strides of 4 doubles
    aos = np.zeros((len(q), 4))
    aos[:, :2] = src
    aos[:, 2] = q
    pa = np.empty(len(q))
    fa = np.empty((len(q), 2))
    pts = lfmm.Points3d(len(q), aos.ctypes.data, 4, aos.ctypes.data + 16, 4)
    rc = lfmm.lib().lfmm_fmm2d(lfmm.default_context().handle, ctypes.byref(pts), None, 1e-8, 16, 5, 0, lfmm.MEM_HOST,
                               0, pa.ctypes.data, fa.ctypes.data, None)
    assert rc == 0 and np.array_equal(pa, p1) and np.array_equal(fa, f1)
    # device path
    pd, fd, _ = lfmm.fmm2d(torch.from_numpy(src).cuda(), torch.from_numpy(q).cuda(), terms=16, tree_depth=5)
    torch.cuda.synchronize()
    assert np.array_equal(pd.cpu().numpy(), p1) and np.array_equal(fd.cpu().numpy(), f1)


def test_fmm2d_auto_parameters(lfmm, request):
    case_name(request)
    x, q = square(100000, 41)
    _, _, rep = lfmm.fmm2d(x, q, eps=1e-8)
    assert rep["terms"] == min(30, int(-1.18 * np.log(1.65e-8)))
    assert rep["tree_depth"] == int(max(3.0, 2 + np.log(100000 / 100.0) / np.log(4.0)))
    _, _, rep = lfmm.fmm2d(x, q, terms=21, leaf_max=64)
    t = lfmm.tree2d(x, q, rep["tree_depth"])
    assert np.diff(t["begin"][-1]).max() <= 64


def test_cpp_dropin_2d_matches_python(lfmm, tmp_path, request):
    case_name(request)
    src = tmp_path / "use2d.cpp"
    src.write_text(r'''
#include "fmm/multipole.h"
#include <cstdio>
#include <random>
int main() {
    std::mt19937_64 g(7);
    std::uniform_real_distribution<double> u(0.0, 1.0), c(-1.0, 1.0);
    std::vector<fmm::particle2d> p(6000);
    for (auto& x : p) x.center = fmm::point2d(u(g), u(g));
    for (auto& x : p) x.q = c(g);
    fmm::gpu::FastMultipole f(p, 1e-8, 14, 5);
    for (size_t i = 0; i < p.size(); ++i)
        std::printf("%.17g %.17g %.17g %.17g %.17g %.17g\n", p[i].center.real(), p[i].center.imag(), p[i].q,
                    f.potentials[i], f.forces[i].real(), f.forces[i].imag());
    return 0;
}
''')
    exe = tmp_path / "use2d"
    r = subprocess.run(["g++", "-O1", "-std=c++17", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe),
                        lfmm.LIB_PATH, "-Wl,-rpath," + os.path.dirname(lfmm.LIB_PATH)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    a = np.loadtxt(out.stdout.splitlines())
    pot, field, _ = lfmm.fmm2d(a[:, :2], a[:, 2], terms=14, tree_depth=5)
    assert np.array_equal(pot, a[:, 3]) and np.array_equal(field, a[:, 4:6])


@pytest.mark.parametrize("n,N,depth", [(200000, 21, 8), (60000, 30, 6)])
def test_fmm2d_fp32_m2l(lfmm, orc, request, n, N, depth):
    """LFMM_FLAG_MIXED in 2D: FP32 per-direction M2L sums at levels >= 4.
    Only those launches change (count pinned), the result stays within 1e-6
    of the FP64 solve on the same tree, and its error against the FP64
    direct sum is the FP64 solve's plus at most 1e-6."""
    case_name(request)
    x, q = square(n, 53)
    p64, f64, r64 = lfmm.fmm2d(x, q, terms=N, tree_depth=depth, exclude_self=True, timings=True)
    p32, f32, r32 = lfmm.fmm2d(x, q, terms=N, tree_depth=depth, exclude_self=True, timings=True, mixed=True)
    assert r32["m2l_mixed_launches"] == depth - 4 and r64["m2l_mixed_launches"] == 0
    ep, ef = rel_l2(p32, p64), rel_l2(f32, f64)
    print("2D n=%d N=%d: fp32-M2L vs fp64 potential %.3g force %.3g" % (n, N, ep, ef), flush=True)
    assert 0 < ep <= 1e-6 and ef <= 1e-6
    idx = np.random.default_rng(4).choice(n, 300, replace=False)
    dp, df = orc.direct2d(x, q, tgt_index=idx, include_self=False)
    assert rel_l2(p32[idx], dp) <= rel_l2(p64[idx], dp) + 1e-6
    assert rel_l2(f32[idx], df) <= rel_l2(f64[idx], df) + 1e-6
