"""GPU parity: the HIP path (through the C ABI) against the CPU oracle.

Bars:
  tree and sort        bit-exact (geometry, stable permutation, keys, ranges)
  potential / field    rel-L2 vs the oracle <= 1e-11 / 1e-10 with the exact
                       geometry, <= 1e-8 with the reference's +1e-12 radius
                       offset in DecToSph (the product does not reproduce that
                       offset); accuracy vs the FP64 direct sum equal to the
                       oracle's own within 10 %.
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, rel_l2, uniform_cube

pytestmark = pytest.mark.gpu


def case_name(request):
    print("\n[case] %s" % request.node.name, flush=True)


def blobs(n, seed):
    rng = np.random.default_rng(seed)
    centres = rng.uniform(0.2, 0.8, (4, 3))
    pick = rng.integers(0, 4, n)
    x = centres[pick] + 0.05 * rng.standard_normal((n, 3))
    q = rng.uniform(-1, 1, n)
    return x, q


TREE_CASES = [
    ("uniform", 20000, 5, 1),
    ("uniform", 3000, 1, 2),
    ("uniform", 3000, 2, 3),
    ("blobs", 12000, 7, 4),
    ("uniform", 1, 3, 5),
    ("dups", 5000, 4, 6),
]


@pytest.mark.parametrize("kind,n,depth,seed", TREE_CASES)
def test_tree_bit_exact(lfmm, orc, request, kind, n, depth, seed):
    case_name(request)
    if kind == "blobs":
        x, q = blobs(n, seed)
    else:
        x, q = uniform_cube(n, seed)
        if kind == "dups":
            x[n // 2:] = x[: n - n // 2]  # exact duplicates
    g = lfmm.tree3d(x, q, depth)
    t = orc.Tree3d(x, q, depth)
    assert g["depth"] == depth
    assert np.array_equal(g["level_sizes"], t.level_sizes)
    assert np.array_equal(g["shift"], t.shift) and g["scale"] == t.scale
    assert np.array_equal(g["pos_map"], t.pos_map)
    for l in range(depth):
        assert np.array_equal(g["keys"][l], t.levels[l]["keys"])
        src = t.levels[l]["src"]
        assert np.array_equal(g["begin"][l][:-1], src[:, 0]) and g["begin"][l][-1] == src[-1, 1]


@pytest.mark.parametrize("N", [1, 2, 5, 9, 11, 14, 17, 18])
def test_fmm_parity_with_oracle(lfmm, orc, request, N):
    case_name(request)
    x, q = uniform_cube(6000, 100 + N)
    depth = 4
    pot, field, rep = lfmm.fmm3d(x, q, terms=N, tree_depth=depth)
    assert rep["terms"] == N and rep["tree_depth"] == depth
    with orc.exact_geometry():
        op, of = orc.fmm3d(x, q, N, depth)
    assert rel_l2(pot, op) <= 1e-11
    assert rel_l2(field, of) <= 1e-10
    rp, rf = orc.fmm3d(x, q, N, depth)
    assert rel_l2(pot, rp) <= 1e-8
    assert rel_l2(field, rf) <= 1e-8


@pytest.mark.parametrize("dist", ["uniform", "blobs"])
def test_accuracy_vs_direct_matches_oracle(lfmm, orc, request, dist):
    case_name(request)
    n = 20000
    x, q = (uniform_cube(n, 7) if dist == "uniform" else blobs(n, 8))
    idx = np.random.default_rng(9).choice(n, 1000, replace=False)
    dp, df = orc.direct3d(x, q, tgt_index=idx, include_self=False)
    for N in (6, 11):
        pot, field, _ = lfmm.fmm3d(x, q, terms=N, tree_depth=5, exclude_self=True)
        op, of = orc.fmm3d(x, q, N, 5)
        eg, eo = rel_l2(pot[idx], dp), rel_l2(op[idx] - q[idx] * 1e4, dp)
        assert eg <= 1.1 * eo + 1e-12
        assert rel_l2(field[idx], df) <= 1.1 * rel_l2(of[idx], df) + 1e-12


def test_exclude_self_flag(lfmm, request):
    case_name(request)
    x, q = uniform_cube(5000, 12)
    p1, f1, _ = lfmm.fmm3d(x, q, terms=9, tree_depth=4)
    p0, f0, _ = lfmm.fmm3d(x, q, terms=9, tree_depth=4, exclude_self=True)
    assert np.allclose(p1 - p0, q * 1e4, rtol=1e-12, atol=1e-9)
    assert np.array_equal(f1, f0)


def test_input_layouts_agree(lfmm, request):
    case_name(request)
    import torch
    x, q = uniform_cube(7000, 13)
    p_ref, f_ref, _ = lfmm.fmm3d(x, q, terms=8, tree_depth=4)
    # reference AoS particle {center[3], q, morton}: strides of 5 doubles
    aos = np.zeros((len(q), 5))
    aos[:, :3] = x
    aos[:, 3] = q
    pot = np.empty(len(q))
    field = np.empty((len(q), 3))
    pts = lfmm.Points3d(len(q), aos.ctypes.data, 5, aos.ctypes.data + 24, 5)
    rep = lfmm.Report()
    import ctypes
    rc = lfmm.lib().lfmm_fmm3d(lfmm.default_context().handle, ctypes.byref(pts), None, 1e-8, 8, 4, 0,
                               lfmm.MEM_HOST, 0, pot.ctypes.data, field.ctypes.data, ctypes.byref(rep))
    assert rc == 0
    assert np.array_equal(pot, p_ref) and np.array_equal(field, f_ref)
    xd = torch.from_numpy(x).cuda()
    qd = torch.from_numpy(q).cuda()
    pd, fd, _ = lfmm.fmm3d(xd, qd, terms=8, tree_depth=4)
    torch.cuda.synchronize()
    assert np.array_equal(pd.cpu().numpy(), p_ref) and np.array_equal(fd.cpu().numpy(), f_ref)


def test_edge_cases(lfmm, orc, request):
    case_name(request)
    # one particle: only the clamped self term
    p, f, _ = lfmm.fmm3d(np.array([[0.3, 0.4, 0.5]]), np.array([2.0]), terms=5, tree_depth=3)
    assert p[0] == pytest.approx(2.0e4) and np.all(f == 0)
    # all particles at one point
    x = np.tile([[0.5, 0.5, 0.5]], (100, 1))
    q = np.linspace(-1, 1, 100)
    p, f, _ = lfmm.fmm3d(x, q, terms=5, tree_depth=3)
    assert np.allclose(p, q.sum() * 1e4) and np.all(f == 0)
    # planar set, depth 1 and 2 (near field only)
    x, q = uniform_cube(500, 14)
    x[:, 2] = 0.25
    for depth in (1, 2, 3):
        p, f, _ = lfmm.fmm3d(x, q, terms=7, tree_depth=depth)
        with orc.exact_geometry():
            op, of = orc.fmm3d(x, q, 7, depth)
        assert rel_l2(p, op) <= 1e-11 and rel_l2(f, of) <= 1e-10


def test_errors_are_reported(lfmm, request):
    case_name(request)
    x, q = uniform_cube(100, 15)
    bad = x.copy()
    bad[3, 1] = np.nan
    for kw, code in ((dict(xyz=bad, q=q, terms=5, tree_depth=3), 6), (dict(xyz=x, q=q, terms=0, tree_depth=3), 2),
                     (dict(xyz=x, q=q, terms=5, tree_depth=0), 3), (dict(xyz=x, q=q, terms=5, tree_depth=17), 3)):
        with pytest.raises(lfmm.LfmmError) as e:
            lfmm.fmm3d(**kw)
        assert e.value.code == code
    import ctypes
    pts = lfmm.Points3d(100, x.ctypes.data, 3, q.ctypes.data, 1)
    tg = lfmm.Points3d(10, None, 3, None, 1)  # targets without positions
    pot = np.empty(10)
    rc = lfmm.lib().lfmm_fmm3d(lfmm.default_context().handle, ctypes.byref(pts), ctypes.byref(tg), 1e-8, 5, 3, 0,
                               lfmm.MEM_HOST, 0, pot.ctypes.data, None, None)
    assert rc == 1


@pytest.mark.parametrize("depth,N", [(4, 9), (5, 12), (2, 6)])
def test_source_target_mode_parity(lfmm, orc, request, depth, N):
    """Targets as q = 0 particles of one tree (src/cpu/multipole3d.cpp:1311-1319)."""
    case_name(request)
    src, q = uniform_cube(9000, 40 + depth)
    tgt, _ = blobs(3000, 50 + depth)
    pot, field, _ = lfmm.fmm3d(src, q, terms=N, tree_depth=depth, targets=tgt)
    with orc.exact_geometry():
        op, of = orc.fmm3d(src, q, N, depth, tgt_xyz=tgt)
    assert rel_l2(pot, op) <= 1e-11 and rel_l2(field, of) <= 1e-10
    dp, df = orc.direct3d(src, q, tgt_xyz=tgt, include_self=True)
    assert rel_l2(pot, dp) < 1e-3
    # the device path agrees with the host path bit for bit
    import torch
    pd, fd, _ = lfmm.fmm3d(torch.from_numpy(src).cuda(), torch.from_numpy(q).cuda(), terms=N, tree_depth=depth,
                           targets=torch.from_numpy(tgt).cuda())
    torch.cuda.synchronize()
    assert np.array_equal(pd.cpu().numpy(), pot) and np.array_equal(fd.cpu().numpy(), field)


def test_tree_partition_of_zero_charges(lfmm, orc, request):
    case_name(request)
    x, q = uniform_cube(8000, 61)
    q[::4] = 0.0
    g = lfmm.tree3d(x, q, 4)
    t = orc.Tree3d(x, q, 4)
    assert np.array_equal(g["pos_map"], t.pos_map)
    for l in range(4):
        assert np.array_equal(g["keys"][l], t.levels[l]["keys"])


def test_auto_parameters(lfmm, request):
    case_name(request)
    x, q = uniform_cube(60000, 16)
    _, _, rep = lfmm.fmm3d(x, q, eps=1e-8)
    assert rep["terms"] == 18
    assert rep["tree_depth"] == int(max(2.0, 1 + np.log(60000 / 100.0) / np.log(8.0)))
    _, _, rep = lfmm.fmm3d(x, q, eps=1e-4)
    assert rep["terms"] == min(18, int(54.2 - 10.73 * np.sqrt(np.log(2.4e-4) + 25.5)))
    _, _, rep = lfmm.fmm3d(x, q, terms=9, leaf_max=64)
    d = rep["tree_depth"]
    key = lambda lev: np.floor((x - (x.min(0) - 1e-7)) / ((x.max(0) - x.min(0)).max() + 1e-6) * 2 ** lev).astype(np.int64)
    occ = lambda lev: np.unique(key(lev) @ np.array([1 << 40, 1 << 20, 1]), return_counts=True)[1].max()
    assert occ(d - 1) <= 64 and occ(d - 2) > 64
    assert rep["max_leaf"] == occ(d - 1)


def test_deterministic(lfmm, request):
    case_name(request)
    x, q = uniform_cube(30000, 17)
    a = lfmm.fmm3d(x, q, terms=11, tree_depth=5)
    b = lfmm.fmm3d(x, q, terms=11, tree_depth=5)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_full_size_4m_sampled_vs_direct(lfmm, orc, request):
    """The headline workload (4,194,304 uniform, p = 10, leaf level 6): sampled
    targets against the FP64 direct sum, self term excluded on both sides."""
    case_name(request)
    import torch
    n = 4194304
    x, q = uniform_cube(n, 1)
    xd, qd = torch.from_numpy(x).cuda(), torch.from_numpy(q).cuda()
    pot, field, rep = lfmm.fmm3d(xd, qd, terms=11, tree_depth=7, exclude_self=True)
    torch.cuda.synchronize()
    assert rep["max_leaf"] <= 64
    idx = np.random.default_rng(2).choice(n, 512, replace=False)
    dp, df = orc.direct3d(x, q, tgt_index=idx, include_self=False)
    p = pot.cpu().numpy()[idx]
    f = field.cpu().numpy()[idx]
    assert rel_l2(p, dp) < 2e-5
    assert rel_l2(f, df) < 1e-4


def test_cpp_dropin_matches_python(lfmm, tmp_path, request):
    case_name(request)
    src = tmp_path / "use.cpp"
    src.write_text(r'''
#include "fmm/multipole3d.h"
#include <cstdio>
#include <random>
int main() {
    std::mt19937_64 g(5);
    std::uniform_real_distribution<double> u(0.0, 1.0), c(-1.0, 1.0);
    std::vector<fmm::particle3d> p(5000);
    for (auto& x : p) { x.center = fmm::Vector3d(u(g), u(g), u(g)); }
    for (auto& x : p) x.q = c(g);
    fmm::gpu::FastMultipole3d<double> f(p, 1e-8, 9, 4);
    for (size_t i = 0; i < p.size(); ++i)
        std::printf("%.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n", p[i].center[0], p[i].center[1],
                    p[i].center[2], p[i].q, f.potentials[i], f.forces[i][0], f.forces[i][1], f.forces[i][2]);
    return 0;
}
''')
    exe = tmp_path / "use"
    r = subprocess.run(["g++", "-O1", "-std=c++17", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe),
                        lfmm.LIB_PATH, "-Wl,-rpath," + os.path.dirname(lfmm.LIB_PATH)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    a = np.loadtxt(out.stdout.splitlines())
    pot, field, _ = lfmm.fmm3d(a[:, :3], a[:, 3], terms=9, tree_depth=4)
    assert np.array_equal(pot, a[:, 4]) and np.array_equal(field, a[:, 5:8])


def test_cpp_options_overload_matches_python(lfmm, tmp_path, request):
    """The (particles, {p, leaf_max, reference_conventions}) constructors
    give the library's own results for terms p + 1, that leaf bound and the
    self-term switch, on the uniform and the adaptive tree."""
    case_name(request)
    src = tmp_path / "use.cpp"
    src.write_text(r'''
#include "fmm/multipole3d.h"
#include <cstdio>
#include <random>
int main() {
    std::mt19937_64 g(9);
    std::uniform_real_distribution<double> u(0.0, 1.0), c(-1.0, 1.0);
    std::vector<fmm::particle3d> p(20000);
    for (auto& x : p) { x.center = fmm::Vector3d(u(g), u(g), u(g)); }
    for (auto& x : p) x.q = c(g);
    fmm::gpu::FastMultipole3d a(p, {10, 32});                // reference conventions (self pair included)
    fmm::gpu::FastMultipole3d b(p, fmm::Options(10, 32, false));
    fmm::Options o(7, 48, false);
    o.adaptive = true;
    fmm::gpu::FastMultipole3d<double> d(p, o);
    for (size_t i = 0; i < p.size(); ++i)
        std::printf("%.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n", p[i].center[0], p[i].center[1],
                    p[i].center[2], p[i].q, a.potentials[i], b.potentials[i], b.forces[i][0], d.potentials[i],
                    d.forces[i][2]);
    return 0;
}
''')
    exe = tmp_path / "use"
    r = subprocess.run(["g++", "-O1", "-std=c++17", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe),
                        lfmm.LIB_PATH, "-Wl,-rpath," + os.path.dirname(lfmm.LIB_PATH)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    a = np.loadtxt(out.stdout.splitlines())
    x, q = a[:, :3], a[:, 3]
    pa, _, ra = lfmm.fmm3d(x, q, terms=11, leaf_max=32)
    pb, fb, _ = lfmm.fmm3d(x, q, terms=11, leaf_max=32, exclude_self=True)
    pd, fd, _ = lfmm.fmm3d(x, q, terms=8, leaf_max=48, adaptive=True, exclude_self=True)
    assert ra["terms"] == 11 and ra["max_leaf"] <= 32
    assert np.array_equal(pa, a[:, 4]) and np.array_equal(pb, a[:, 5]) and np.array_equal(fb[:, 0], a[:, 6])
    assert np.array_equal(pd, a[:, 7]) and np.array_equal(fd[:, 2], a[:, 8])
    # the self pair at the clamped distance: q_i / 1e-4 apart
    assert np.allclose(a[:, 4] - a[:, 5], q / 1e-4, rtol=1e-9, atol=1e-9)
