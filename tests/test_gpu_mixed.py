"""GPU: M2L on the matrix cores in split FP16 (LFMM_FLAG_MIXED) against the
FP64 path and the FP64 direct sum.

Stated error of the mixed path: operands carry 22 significant bits (fp16 hi +
lo, the lo*lo product dropped) and sums are FP32, so each local expansion
coefficient of levels >= 4 is within ~1e-7 relative of its FP64 value.  The
bars below (self term excluded, so the relative norms measure the
interaction sums):
  potential  rel-L2(mixed - fp64) <= 1e-6,  field <= 2e-6
  accuracy vs the FP64 direct sum: mixed error <= fp64 error + 1e-6
"""
import numpy as np
import pytest

from conftest import rel_l2, uniform_cube

pytestmark = pytest.mark.gpu


def case_name(request):
    print("\n[case] %s" % request.node.name, flush=True)


def blobs(n, seed):
    rng = np.random.default_rng(seed)
    centres = rng.uniform(0.2, 0.8, (4, 3))
    pick = rng.integers(0, 4, n)
    x = centres[pick] + 0.05 * rng.standard_normal((n, 3))
    return x, rng.uniform(-1, 1, n)


@pytest.mark.parametrize("N", [1, 2, 4, 9, 11, 12, 13, 14, 15, 17, 18])
def test_mixed_matches_fp64(lfmm, request, N):
    case_name(request)
    x, q = uniform_cube(40000, 300 + N)
    p64, f64, _ = lfmm.fmm3d(x, q, terms=N, tree_depth=6, exclude_self=True)
    p16, f16, rep = lfmm.fmm3d(x, q, terms=N, tree_depth=6, exclude_self=True, mixed=True, timings=True)
    assert rep["m2l_mixed_launches"] == 2  # levels 4 and 5
    assert rel_l2(p16, p64) <= 1e-6 and rel_l2(f16, f64) <= 2e-6
    assert not np.array_equal(p16, p64)  # the matrix-core path did run


@pytest.mark.parametrize("dist,n,depth", [("uniform", 60000, 6), ("blobs", 30000, 7), ("blobs", 20000, 10)])
def test_mixed_accuracy_vs_direct(lfmm, orc, request, dist, n, depth):
    """Sparse trees (missing sources read the zero box) and, at depth 10, the
    binary-search lookup of levels above the dense key tables."""
    case_name(request)
    x, q = uniform_cube(n, 31) if dist == "uniform" else blobs(n, 32)
    idx = np.random.default_rng(33).choice(n, 800, replace=False)
    dp, df = orc.direct3d(x, q, tgt_index=idx, include_self=False)
    for N in (8, 11):
        p64, f64, _ = lfmm.fmm3d(x, q, terms=N, tree_depth=depth, exclude_self=True)
        p16, f16, rep = lfmm.fmm3d(x, q, terms=N, tree_depth=depth, exclude_self=True, mixed=True)
        assert rep["m2l_mixed_launches"] == depth - 4  # levels 4 .. depth - 1
        assert rel_l2(p16, p64) <= 1e-6 and rel_l2(f16, f64) <= 2e-6
        assert rel_l2(p16[idx], dp) <= rel_l2(p64[idx], dp) + 1e-6
        assert rel_l2(f16[idx], df) <= rel_l2(f64[idx], df) + 1e-6


def test_mixed_source_target_and_device_path(lfmm, request):
    case_name(request)
    import torch
    src, q = uniform_cube(30000, 41)
    tgt, _ = blobs(5000, 42)
    p64, f64, _ = lfmm.fmm3d(src, q, terms=10, tree_depth=6, targets=tgt)
    p16, f16, _ = lfmm.fmm3d(src, q, terms=10, tree_depth=6, targets=tgt, mixed=True)
    assert rel_l2(p16, p64) <= 1e-6 and rel_l2(f16, f64) <= 2e-6
    pd, fd, _ = lfmm.fmm3d(torch.from_numpy(src).cuda(), torch.from_numpy(q).cuda(), terms=10, tree_depth=6,
                           targets=torch.from_numpy(tgt).cuda(), mixed=True)
    torch.cuda.synchronize()
    assert np.array_equal(pd.cpu().numpy(), p16) and np.array_equal(fd.cpu().numpy(), f16)


def test_mixed_is_deterministic_and_shallow_trees_stay_fp64(lfmm, request):
    case_name(request)
    x, q = uniform_cube(20000, 51)
    a = lfmm.fmm3d(x, q, terms=11, tree_depth=6, mixed=True)
    b = lfmm.fmm3d(x, q, terms=11, tree_depth=6, mixed=True)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    # depth 4: M2L only at levels 2-3, which stay FP64 -> identical to the FP64 solve
    p64, f64, _ = lfmm.fmm3d(x, q, terms=11, tree_depth=4)
    p16, f16, rep = lfmm.fmm3d(x, q, terms=11, tree_depth=4, mixed=True)
    assert rep["m2l_mixed_launches"] == 0
    assert np.array_equal(p16, p64) and np.array_equal(f16, f64)
