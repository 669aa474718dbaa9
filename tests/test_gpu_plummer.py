"""GPU: Plummer-sphere inputs (scale radius 0.05, truncated at radius 1,
centred in the unit cube; the bounding box grows to hold the points outside
it).  A uniform-depth tree chosen by leaf_max is deep and sparse here, so the
levels beyond the dense key tables (binary-search lookups) and the sparse
matrix-core packing are exercised.

Bars: FP64 path vs the CPU oracle on the same tree rel-L2 <= 1e-11; mixed vs
FP64 potential <= 1e-6, field <= 2e-6; sampled targets vs the FP64 direct sum
no worse than the oracle's own error (+1e-6 for the mixed path).
"""
import numpy as np
import pytest

from conftest import plummer, rel_l2

pytestmark = pytest.mark.gpu


def case_name(request):
    print("\n[case] %s" % request.node.name, flush=True)


@pytest.mark.parametrize("n,N", [(200000, 11), (60000, 14)])
def test_plummer_parity_accuracy_and_mixed(lfmm, orc, request, n, N):
    case_name(request)
    x, q = plummer(n, 7 + N)
    pot, field, rep = lfmm.fmm3d(x, q, terms=N, leaf_max=64, exclude_self=True)
    depth = rep["tree_depth"]
    print("plummer n=%d N=%d depth=%d max_leaf=%d" % (n, N, depth, rep["max_leaf"]), flush=True)
    assert rep["max_leaf"] <= 64 and depth >= 8  # deeper than the dense key tables (level <= 8)
    with orc.exact_geometry():
        op, of = orc.fmm3d(x, q, N, depth)
    op = op - q * 1e4  # the oracle includes the clamped self term
    assert rel_l2(pot, op) <= 1e-11 and rel_l2(field, of) <= 1e-11
    p16, f16, r16 = lfmm.fmm3d(x, q, terms=N, tree_depth=depth, exclude_self=True, mixed=True, timings=True)
    assert r16["m2l_mixed_launches"] == depth - 4
    assert rel_l2(p16, pot) <= 1e-6 and rel_l2(f16, field) <= 2e-6
    idx = np.random.default_rng(3).choice(n, 400, replace=False)
    dp, df = orc.direct3d(x, q, tgt_index=idx, include_self=False)
    e64, e16, eo = rel_l2(pot[idx], dp), rel_l2(p16[idx], dp), rel_l2(op[idx], dp)
    print("rel-L2 vs direct: fp64 %.3g mixed %.3g oracle %.3g" % (e64, e16, eo), flush=True)
    assert e64 <= 1.01 * eo + 1e-13 and e16 <= e64 + 1e-6
