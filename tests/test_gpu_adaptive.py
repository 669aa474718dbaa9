"""GPU: adaptive octree (LFMM_FLAG_ADAPTIVE).

The reference has uniform-depth trees only; the adaptive solve adds P2L
(X lists) and M2P (W lists) to M2L and P2P.  Pinned by:
- uniform input, where every leaf of the adaptive tree sits on one level:
  the same tree, lists without X or W pairs, and the same results as the
  uniform solve to round-off (<= 1e-14 relative: the same operators and
  summation order, but the uniform L2P kernel is compiled with its
  coefficient loads grouped differently, so the compiler's FMA contraction
  of the field may differ in the last bits);
- clustered input against the FP64 direct sum: geometric convergence with
  the number of terms (1.3e-3, 2.4e-5, 8.2e-7, 2.5e-8 at 5, 9, 13, 18 terms);
- the Plummer sphere of the headline workload, at 300,000 and at the full
  4,194,304 particles, against the direct sum at <= 1e-6, FP64 and mixed.
"""
import numpy as np
import pytest

from conftest import plummer, rel_l2, uniform_cube

pytestmark = pytest.mark.gpu


def case_name(request):
    print("\n[case] %s" % request.node.name, flush=True)


def clusters(n, seed):
    """Two dense Plummer-like clusters of very different size plus a uniform
    background: leaves on many levels, many X and W pairs."""
    rng = np.random.default_rng(seed)
    n1, n2 = n // 2, n // 4
    n3 = n - n1 - n2
    x1, _ = plummer(n1, seed + 1, a=0.02, rmax=0.3)
    x2, _ = plummer(n2, seed + 2, a=0.004, rmax=0.05)
    x2 = x2 - 0.5 + np.array([0.8, 0.2, 0.7])
    x3 = rng.random((n3, 3))
    x = np.concatenate([x1, x2, x3])
    q = rng.uniform(-1.0, 1.0, n)
    return x, q


def test_adaptive_uniform_input_equals_uniform_tree(lfmm, request):
    case_name(request)
    x, q = uniform_cube(100000, 31)
    pu, fu, ru = lfmm.fmm3d(x, q, terms=9, leaf_max=64)
    pa, fa, ra = lfmm.fmm3d(x, q, terms=9, leaf_max=64, adaptive=True, count_ops=True)
    print("uniform: depth %d leaves %d / adaptive: depth %d leaves %d boxes %d xw %d" %
          (ru["tree_depth"], ru["num_leaves"], ra["tree_depth"], ra["num_leaves"], ra["num_boxes"], ra["xw_pairs"]),
          flush=True)
    assert ra["tree_depth"] == ru["tree_depth"] and ra["num_leaves"] == ru["num_leaves"]
    assert ra["xw_pairs"] == 0
    dp, df = rel_l2(pa, pu), rel_l2(fa, fu)
    print("adaptive vs uniform: potential %.2e field %.2e" % (dp, df), flush=True)
    assert dp <= 1e-14 and df <= 1e-14


def test_adaptive_clustered_converges_to_direct(lfmm, orc, request):
    case_name(request)
    n = 30000
    x, q = clusters(n, 51)
    dp, df = orc.direct3d(x, q, tgt_index=np.arange(n), include_self=False)
    errs = []
    for N in (5, 9, 13, 18):
        p, f, r = lfmm.fmm3d(x, q, terms=N, leaf_max=16, adaptive=True, exclude_self=True, count_ops=True)
        e, ef = rel_l2(p, dp), rel_l2(f, df)
        errs.append(e)
        print("N=%d depth %d leaves %d boxes %d xw %d max_leaf %d: potential %.3e field %.3e" %
              (N, r["tree_depth"], r["num_leaves"], r["num_boxes"], r["xw_pairs"], r["max_leaf"], e, ef), flush=True)
        assert r["max_leaf"] <= 16 and r["xw_pairs"] > 0
    # geometric convergence (about 0.4-0.5 per term on this input)
    assert errs[1] < 0.1 * errs[0] and errs[2] < 0.1 * errs[1] and errs[3] < 0.1 * errs[2]
    assert errs[3] <= 1e-7


@pytest.mark.parametrize("mixed", [False, True])
def test_adaptive_plummer_vs_direct(lfmm, orc, request, mixed):
    case_name(request)
    n = 300000
    x, q = plummer(n, 61)
    p, f, r = lfmm.fmm3d(x, q, terms=13, leaf_max=64, adaptive=True, exclude_self=True, count_ops=True, mixed=mixed)
    idx = np.random.default_rng(7).choice(n, 2000, replace=False)
    dp, df = orc.direct3d(x, q, tgt_index=idx, include_self=False)
    e, ef = rel_l2(p[idx], dp), rel_l2(f[idx], df)
    print("plummer n=%d mixed=%s depth %d leaves %d boxes %d xw %d m2l %d p2p %d: potential %.3e field %.3e" %
          (n, mixed, r["tree_depth"], r["num_leaves"], r["num_boxes"], r["xw_pairs"], r["m2l_ops"], r["p2p_pairs"],
           e, ef), flush=True)
    assert r["max_leaf"] <= 64 and r["xw_pairs"] > 0
    assert e <= 1e-6


def test_adaptive_plummer_4m_full_size(lfmm, orc, request):
    """The headline Plummer workload at full size: 4,194,304 particles
    (a = 0.05, r <= 1), adaptive octree with leaf_max 64, 13 terms, FP64 and
    mixed, on 4,096 sampled targets against the FP64 direct sum."""
    case_name(request)
    n = 4194304
    x, q = plummer(n, 23)
    idx = np.random.default_rng(24).choice(n, 4096, replace=False)
    dp, df = orc.direct3d(x, q, tgt_index=idx, include_self=False)
    for mixed in (False, True):
        p, f, r = lfmm.fmm3d(x, q, terms=13, leaf_max=64, adaptive=True, exclude_self=True, mixed=mixed)
        e, ef = rel_l2(p[idx], dp), rel_l2(f[idx], df)
        print("plummer 4M mixed=%s depth %d leaves %d boxes %d xw %d: potential %.3e field %.3e" %
              (mixed, r["tree_depth"], r["num_leaves"], r["num_boxes"], r["xw_pairs"], e, ef), flush=True)
        assert r["max_leaf"] <= 64 and r["tree_depth"] >= 10
        assert e <= 1e-6 and ef <= 1e-5


@pytest.mark.parametrize("seed,n,leaf_max", [(3, 1500, 8), (8, 6000, 32), (9, 20000, 64)])
def test_adaptive_lists_match_definitions(lfmm, request, seed, n, leaf_max):
    """Boxes, U, W and X lists of the GPU build against the brute-force
    restatement of their definitions (oracle/adaptive.py); V lists (key
    lookups inside the M2L kernels) through the M2L operation count."""
    case_name(request)
    import adaptive as ad
    x, q = plummer(n, seed, a=0.03)
    t = lfmm.tree3d_adaptive(x, leaf_max=leaf_max)
    T = ad.Tree(x, t["geom"], leaf_max, t["depth"])
    gpu_boxes = {(int(l), int(k)): int(r[1] - r[0]) for l, k, r in zip(t["box_level"], t["box_key"], t["box_range"])}
    ora_boxes = {(int(l), int(k)): int(c) for l, k, c in zip(T.level, T.key, T.count)}
    assert gpu_boxes == ora_boxes
    oid = {(int(l), int(k)): i for i, (l, k) in enumerate(zip(T.level, T.key))}
    gb = [oid[(int(l), int(k))] for l, k in zip(t["box_level"], t["box_key"])]
    leaf_o = [gb[g] for g in t["leaf_box"]]
    U, V, W, X = T.lists()
    assert sorted(leaf_o) == sorted(int(b) for b in np.nonzero(T.leaf)[0])
    for i, b in enumerate(leaf_o):
        assert sorted(leaf_o[j] for j in t["u_leaf"][t["u_off"][i]:t["u_off"][i + 1]]) == U[b]
        assert sorted(gb[g] for g in t["w_box"][t["w_off"][i]:t["w_off"][i + 1]]) == W[b]
    for g in range(len(gb)):
        assert sorted(leaf_o[j] for j in t["x_leaf"][t["x_off"][g]:t["x_off"][g + 1]]) == X[gb[g]]
    _, _, r = lfmm.fmm3d(x, q, terms=4, leaf_max=leaf_max, adaptive=True, count_ops=True)
    assert r["m2l_ops"] == sum(len(v) for v in V.values())
    assert r["xw_pairs"] == sum(len(v) for v in X.values())
    # the pair kernel's near lists: U, then X into a leaf and W boxes of at
    # most 96 particles as direct sums (cheaper than P2L / M2P, and exact)
    near = {b: sum(T.count[c] for c in us) + sum(T.count[c] for c in X[b]) +
            sum(T.count[c] for c in W[b] if T.count[c] <= 96) for b, us in U.items()}
    assert r["p2p_pairs"] == sum(T.count[b] * near[b] for b in U)
    print("n=%d leaf_max=%d: %d boxes, %d leaves, %d X pairs, %d V pairs" %
          (n, leaf_max, len(gb), len(leaf_o), r["xw_pairs"], r["m2l_ops"]), flush=True)
