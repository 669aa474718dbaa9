"""GPU: per-box expansion orders and the learned order selector.

Pinned by:
- the order network's forward pass, loss and gradient (HIP kernels in FP32)
  against the FP64 numpy restatement oracle/order_mlp.py (itself pinned by
  finite differences): loss and gradient within 1e-5 relative;
- training on a synthetic rule the network can represent (accuracy);
- per-box orders requested per particle against the oracle's per-box
  truncation restatement on the same uniform tree: rel-L2 <= 1e-11;
- the trained selector on a held-out Plummer sphere: error <= 1e-6 against
  the FP64 direct sum with fewer M2L flops than the smallest fixed order that
  meets 1e-6 there.
"""
import os
import sys

import numpy as np
import pytest

from conftest import plummer, rel_l2, uniform_cube

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import order_mlp as om  # noqa: E402

pytestmark = pytest.mark.gpu


def case_name(request):
    print("\n[case] %s" % request.node.name, flush=True)


def random_model(rng):
    m = np.concatenate([rng.normal(0, 1, 8), rng.uniform(0.5, 2.0, 8), rng.normal(0, 0.4, 512),
                        rng.normal(0, 0.1, 64), rng.normal(0, 0.2, 4096), rng.normal(0, 0.1, 64),
                        rng.normal(0, 0.3, 832), rng.normal(0, 0.1, 13)])
    return m.astype(np.float32)


def test_order_network_matches_fp64(lfmm, request):
    case_name(request)
    rng = np.random.default_rng(11)
    model = random_model(rng)
    feats = rng.normal(0, 1.5, (1000, 8)).astype(np.float32)
    labels = rng.integers(0, 13, 1000).astype(np.int32)
    loss, g = lfmm.order_loss_grad(model, feats, labels)
    l64, g64 = om.loss_grad(model.astype(np.float64), feats.astype(np.float64), labels)
    print("loss gpu %.9g fp64 %.9g; grad rel %.3g" % (loss, l64, np.linalg.norm(g - g64) / np.linalg.norm(g64)),
          flush=True)
    assert abs(loss - l64) <= 1e-5 * abs(l64)
    assert np.all(g[:16] == 0.0)
    for a, b in ((16, 528), (528, 592), (592, 4688), (4688, 4752), (4752, 5584), (5584, 5597)):
        assert np.linalg.norm(g[a:b] - g64[a:b]) <= 1e-5 * np.linalg.norm(g64[a:b]), (a, b)
    probs, terms = lfmm.order_predict(model, feats, confidence=0.9)
    _, _, _, p64 = om.forward(model.astype(np.float64), feats.astype(np.float64))
    assert np.abs(probs - p64).max() <= 1e-5
    want = om.decide(p64, 0.9)
    cum = np.cumsum(p64, axis=1)
    close = np.any(np.abs(cum - 0.9) < 1e-5, axis=1)
    assert np.all((terms == want) | close)


def test_order_training_learns_a_rule(lfmm, request):
    case_name(request)
    rng = np.random.default_rng(12)
    f = rng.normal(0, 1, (20000, 8)).astype(np.float32)
    y = np.clip(np.round(2.0 * f[:, 0] - 1.5 * f[:, 3] + 6.0), 0, 12).astype(np.int32)
    model = lfmm.order.init_model(f, seed=5)
    trained, hist = lfmm.order_train(model, f, y, epochs=40, batch=512, lr=3e-3, seed=5)
    probs, _ = lfmm.order_predict(trained, f)
    acc = float(np.mean(np.argmax(probs, axis=1) == y))
    print("loss %.4f -> %.4f, accuracy %.3f" % (hist[0], hist[-1], acc), flush=True)
    assert hist[-1] < 0.5 * hist[0] and acc >= 0.85
    # deterministic: the same seed trains the same weights
    again, _ = lfmm.order_train(model, f, y, epochs=40, batch=512, lr=3e-3, seed=5)
    assert np.array_equal(again, trained)


@pytest.mark.parametrize("mixed", [False, True])
def test_per_box_orders_match_oracle(lfmm, orc, request, mixed):
    case_name(request)
    n = 200000
    x, q = uniform_cube(n, 71)
    pt = (3 + np.floor(7 * x[:, 0])).astype(np.int32)  # 3 .. 9 terms across x
    p, f, r = lfmm.fmm3d(x, q, tree_depth=6, leaf_max=8, particle_terms=pt, count_ops=True, mixed=mixed)
    print("depth %d leaves %d terms %d..%d m2l_flops %.4g launches %d" %
          (r["tree_depth"], r["num_leaves"], r["terms_min"], r["terms"], r["m2l_flops"], r["m2l_launches"]),
          flush=True)
    assert r["tree_depth"] == 6 and r["terms"] == 9 and r["terms_min"] == 3 and r["num_leaves"] > 30000
    with orc.exact_geometry():
        op, of = orc.fmm3d_orders(x, q, pt, 6)
    ep, ef = rel_l2(p, op), rel_l2(f, of)
    print("vs oracle: potential %.3e field %.3e" % (ep, ef), flush=True)
    if mixed:
        assert ep <= 2e-6 and ef <= 2e-6
    else:
        assert ep <= 1e-11 and ef <= 1e-10
    # fewer flops than the largest order everywhere
    _, _, r9 = lfmm.fmm3d(x, q, terms=9, tree_depth=6, leaf_max=8, adaptive=True, count_ops=True)
    assert r["m2l_flops"] < r9["m2l_flops"]


def test_learned_orders_gain_over_fixed(lfmm, orc, request):
    case_name(request)
    sets = []
    for seed, n_train in ((101, 200000), (102, 300000), (104, 400000)):
        x, q = plummer(n_train, seed)
        feats, labels, info = lfmm.order.make_labels(lfmm, x, q, leaf_max=64, tol=1e-6, safety=0.7)
        sets.append((feats, labels))
        print("seed %d: %d leaves, planned rel-L2 %.3g, label histogram %s" %
              (seed, len(labels), info["planned_rel_l2"], np.bincount(labels, minlength=13)), flush=True)
    model, hist = lfmm.order.train(lfmm, sets, epochs=40)
    print("training loss %.4f -> %.4f" % (hist[0], hist[-1]), flush=True)
    lfmm.set_order_model(model, confidence=0.5, check_tol=1e-6, check_samples=256)
    n = 300000
    x, q = plummer(n, 103)
    idx = np.random.default_rng(5).choice(n, 2000, replace=False)
    dp, _ = orc.direct3d(x, q, tgt_index=idx, include_self=False)
    p, _, r = lfmm.fmm3d(x, q, leaf_max=64, learned_order=True, exclude_self=True, count_ops=True)
    e = rel_l2(p[idx], dp)
    print("learned: terms %d..%d, m2l_flops %.4g, fallback %d, check %.3e, rel-L2 vs direct %.3e" %
          (r["terms_min"], r["terms"], r["m2l_flops"], r["fallback"], r["check_rel_l2"], e), flush=True)
    assert e <= 1e-6 and r["fallback"] == 0
    best = None
    for N in range(10, 18):
        pf, _, rf = lfmm.fmm3d(x, q, terms=N, leaf_max=64, adaptive=True, exclude_self=True, count_ops=True)
        ef = rel_l2(pf[idx], dp)
        print("fixed %d terms: rel-L2 %.3e m2l_flops %.4g" % (N, ef, rf["m2l_flops"]), flush=True)
        if ef <= 1e-6:
            best = rf
            break
    assert best is not None
    print("learned / fixed M2L flops: %.3f" % (r["m2l_flops"] / best["m2l_flops"]), flush=True)
    assert r["m2l_flops"] < best["m2l_flops"]
    lfmm.set_order_model(None)
