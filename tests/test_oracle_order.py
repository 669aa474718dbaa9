"""CPU: the FP64 numpy restatement of the order network (oracle/order_mlp.py),
the checker of the GPU forward / backward kernels.

Pinned by central finite differences of the mean cross-entropy on every
trained parameter group, by a planted mistake (a transposed layer) that the
difference check must catch, and by the decision rule on hand-made
probabilities.
"""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import order_mlp as om  # noqa: E402


def random_model(rng):
    m = np.concatenate([rng.normal(0, 1, 8), rng.uniform(0.5, 2.0, 8), rng.normal(0, 0.4, om.H * om.F),
                        rng.normal(0, 0.1, om.H), rng.normal(0, 0.2, om.H * om.H), rng.normal(0, 0.1, om.H),
                        rng.normal(0, 0.3, om.C * om.H), rng.normal(0, 0.1, om.C)])
    assert len(m) == 5597
    return m


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    m = random_model(rng)
    f = rng.normal(0, 1.5, (40, 8))
    y = rng.integers(0, 13, 40)
    _, g = om.loss_grad(m, f, y)
    assert np.all(g[:16] == 0.0)
    h = 1e-6
    # a sample of entries from every trained group: W1, b1, W2, b2, W3, b3
    for e in list(rng.integers(16, 528, 6)) + [530, 600, 3000, 4650, 4700, 5000, 5590]:
        mp, mm = m.copy(), m.copy()
        mp[e] += h
        mm[e] -= h
        fd = (om.loss_grad(mp, f, y)[0] - om.loss_grad(mm, f, y)[0]) / (2 * h)
        assert abs(fd - g[e]) <= 1e-6 * max(1.0, abs(fd)), (e, fd, g[e])


def test_planted_transpose_is_caught():
    rng = np.random.default_rng(4)
    m = random_model(rng)
    f = rng.normal(0, 1.5, (20, 8))
    y = rng.integers(0, 13, 20)
    _, g = om.loss_grad(m, f, y)
    # a wrong backward pass through layer 2 (W2 not transposed) differs
    mu, sd, w1, b1, w2, b2, w3, b3 = om.unpack(m)
    x, h1, h2, p = om.forward(m, f)
    dz = p.copy()
    dz[np.arange(20), y] -= 1.0
    dz /= 20
    d2 = (dz @ w3) * (h2 > 0)
    d1_wrong = (d2 @ w2.T) * (h1 > 0)
    gw1_wrong = d1_wrong.T @ x
    assert np.abs(gw1_wrong.ravel() - g[16:16 + 512]).max() > 1e-4


def test_decision_rule():
    p = np.zeros((3, 13))
    p[0, 0] = 1.0
    p[1, [2, 5]] = [0.5, 0.5]
    p[2, :] = 1.0 / 13
    assert list(om.decide(p, 0.9)) == [5, 10, 16]
    assert list(om.decide(p, 0.5)) == [5, 7, 11]
