"""CPU: the adaptive-octree restatement in oracle/adaptive.py (the checker of
the GPU adaptive tree).

Pinned by the defining property of the four lists: every (target leaf,
source leaf) pair is covered exactly once by U, the V pairs of their
ancestors, X of a target ancestor or W of a source ancestor.  A planted
mistake (a dropped X pair, a dropped U entry, a W pair counted twice) must
break it.  Uniform input gives the uniform tree: one leaf level, empty X and
W lists, 27-neighbourhood U lists.
"""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import adaptive as ad  # noqa: E402

from conftest import plummer, uniform_cube  # noqa: E402


def geom_of(x):
    # the reference's extent rule (src/cpu/morton_tree.cpp:293-301)
    lo, hi = x.min(0), x.max(0)
    return np.r_[lo - 1e-7, (hi - lo).max() + 1e-6]


def depth_for(x, geom, leaf_max):
    for level in range(16):
        _, cnt = np.unique(ad.morton(ad.cells(x, geom, level)), return_counts=True)
        if cnt.max() <= leaf_max:
            return level + 1
    return 16


def build(x, leaf_max):
    g = geom_of(x)
    t = ad.Tree(x, g, leaf_max, depth_for(x, g, leaf_max))
    return t, t.lists()


def test_clustered_lists_cover_every_pair_once():
    x, _ = plummer(1500, 3, a=0.03)
    t, lists = build(x, 8)
    U, V, W, X = lists
    assert t.level.max() >= 6 and len(set(t.level[t.leaf])) >= 4  # leaves on many levels
    assert sum(len(v) for v in X.values()) > 100
    assert ad.partition_check(t, lists) == 0
    # leaves hold at most leaf_max particles unless on the deepest level
    assert np.all((t.count[t.leaf] <= 8) | (t.level[t.leaf] == t.level.max()))
    # X and W are transposes
    assert sum(len(v) for v in X.values()) == sum(len(v) for v in W.values())


def test_planted_mistakes_are_detected():
    x, _ = plummer(800, 4, a=0.03)
    t, (U, V, W, X) = build(x, 8)
    b = next(k for k, v in X.items() if v)
    X2 = dict(X)
    X2[b] = X[b][1:]
    assert ad.partition_check(t, (U, V, W, X2)) > 0
    leaf = next(k for k, v in U.items() if len(v) > 1)
    U2 = dict(U)
    U2[leaf] = [c for c in U[leaf] if c != leaf]
    assert ad.partition_check(t, (U2, V, W, X)) > 0
    w = next(k for k, v in W.items() if v)
    W2 = dict(W)
    W2[w] = W[w] + W[w][:1]
    assert ad.partition_check(t, (U, V, W2, X)) > 0


def test_uniform_input_gives_uniform_tree():
    x, _ = uniform_cube(12000, 5)  # level 2 boxes all above 64, level 3 all below
    t, (U, V, W, X) = build(x, 64)
    assert len(set(t.level[t.leaf])) == 1
    assert not any(X.values()) and not any(W.values())
    assert max(len(v) for v in U.values()) == 27
