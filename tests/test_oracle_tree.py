"""Pins the oracle Morton tree against a brute-force restatement written from
the definitions (keys, stable order, levels, parents, near/far sets)."""
import numpy as np
import pytest

from conftest import uniform_cube


def spread3(v):
    v = v.astype(np.uint64) & np.uint64(0x1FFFFF)
    out = np.zeros_like(v)
    for b in range(21):
        out |= ((v >> np.uint64(b)) & np.uint64(1)) << np.uint64(3 * b)
    return out


def brute_keys(xyz, depth):
    mn = xyz.min(0)
    shift = mn - 1e-7
    scale = (xyz.max(0) - mn).max() + 1e-6
    bits = depth - 1
    c = np.floor((2.0 ** bits) * ((xyz - shift) / scale)).astype(np.uint64)
    key = (spread3(c[:, 0]) << np.uint64(2)) | (spread3(c[:, 1]) << np.uint64(1)) | spread3(c[:, 2])
    return key, shift, scale


def coords(key, level):
    out = np.zeros((len(key), 3), dtype=np.int64)
    for b in range(level):
        for a, sh in ((0, 2), (1, 1), (2, 0)):
            out[:, a] |= ((key >> np.uint64(3 * b + sh)) & np.uint64(1)).astype(np.int64) << b
    return out


@pytest.mark.parametrize("n,depth,seed", [(3000, 4, 1), (5000, 3, 2), (700, 5, 3), (1, 3, 4), (64, 2, 5)])
def test_tree_matches_brute_force(orc, n, depth, seed):
    xyz, q = uniform_cube(n, seed)
    t = orc.Tree3d(xyz, q, depth)
    key, shift, scale = brute_keys(xyz, depth)
    assert np.array_equal(t.shift, shift) and t.scale == scale
    order = np.argsort(key, kind="stable")
    pos_map = np.empty(n, dtype=np.int64)
    pos_map[order] = np.arange(n)
    assert np.array_equal(t.pos_map, pos_map)
    assert np.array_equal(t.keys, key[order])
    for l in range(depth):
        lk = key[order] >> np.uint64(3 * (depth - 1 - l))
        u, first = np.unique(lk, return_index=True)
        lv = t.levels[l]
        assert np.array_equal(lv["keys"], u)
        if l == depth - 1:
            begin = np.append(first, n)
            assert np.array_equal(lv["src"][:, 0], begin[:-1]) and np.array_equal(lv["src"][:, 1], begin[1:])
        # centres are the exact dyadic cell centres
        c = coords(u, l)
        cen = ((2 * c + 1) / 2.0 ** (l + 1)) * scale + shift
        assert np.array_equal(lv["centre"], cen)
        if l + 1 < depth:
            child = t.levels[l + 1]["keys"] >> np.uint64(3)
            assert np.array_equal(u[t.levels[l + 1]["parent"]], child)


def test_near_far_sets_by_definition(orc):
    n, depth = 4000, 4
    xyz, q = uniform_cube(n, 11)
    t = orc.Tree3d(xyz, q, depth)
    for l in range(2, depth):
        lv = t.levels[l]
        c = coords(lv["keys"], l)
        pc = c // 2
        for j in range(0, len(c), max(1, len(c) // 60)):
            d = np.abs(c - c[j]).max(1)
            pd = np.abs(pc - pc[j]).max(1)
            near = set(np.nonzero((d <= 1) & (np.arange(len(c)) != j))[0])
            far = set(np.nonzero((d > 1) & (pd <= 1))[0])
            got_far = set(lv["far_idx"][lv["far_off"][j]:lv["far_off"][j + 1]])
            assert got_far == far
            got_near = list(lv["near_idx"][lv["near_off"][j]:lv["near_off"][j + 1]])
            if l == depth - 1:  # single-set tree keeps the half lists (k > j)
                assert got_near == sorted(got_near) or set(got_near) == {k for k in near if k > j}
                assert set(got_near) == {k for k in near if k > j}
            else:
                assert set(got_near) == near


def test_zero_charges_are_targets(orc):
    """Inside a leaf, q != 0 particles come first (morton_tree.cpp:426-468)."""
    n, depth = 2000, 3
    xyz, q = uniform_cube(n, 13)
    q[::3] = 0.0
    t = orc.Tree3d(xyz, q, depth)
    leaves = t.levels[-1]
    assert t.targets_num == int((q == 0).sum())
    for (sb, se), (tb, te) in zip(leaves["src"], leaves["tgt"]):
        assert se == tb
        assert np.all(t.xyzq[sb:se, 3] != 0) and np.all(t.xyzq[tb:te, 3] == 0)
