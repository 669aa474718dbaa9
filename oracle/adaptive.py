"""Adaptive octree and its interaction lists, restated from their definitions
(test infrastructure: the checker of the GPU adaptive tree; never shipped).

The reference builds uniform-depth trees only (src/cpu/morton_tree.cpp:
371-423, lists :147-208), so there is no reference code for this part.  The
restatement follows the published adaptive FMM (Carrier, Greengard and
Rokhlin 1988), on the reference's Morton grid (keys as
src/cpu/morton_tree.cpp:66-99, geometry :293-301):
  - a box is split while it holds more than leaf_max particles and is above
    the deepest level; its children are its non-empty octants;
  - U(b), leaf b:  leaves touching b (b included);
  - V(b), level >= 2: boxes of b's level not touching b whose parents touch;
  - W(b), leaf b:  boxes c finer than b, not touching b, whose parent touches b;
  - X(b):          leaves c with b in W(c).
"Touch" is closed-interval contact of the boxes' cells.  Everything is brute
force over all box pairs: small inputs only.

Pinned by `partition_check`: every (target leaf, source leaf) pair is covered
by exactly one of U, the V pairs of their ancestors, X of a target ancestor
and W of a source ancestor -- the condition under which the FMM sums every
pair once.
"""
import numpy as np


def cells(x, geom, level):
    """Integer cell coordinates of points at a level, as the GPU computes them
    (floor(2^level * (p - shift) / scale), src/cpu/morton_tree.cpp:66-82)."""
    shift = np.asarray(geom[:3], dtype=np.float64)
    scale = float(geom[3])
    u = (np.asarray(x, dtype=np.float64) - shift) / scale
    return np.floor(u * float(1 << level)).astype(np.int64)


def spread3(v):
    v = np.asarray(v, dtype=np.uint64) & np.uint64(0x1fffff)
    out = np.zeros_like(v)
    for b in range(21):
        out |= ((v >> np.uint64(b)) & np.uint64(1)) << np.uint64(3 * b)
    return out


def morton(c):
    c = np.asarray(c)
    return (spread3(c[..., 0]) << np.uint64(2)) | (spread3(c[..., 1]) << np.uint64(1)) | spread3(c[..., 2])


class Tree:
    """Boxes as arrays: level, cell (3 ints), key, particle indices, parent,
    children, leaf flag.  Box 0 is the root."""

    def __init__(self, x, geom, leaf_max, depth):
        self.level, self.cell, self.members, self.parent, self.children = [], [], [], [], []
        n = len(x)
        cell_at = [cells(x, geom, l) for l in range(depth)]
        self._add(0, np.zeros(3, np.int64), np.arange(n), -1)
        i = 0
        while i < len(self.level):
            l, m = self.level[i], self.members[i]
            if len(m) > leaf_max and l + 1 < depth:
                c = cell_at[l + 1][m]
                k = morton(c)
                for key in np.unique(k):
                    sel = m[k == key]
                    self.children[i].append(len(self.level))
                    self._add(l + 1, cell_at[l + 1][sel[0]], sel, i)
            i += 1
        self.level = np.array(self.level)
        self.cell = np.array(self.cell)
        self.key = morton(self.cell)
        self.parent = np.array(self.parent)
        self.leaf = np.array([len(ch) == 0 for ch in self.children])
        self.count = np.array([len(m) for m in self.members])

    def _add(self, level, cell, members, parent):
        self.level.append(level)
        self.cell.append(np.asarray(cell, np.int64))
        self.members.append(members)
        self.parent.append(parent)
        self.children.append([])

    def touch_matrix(self, a_idx, b_idx):
        """touch[i, j] for boxes a_idx[i], b_idx[j] (closed contact or overlap)."""
        la, lb = self.level[a_idx][:, None], self.level[b_idx][None, :]
        lf = np.maximum(la, lb)
        sa, sb = 1 << (lf - la), 1 << (lf - lb)
        ok = np.ones((len(a_idx), len(b_idx)), bool)
        for ax in range(3):
            a0 = self.cell[a_idx, ax][:, None] * sa
            b0 = self.cell[b_idx, ax][None, :] * sb
            ok &= (a0 <= b0 + sb) & (b0 <= a0 + sa)
        return ok

    def lists(self):
        """U, V, W, X as dicts box -> sorted list of boxes."""
        B = len(self.level)
        allb = np.arange(B)
        T = self.touch_matrix(allb, allb)
        par = self.parent
        leaves = np.nonzero(self.leaf)[0]
        U = {int(b): sorted(int(c) for c in leaves[T[b, leaves]]) for b in leaves}
        V, W, X = {}, {}, {}
        for b in range(B):
            if self.level[b] >= 2:
                same = np.nonzero((self.level == self.level[b]) & ~T[b])[0]
                V[b] = sorted(int(c) for c in same if T[par[b], par[c]])
        for b in leaves:
            fin = np.nonzero((self.level > self.level[b]) & ~T[b])[0]
            W[int(b)] = sorted(int(c) for c in fin if T[b, par[c]])
        for b in range(B):
            X[b] = []
        for b, ws in W.items():
            for c in ws:
                X[c].append(b)
        X = {b: sorted(v) for b, v in X.items()}
        return U, V, W, X

    def ancestors(self, b):
        out = []
        while b >= 0:
            out.append(b)
            b = self.parent[b]
        return out


def partition_check(tree, lists):
    """Number of (target leaf, source leaf) pairs covered other than exactly once."""
    U, V, W, X = lists
    leaves = np.nonzero(tree.leaf)[0]
    li = {int(b): i for i, b in enumerate(leaves)}
    # leaf descendants (as leaf numbers) of every box
    desc = [[] for _ in range(len(tree.level))]
    for b in leaves:
        for a in tree.ancestors(int(b)):
            desc[a].append(li[int(b)])
    desc = [np.array(d, dtype=np.int64) for d in desc]
    C = np.zeros((len(leaves), len(leaves)), np.int64)
    for t, us in U.items():
        C[li[t], [li[s] for s in us]] += 1
    for a, vs in V.items():
        for b in vs:
            C[np.ix_(desc[a], desc[b])] += 1
    for a, xs in X.items():
        for s in xs:
            C[desc[a], li[s]] += 1
    for t, ws in W.items():
        for b in ws:
            C[li[t], desc[b]] += 1
    return int(np.count_nonzero(C != 1))
