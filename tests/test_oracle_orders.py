"""CPU: the oracle's per-box truncation orders (fmm3d_orders in
oracle/fmm_oracle.c, the checker of the GPU per-box orders).

Rule restated there: a leaf takes the largest order requested by its
particles, a box with children the largest of its children, and the M2L and
L2L into a box of level >= 4 run at the box's own order; levels 2-3,
P2M, M2M and L2P use the largest order.  Pinned by what the rule fixes
exactly:
- one order everywhere is the fixed-order solve, bit for bit;
- with order A on one half of the domain (x < 1/2) and B > A on the other,
  a particle whose level-3 ancestor lies in the B half meets the same
  operations as in the fixed-B solve: equal bit for bit.  Taking the M2L
  order from the source box, or the smaller of two orders, would truncate
  the sources of the A half there and break this;
- the A half is as accurate as the fixed-A solve there (within 2x), which a
  rule that keeps B there (or drops below A) would not be.
"""
import numpy as np

from conftest import uniform_cube


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_one_order_everywhere_is_the_fixed_solve(orc):
    x, q = uniform_cube(6000, 5)
    for N in (5, 9):
        with orc.exact_geometry(), orc.threads(1):
            p, f = orc.fmm3d_orders(x, q, np.full(len(q), N, dtype=np.int32), 5)
            p0, f0 = orc.fmm3d(x, q, N, 5)[:2]
        assert np.array_equal(p, p0) and np.array_equal(f, f0)


def test_two_orders_by_halves_of_the_domain(orc):
    x, q = uniform_cube(30000, 6)
    A, B, depth = 4, 11, 6
    pt = np.where(x[:, 0] < 0.5, A, B).astype(np.int32)
    with orc.exact_geometry(), orc.threads(1):
        p, f = orc.fmm3d_orders(x, q, pt, depth)
        pa, _ = orc.fmm3d(x, q, A, depth)[:2]
        pb, fb = orc.fmm3d(x, q, B, depth)[:2]
    # the tree's level-3 cells (edge scale / 8): a particle's level-3 ancestor
    # lies in the B half when its cell index along x is >= 4
    lo = x.min(0) - 1e-7
    scale = (x.max(0) - x.min(0)).max() + 1e-6
    cx3 = np.floor(8 * (x[:, 0] - lo[0]) / scale).astype(int)
    hi_half = cx3 >= 4
    assert hi_half.sum() > 10000
    assert np.array_equal(p[hi_half], pb[hi_half]) and np.array_equal(f[hi_half], fb[hi_half])
    # the A half: accuracy of the fixed-A solve
    idx = np.flatnonzero(~hi_half)[::7]
    dp, _ = orc.direct3d(x, q, tgt_index=idx, include_self=True)
    ea, e = rel(pa[idx], dp), rel(p[idx], dp)
    assert 0.5 * ea <= e <= 2.0 * ea, (e, ea)
    assert rel(pb[idx], dp) < 0.1 * e
