"""Pins the oracle's single FMM operators by what the mathematics fixes:
exactness of multipole-to-multipole and local-to-local translation,
convergence of P2M -> M2L -> L2P to the direct sum, and the pair kernel."""
import numpy as np
import pytest


def tri(n):
    return n * (n + 1) // 2


def direct(src, q, tgt):
    d = tgt[:, None, :] - src[None, :, :]
    r = np.sqrt((d * d).sum(-1))
    pot = (q[None, :] / r).sum(1)
    field = (q[None, :, None] * d / r[..., None] ** 3).sum(1)
    return pot, field


@pytest.fixture
def cluster():
    rng = np.random.default_rng(7)
    src = 0.5 + 0.25 * (rng.random((40, 3)) - 0.5)  # inside a box of edge 0.25 at (0.5,0.5,0.5)
    q = rng.uniform(-1, 1, 40)
    return src, q


def test_m2m_is_exact(orc, cluster):
    """M2M of a truncated multipole equals P2M about the new centre."""
    src, q = cluster
    N = 10
    child = np.array([0.5 - 0.0625, 0.5 + 0.0625, 0.5 - 0.0625])
    parent = np.array([0.5, 0.5, 0.5])
    # a child-cell multipole translated to the parent (offset of a child: +-h/4 per axis)
    with orc.exact_geometry():
        Mc = orc.p2m3d(N, src, q, child)
        Mp = orc.m2m3d(N, Mc, child - parent)
        Md = orc.p2m3d(N, src, q, parent)
    assert np.max(np.abs(Mp - Md)) <= 1e-13 * np.max(np.abs(Md))
    # with the reference's radius offset the identity holds to ~1e-11 only
    Mc = orc.p2m3d(N, src, q, child)
    Mp = orc.m2m3d(N, Mc, child - parent)
    Md = orc.p2m3d(N, src, q, parent)
    assert np.max(np.abs(Mp - Md)) <= 1e-10 * np.max(np.abs(Md))


def test_p2m_m2l_l2p_converges_to_direct(orc, cluster):
    src, q = cluster
    centre = np.array([0.5, 0.5, 0.5])
    h = 0.25
    tcentre = centre + np.array([2 * h, -h, 0.0])  # well separated: offset (2,-1,0) boxes
    rng = np.random.default_rng(3)
    tgt = tcentre + h * (rng.random((30, 3)) - 0.5)
    dp, df = direct(src, q, tgt)
    errs = []
    for N in (4, 8, 12, 16):
        M = orc.p2m3d(N, src, q, centre)
        L = orc.m2l3d(N, M, centre - tcentre)
        pot, field = orc.l2p3d(N, L, tcentre, tgt)
        e = np.linalg.norm(pot - dp) / np.linalg.norm(dp)
        ef = np.linalg.norm(field - df) / np.linalg.norm(df)
        errs.append((e, ef))
    assert errs[-1][0] < 1e-6 and errs[-1][1] < 1e-5
    for a, b in zip(errs, errs[1:]):  # geometric decrease with the order
        assert b[0] < 0.5 * a[0] and b[1] < 0.5 * a[1]


def test_l2l_is_exact(orc, cluster):
    """A local expansion is a polynomial: L2L reproduces it exactly."""
    src, q = cluster
    N = 11
    centre = np.array([0.5, 0.5, 0.5])
    far = centre + np.array([0.75, 0.5, -0.5])
    M = orc.p2m3d(N, src, q, centre)
    Lp = orc.m2l3d(N, M, centre - far)        # local about `far` (parent)
    child = far + np.array([0.0625, -0.0625, 0.0625])
    Lc = orc.l2l3d(N, Lp, far - child)
    rng = np.random.default_rng(5)
    pts = child + 0.1 * (rng.random((20, 3)) - 0.5)
    with orc.exact_geometry():
        Lc = orc.l2l3d(N, Lp, far - child)
        p1, f1 = orc.l2p3d(N, Lp, far, pts)
        p2, f2 = orc.l2p3d(N, Lc, child, pts)
    assert np.max(np.abs(p1 - p2)) <= 1e-13 * np.max(np.abs(p1))
    assert np.max(np.abs(f1 - f2)) <= 1e-11 * np.max(np.abs(f1))


def test_m2l_rotation_invariance(orc, cluster):
    """The same physical interaction through two different rotation keys
    (mirror images of the configuration) gives mirror-image potentials."""
    src, q = cluster
    N = 12
    c = np.array([0.5, 0.5, 0.5])
    d = np.array([0.5, 0.25, -0.5])
    M = orc.p2m3d(N, src, q, c)
    L = orc.m2l3d(N, M, c - (c + d))
    mir = np.array([1.0, -1.0, 1.0])  # reflect y about the centre plane
    src_m = c + (src - c) * mir
    M2 = orc.p2m3d(N, src_m, q, c)
    L2 = orc.m2l3d(N, M2, c - (c + d * mir))
    pts = c + d + 0.05 * np.array([[0.3, -0.2, 0.1], [-0.4, 0.1, 0.2]])
    p1, _ = orc.l2p3d(N, L, c + d, pts)
    p2, _ = orc.l2p3d(N, L2, c + d * mir, c + (pts - c) * mir)
    assert np.allclose(p1, p2, rtol=1e-12, atol=0)
