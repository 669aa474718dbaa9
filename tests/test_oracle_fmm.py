"""Pins the oracle's whole solves against FP64 direct sums: 3D and 2D, single
set and source/target, plus the reference's self-term convention and its
ReadError statistics."""
import numpy as np
import pytest

from conftest import rel_l2, uniform_cube


@pytest.fixture(scope="module")
def case3d():
    xyz, q = uniform_cube(6000, 21)
    idx = np.arange(0, 6000, 17)
    return xyz, q, idx


def test_fmm3d_converges(orc, case3d):
    xyz, q, idx = case3d
    dp, df = orc.direct3d(xyz, q, tgt_index=idx, include_self=False)
    prev = None
    for N in (3, 6, 9, 12, 15):
        pot, field = orc.fmm3d(xyz, q, N, 4)
        # remove the reference's clamped self term q_i / 1e-4 before comparing
        e = rel_l2(pot[idx] - q[idx] * 1e4, dp)
        ef = rel_l2(field[idx], df)
        if prev is not None:
            assert e < 0.6 * prev[0] and ef < 0.6 * prev[1]
        prev = (e, ef)
    assert prev[0] < 1e-6 and prev[1] < 1e-5


def test_fmm3d_self_term_convention(orc, case3d):
    """The reference includes i == j with r clamped to 1e-4 (simple_math.h:130-133)."""
    xyz, q, idx = case3d
    pot, _ = orc.fmm3d(xyz, q, 9, 4)
    dp_incl, _ = orc.direct3d(xyz, q, tgt_index=idx, include_self=True)
    dp_excl, _ = orc.direct3d(xyz, q, tgt_index=idx, include_self=False)
    assert np.allclose(dp_incl - dp_excl, q[idx] * 1e4, rtol=1e-12, atol=1e-9)
    assert rel_l2(pot[idx], dp_incl) < 1e-5


def test_fmm3d_source_target_mode(orc):
    src, q = uniform_cube(4000, 31)
    tgt, _ = uniform_cube(900, 32)
    pot, field = orc.fmm3d(src, q, 12, 4, tgt_xyz=tgt)
    dp, df = orc.direct3d(src, q, tgt_xyz=tgt, include_self=True)
    assert rel_l2(pot, dp) < 1e-5 and rel_l2(field, df) < 1e-4


def test_fmm2d_converges(orc):
    xy, q = uniform_cube(8000, 41, dim=2)
    idx = np.arange(0, 8000, 13)
    dp, df = orc.direct2d(xy, q, tgt_index=idx, include_self=True)
    errs = []
    for N in (6, 12, 20):
        pot, field = orc.fmm2d(xy, q, N, 6)
        errs.append((rel_l2(pot[idx], dp), rel_l2(field[idx], df)))
    assert errs[-1][0] < 1e-9 and errs[-1][1] < 1e-8
    assert errs[0][0] > errs[1][0] > errs[2][0]


def test_read_error_statistics(orc):
    exact = np.array([1.0, -2.0, 4.0])
    res = np.array([1.1, -2.0, 3.0])
    mx, l2, rel = orc.read_error(exact, res)
    e = np.abs(exact - res)
    assert mx == pytest.approx(e.max())
    assert l2 == pytest.approx(np.sqrt(np.mean((e / np.abs(exact)) ** 2)))
    assert rel == pytest.approx(e.sum() / np.abs(exact).sum())
