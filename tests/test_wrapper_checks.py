"""Argument checks of the Python binding that run before any library call
(no GPU needed): mismatched shapes and misplaced outputs raise."""
import numpy as np
import pytest


def test_charges_must_match_positions(lfmm):
    x = np.zeros((10, 3))
    with pytest.raises(ValueError, match="charges"):
        lfmm.fmm3d(x, np.zeros(9), terms=5, tree_depth=3)
    with pytest.raises(ValueError, match="charges"):
        lfmm.direct3d(x, np.zeros(11))
    with pytest.raises(ValueError, match="charges"):
        lfmm.fmm2d(np.zeros((10, 2)), np.zeros(12), terms=5, tree_depth=3)


def test_positions_must_have_dim_columns(lfmm):
    with pytest.raises(ValueError, match="positions"):
        lfmm.fmm3d(np.zeros((10, 4)), np.zeros(10), terms=5, tree_depth=3)
    with pytest.raises(ValueError, match="targets"):
        lfmm.fmm3d(np.zeros((10, 3)), np.zeros(10), terms=5, tree_depth=3, targets=np.zeros((4, 2)))


def test_vector_charges_need_three_components(lfmm):
    with pytest.raises(ValueError, match="charge vectors"):
        lfmm.fmm3d_vec(np.zeros((10, 3)), np.zeros((10, 2)), terms=5, tree_depth=3)


def test_host_calls_reject_out(lfmm):
    x = np.zeros((10, 3))
    with pytest.raises(ValueError, match="out="):
        lfmm.fmm3d(x, np.zeros(10), terms=5, tree_depth=3, out=(np.zeros(10), np.zeros((10, 3))))
    with pytest.raises(ValueError, match="out="):
        lfmm.fmm3d_batch(x, np.zeros(10), [0, 10], terms=5, out=(np.zeros(10), np.zeros((10, 3))))


def test_batch_offsets_are_checked(lfmm):
    x = np.zeros((10, 3))
    with pytest.raises(ValueError, match="offsets"):
        lfmm.fmm3d_batch(x, np.zeros(10), [0, 4, 9], terms=5)
    with pytest.raises(ValueError, match="offsets"):
        lfmm.fmm3d_batch(x, np.zeros(10), [0, 6, 4, 10], terms=5)
    with pytest.raises(ValueError, match="charges"):
        lfmm.fmm3d_batch(x, np.zeros(8), [0, 10], terms=5)

