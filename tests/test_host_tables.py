"""Pins the product's host-built FP64 operator tables (tables.cpp) against the
oracle's reference-convention Wigner matrices and coefficient tables, and
against an independent scipy construction.  No GPU needed."""
import math

import numpy as np
import pytest
import scipy.linalg as sl

from test_oracle_tables import ref_block, std_wigner_d


def tri(n):
    return n * (n + 1) // 2


def rot_base(n):
    return n * (n + 1) * (2 * n + 1) // 6


def tr_base(N, k):
    return rot_base(N) - rot_base(N - k)


@pytest.mark.parametrize("j,beta", [(0, 0.3), (1, 1.1), (5, 0.9553), (11, 2.4), (17, 3.0), (17, 1.5707963267948966)])
def test_host_wigner_matches_expm(lfmm, j, beta):
    assert np.allclose(lfmm.host_wigner_d(j, beta), std_wigner_d(j, beta), atol=5e-13, rtol=0)


@pytest.mark.parametrize("j", [1, 6, 11, 17])
def test_host_wigner_group_properties(lfmm, j):
    """d^j is orthogonal, d(a) d(b) = d(a + b), d(-b) = d(b)^T, and degree 1
    has the closed form d^1_{00} = cos b, d^1_{11} = (1 + cos b)/2."""
    a, b = 0.7, 1.9
    da, db, dab = lfmm.host_wigner_d(j, a), lfmm.host_wigner_d(j, b), lfmm.host_wigner_d(j, a + b)
    I = np.eye(2 * j + 1)
    assert np.abs(da @ da.T - I).max() < 1e-15
    assert np.abs(da @ db - dab).max() < 1e-14
    assert np.abs(lfmm.host_wigner_d(j, -b) - db.T).max() < 1e-15
    if j == 1:
        assert da[1, 1] == pytest.approx(math.cos(a), abs=3e-16)
        assert da[2, 2] == pytest.approx((1 + math.cos(a)) / 2, abs=3e-16)


@pytest.mark.parametrize("N", [9, 11, 18])
def test_rotation_pairs_match_reference_tables(lfmm, orc, N):
    rot = lfmm.host_op_tables(N, "rot").reshape(-1, rot_base(N), 2)
    dir_theta = lfmm.host_op_tables(N, "dir_theta")
    dir_vec = lfmm.host_op_tables(N, "dir_vec").reshape(-1, 4)
    seen = set()
    for d in range(len(dir_theta)):
        t = int(dir_theta[d])
        if t in seen:
            continue
        seen.add(t)
        x, y, z = dir_vec[d, :3]
        theta = math.acos(z / math.sqrt(x * x + y * y + z * z))
        D = orc.wignerd(N, theta)
        for n in range(N):
            B = ref_block(D, n)
            d_std = std_wigner_d(n, theta)
            # the reference's recurrence loses accuracy with the degree
            # (~1e-12 at n = 17); the independent expm table pins tightly
            tol_ref = 1e-13 if n < 12 else 5e-12
            for m in range(n + 1):
                for k in range(n + 1):
                    P, Q = rot[t, rot_base(n) + m * (n + 1) + k]
                    if k == 0:
                        eP = eQ = B[m, n]
                        sP = sQ = d_std[n, m + n]
                    else:
                        eP = B[m, k + n] + B[m, -k + n]
                        eQ = B[m, k + n] - B[m, -k + n]
                        a, b = d_std[k + n, m + n], (-1) ** k * d_std[-k + n, m + n]
                        sP, sQ = a + b, a - b
                    assert abs(P - eP) < tol_ref and abs(Q - eQ) < tol_ref, (d, n, m, k)
                    assert abs(P - sP) < 5e-13 and abs(Q - sQ) < 5e-13, (d, n, m, k)
    assert len(seen) == rot.shape[0]


def test_translation_tables(lfmm, orc):
    N = 11
    tr = lfmm.host_op_tables(N, "trans").reshape(-1, rot_base(N))
    dir_trans = lfmm.host_op_tables(N, "dir_trans")
    dir_vec = lfmm.host_op_tables(N, "dir_vec").reshape(-1, 4)
    A = orc.anm(N)
    a = lambda n, m: A[tri(n) + m]
    for d in (0, 57, 123, 200, 315):
        D = math.sqrt(float((dir_vec[d, :3] ** 2).sum()))
        t = tr[dir_trans[d]]
        for k in range(N):
            for j in range(k, N):
                for i in range(k, N):
                    exp = orc.m2lcoef(j, k, i) * D ** (-(j + i + 1))
                    assert t[tr_base(N, k) + (j - k) * (N - k) + (i - k)] == pytest.approx(exp, rel=1e-13)
    m2m, l2l = tr[16], tr[17]
    rr = math.sqrt(3) / 4
    for k in range(N):
        for j in range(k, N):
            for i in range(k, N):
                o = tr_base(N, k) + (j - k) * (N - k) + (i - k)
                em = a(j - i, 0) * a(i, k) / a(j, k) * rr ** (j - i) * 0.5 ** i if i <= j else 0.0
                el = a(i - j, 0) * a(j, k) * (-1) ** (i + j) / a(i, k) * rr ** (i - j) * 0.5 ** (j + 1) if i >= j else 0.0
                assert m2m[o] == pytest.approx(em, rel=1e-13, abs=1e-300)
                assert l2l[o] == pytest.approx(el, rel=1e-13, abs=1e-300)


def test_direction_lists(lfmm):
    N = 9
    dv = lfmm.host_op_tables(N, "dir_vec").reshape(-1, 4)
    cd = lfmm.host_op_tables(N, "class_dirs").reshape(8, 189)
    ephi = lfmm.host_op_tables(N, "ephi").reshape(-1, N, 2)
    assert len(dv) == 324
    union = set()
    for c in range(8):
        cx, cy, cz = (c >> 2) & 1, (c >> 1) & 1, c & 1
        ds = set(cd[c].tolist())
        assert len(ds) == 189 and -1 not in ds
        for d in ds:
            x, y, z = dv[d, :3]
            assert max(abs(x), abs(y), abs(z)) >= 2
            assert -2 - cx <= x <= 3 - cx and -2 - cy <= y <= 3 - cy and -2 - cz <= z <= 3 - cz
        union |= ds
    assert union == set(range(316))
    for d in (3, 100, 316, 323):
        x, y, _ = dv[d, :3]
        phi = 0.0 if x == 0 and y == 0 else math.atan2(y, x)
        for k in range(N):
            assert ephi[d, k, 0] == pytest.approx(math.cos(k * phi), abs=5e-15)
            assert ephi[d, k, 1] == pytest.approx(math.sin(k * phi), abs=5e-15)


def test_knm(lfmm, orc):
    K = lfmm.host_op_tables(17, "knm")
    assert np.allclose(K[:tri(17)], orc.knm(17), rtol=4e-16, atol=0)
