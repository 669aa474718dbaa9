"""Pins the oracle's constant tables and special functions against closed forms
and independent library routines (no reference golden vectors exist for this
path: the reference repository holds no tests or fixtures)."""
import math

import numpy as np
import pytest
import scipy.linalg as sl
from scipy.special import lpmv

N18 = 18


def tri(n):
    return n * (n + 1) // 2


def test_knm_closed_form(orc):
    K = orc.knm(N18)
    for n in range(N18):
        for m in range(n + 1):
            ref = math.sqrt(math.factorial(n - m) / math.factorial(n + m))
            assert abs(K[tri(n) + m] - ref) <= 4e-16 * ref


def test_anm_closed_form(orc):
    A = orc.anm(N18)
    for n in range(2 * N18):
        for m in range(n + 1):
            ref = (-1) ** n / math.sqrt(math.factorial(n - m) * math.factorial(n + m))
            assert abs(A[tri(n) + m] - ref) <= 6e-16 * abs(ref)


def test_m2lcoef_closed_form(orc):
    for j in range(0, N18, 3):
        for k in range(j + 1):
            for n in range(k, N18, 2):
                a = lambda n_, m_: (-1) ** n_ / math.sqrt(math.factorial(n_ - m_) * math.factorial(n_ + m_))
                ref = a(n, k) * a(j, k) * (-1) ** (n + k) / a(j + n, 0)
                assert orc.m2lcoef(j, k, n) == pytest.approx(ref, rel=2e-15)


@pytest.mark.parametrize("theta", [0.0, 0.1, 0.7, 1.3, math.pi / 2, 2.2, 3.0])
def test_pnm_vs_scipy(orc, theta):
    P = orc.pnm(N18, theta)
    K = orc.knm(N18)
    x = math.cos(theta)
    for n in range(N18):
        for m in range(n + 1):
            ref = lpmv(m, n, x)  # includes the Condon-Shortley phase
            # compare the normalised K_nm P_n^m (|.| <= 1), the quantity the FMM uses
            assert abs(K[tri(n) + m] * (P[tri(n) + m] - ref)) <= 1e-13
    # special value on the diagonal: P_n^n = (-1)^n (2n-1)!! sin^n
    for n in range(N18):
        df = 1.0
        for k in range(1, 2 * n, 2):
            df *= k
        assert P[tri(n) + n] == pytest.approx((-1) ** n * df * math.sin(theta) ** n, rel=1e-13, abs=1e-300)


def std_wigner_d(j, beta):
    """Standard small Wigner matrix d^j_{m'm}(beta) = <j m'|exp(-i beta J_y)|j m> via expm."""
    D = 2 * j + 1
    A = np.zeros((D, D))
    for a in range(D - 1):
        m = a - j
        v = math.sqrt(j * (j + 1) - m * (m + 1))
        A[a + 1, a] = -0.5 * beta * v
        A[a, a + 1] = 0.5 * beta * v
    return sl.expm(A)


def ref_block(D, n):
    base = n * (n + 1) * (4 * n - 1) // 6
    return np.array([[D[base + m * (2 * n + 1) + k + n] for k in range(-n, n + 1)] for m in range(n + 1)])


@pytest.mark.parametrize("theta", [0.25, 0.9553166181245093, 1.2309594173407747, 2.0, 2.6, -0.5, -2.9])
def test_wignerd_matches_standard_convention(orc, theta):
    """Reference table D_n(m,k) = eps(k) d^n_{k,m}(theta), eps(k) = (-1)^k for k < 0."""
    N = 12
    D = orc.wignerd(N, theta)
    for n in range(N):
        B = ref_block(D, n)
        d = std_wigner_d(n, theta)
        for m in range(n + 1):
            for k in range(-n, n + 1):
                eps = (-1) ** k if k < 0 else 1
                assert abs(B[m, k + n] - eps * d[k + n, m + n]) < 5e-13


def test_wignerd_identity_and_degree_one(orc):
    N = 8
    D = orc.wignerd(N, 0.0)
    for n in range(N):
        B = ref_block(D, n)
        for m in range(n + 1):
            for k in range(-n, n + 1):
                assert B[m, k + n] == (1.0 if k == m else 0.0)
    th = 0.8
    B = ref_block(orc.wignerd(N, th), 1)
    # degree 1 closed forms: d^1_{00} = cos, d^1_{1,1} = (1+cos)/2, d^1_{-1,1} = (1-cos)/2
    assert B[0, 1] == pytest.approx(math.cos(th), abs=1e-14)
    assert B[1, 2] == pytest.approx((1 + math.cos(th)) / 2, abs=1e-14)
    assert B[1, 0] == pytest.approx(-(1 - math.cos(th)) / 2, abs=1e-14)  # eps(-1) = -1


def test_wignerd_full_rows_orthonormal(orc):
    # rows m >= 0 of d^n are orthonormal over k in [-n, n] (undo the eps sign)
    N = 10
    for th in (0.4, 1.7, 2.9):
        D = orc.wignerd(N, th)
        for n in range(N):
            B = ref_block(D, n).copy()
            for k in range(-n, 0):
                B[:, k + n] *= (-1) ** k
            G = B @ B.T
            assert np.allclose(G, np.eye(n + 1), atol=1e-12)


def test_direction_tables_cover_all_offsets(orc):
    """Every M2L offset, and the +-theta keys of M2M/L2L child offsets, resolve
    in the hash-keyed table (special_functions.cpp:363-381)."""
    N = 6
    for i in range(-3, 4):
        for j in range(-3, 4):
            for k in range(-3, 4):
                if max(abs(i), abs(j), abs(k)) <= 1:
                    continue
                th = orc.dec_to_sph([i, j, k])[1]
                assert orc.dmatrix_lookup(th, N) is not None
                assert orc.dmatrix_lookup(-th, N) is not None
    for sx in (-1, 1):
        for sy in (-1, 1):
            for sz in (-1, 1):
                th = orc.dec_to_sph([0.25 * sx, 0.25 * sy, 0.25 * sz])[1]
                assert orc.dmatrix_lookup(th, N) is not None


def test_dec_to_sph_special_cases(orc):
    r = orc.dec_to_sph([0.0, 0.0, 2.0])
    assert r[2] == 0.0 and r[0] == 2.0 + 1e-12
    r = orc.dec_to_sph([0.0, -1.0, 0.0])
    assert r[2] == pytest.approx(-math.pi / 2)
    r = orc.dec_to_sph([-1.0, 1.0, 0.0])
    assert r[2] == pytest.approx(3 * math.pi / 4)
    assert r[1] == pytest.approx(math.pi / 2)
