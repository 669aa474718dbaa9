"""Dense real M2L operators and their split-fp16 packing (host side of the
matrix-core M2L), pinned against the oracle's rotation-based M2L.

With box edge h = 1 the per-level normalisation (M/h^n, L h^(n+1)) is the
identity, so K_d applied to the real coordinates of a multipole must equal
the oracle's M2L (reference src/cpu/multipole3d.cpp:152-190) for the
translation vector d = source centre - target centre.
"""
import numpy as np
import pytest


def tri(n):
    return n * (n + 1) // 2


def real_index(n, m, part):
    return n * n + (0 if m == 0 else 2 * m - 1 + part)


def to_real(N, M):
    x = np.zeros(N * N)
    for n in range(N):
        for m in range(n + 1):
            x[real_index(n, m, 0)] = M[tri(n) + m].real
            if m:
                x[real_index(n, m, 1)] = M[tri(n) + m].imag
    return x


def random_multipole(N, rng):
    M = rng.standard_normal(tri(N)) + 1j * rng.standard_normal(tri(N))
    for n in range(N):
        M[tri(n)] = M[tri(n)].real  # m = 0 coefficients are real for real charges
    return M


@pytest.mark.parametrize("N", [1, 3, 6, 11, 14])
def test_dense_m2l_equals_oracle_m2l(lfmm, orc, N):
    rng = np.random.default_rng(N)
    dv = lfmm.host_op_tables(N, "dir_vec").reshape(-1, 4)
    dirs = [0, 7, 100, 157, 200, 315] + list(rng.choice(316, 6, replace=False))
    for d in dirs:
        K = lfmm.host_m2l_dense(N, int(d))
        M = random_multipole(N, rng)
        with orc.exact_geometry():
            L = orc.m2l3d(N, M, dv[d, :3].astype(float))
        x = to_real(N, M)
        y = K @ x
        ref = to_real(N, L)
        # bound relative to the size of the summed terms; the oracle's Wigner
        # recurrence is accurate to ~2.5e-12 for the |d|^2 = 5 directions
        scale = np.max(np.abs(K) @ np.abs(x))
        assert np.max(np.abs(y - ref)) <= 5e-12 * scale, (N, d)
        # the imaginary part of every m = 0 output vanishes
        assert np.max(np.abs([L[tri(n)].imag for n in range(N)])) <= 5e-12 * scale


def test_dense_m2l_reflection_symmetry(lfmm):
    """Mirroring z maps M_nm to (-1)^(n+m) M_nm on both sides: K_{-d_z} = S K_d S."""
    N = 9
    dv = lfmm.host_op_tables(N, "dir_vec").reshape(-1, 4)[:316, :3]
    s = np.ones(N * N)
    for n in range(N):
        for m in range(n + 1):
            for p in ((0,) if m == 0 else (0, 1)):
                s[real_index(n, m, p)] = (-1) ** (n + m)
    idx = {tuple(v): i for i, v in enumerate(dv)}
    for d in (3, 50, 120, 250):
        v = dv[d]
        e = idx[(v[0], v[1], -v[2])]
        K, Km = lfmm.host_m2l_dense(N, d), lfmm.host_m2l_dense(N, e)
        assert np.allclose(Km, s[:, None] * K * s[None, :], rtol=0, atol=1e-14 * np.abs(K).max())


@pytest.mark.parametrize("N", [4, 11])
def test_m2l16_packing_reconstructs_operators(lfmm, N):
    NR = N * N
    RT, KC = (NR + 31) // 32, (NR + 15) // 16
    pk, kinv, kpre = lfmm.host_m2l16_pack(N)
    assert pk.size == 316 * RT * KC * 2 * 64 * 8
    assert kinv > 0 and np.log2(kinv) == np.round(np.log2(kinv))  # a power of two
    h = pk.view(np.float16).astype(np.float64).reshape(316, RT, KC, 2, 64, 8)
    lane = np.arange(64)
    rows = (np.arange(RT)[:, None, None, None] * 32 + (lane % 32)[None, None, :, None])  # (RT,1,64,1)
    ks = (np.arange(KC)[None, :, None, None] * 16 + 8 * (lane // 32)[None, None, :, None]
          + np.arange(8)[None, None, None, :])                                            # (1,KC,64,8)
    rows = np.broadcast_to(rows, (RT, KC, 64, 8))
    ks = np.broadcast_to(ks, (RT, KC, 64, 8))
    for d in (0, 91, 315):
        K = lfmm.host_m2l_dense(N, d)
        Kp = np.zeros((RT * 32, KC * 16))
        Kp[:NR, :NR] = K
        want = Kp[rows, ks]
        got = (h[d, :, :, 0] + h[d, :, :, 1]) * kinv
        tol = 2.0 ** -21 * np.abs(want) + 2.0 ** -24 * kinv
        assert np.all(np.abs(got - want) <= tol), d
        assert np.all(np.abs(h[d, :, :, 0]) <= 2.0 ** 14)
        # hi alone is the fp16 rounding of the scaled entry
        assert np.allclose(h[d, :, :, 0], (want / kinv).astype(np.float32).astype(np.float16), rtol=0, atol=0)


@pytest.mark.parametrize("N", [5, 11])
def test_m2l16_block_prefixes_bound_dropped_terms(lfmm, N):
    """Blocks at or beyond k1 are below 2^-22 of the largest entry, blocks at
    or beyond k3 below 2^-11; k3 <= k1 <= KC, and the prefixes are tight."""
    NR = N * N
    RT, KC = (NR + 31) // 32, (NR + 15) // 16
    _, _, kpre = lfmm.host_m2l16_pack(N)
    Ks = [lfmm.host_m2l_dense(N, d) for d in range(316)]
    kmax = max(np.abs(K).max() for K in Ks)
    for d in range(0, 316, 7):
        P = np.zeros((RT * 32, KC * 16))
        P[:NR, :NR] = np.abs(Ks[d])
        B = P.reshape(RT, 32, KC, 16).max(axis=(1, 3))
        for rt in range(RT):
            k3, k1 = kpre[d, rt]
            assert 0 <= k3 <= k1 <= KC
            assert np.all(B[rt, k1:] < kmax * 2.0 ** -22)
            assert np.all(B[rt, k3:] < kmax * 2.0 ** -11)
            if k1 > 0:
                assert B[rt, k1 - 1] >= kmax * 2.0 ** -22 or k1 == k3
            if k3 > 0:
                assert B[rt, k3 - 1] >= kmax * 2.0 ** -11


@pytest.mark.parametrize("N", [1, 4, 9, 13])
def test_dense_m2m_equals_oracle_m2m(lfmm, orc, N):
    """The matrix-core M2M operator of child class c (bits x y z) maps the
    child's normalised multipole (box edge 1/2: M_n / 2^-n) to the parent's
    (edge 1), as the oracle's rotation-based M2M (reference M2M,
    src/cpu/multipole3d.cpp:115-150) translates it by the child centre minus
    the parent centre, (+-1/4, +-1/4, +-1/4)."""
    rng = np.random.default_rng(100 + N)
    for c in range(8):
        K = lfmm.host_m2m_dense(N, c)
        M = random_multipole(N, rng)
        off = np.array([(c >> 2) & 1, (c >> 1) & 1, c & 1], dtype=float) / 2.0 - 0.25
        with orc.exact_geometry():
            Mp = orc.m2m3d(N, M, off)
        Mn = np.array([M[tri(n) + m] / 0.5 ** n for n in range(N) for m in range(n + 1)])
        x = to_real(N, Mn)
        y = K @ x
        ref = to_real(N, Mp)
        scale = np.max(np.abs(K) @ np.abs(x))
        assert np.max(np.abs(y - ref)) <= 5e-12 * scale, (N, c)
