"""Operand formats of the matrix-core M2L, compared on one target box.

The matrix-core M2L (m2l16_kernels.h, k_m2l16r) multiplies dense per-direction
operators K_d by the real coordinates of the source multipoles in split FP16:
both operands scaled by powers of two and split as hi = fp16(x), lo =
fp16(x - hi); per 16-wide K chunk and 32-row tile the products hi*hi (below the
tile's bound k1) and hi*lo + lo*hi (below its bound k3), each MFMA result
rounded to FP32; row tile 0 restarts every direction and is summed into a
second FP32 accumulator.  This test replays that sequence in numpy for the 189
directions of child class 0 (operators from the library's host builder,
multipoles of 64 random charges per source box from the oracle's P2M), and the
same sequence with split BF16 operands and with plain FP32 operands (every
product, no skipping: the best case of each), and evaluates the potential of
each local expansion at 400 points of the target box against the FP64 product.

At 13 terms split FP16 stays below 3e-7; split BF16, at the same matrix-core
rate, errs more than twice as much (its 8-bit halves carry 16 bits); FP32
operands err less but run on the 157 TF FP32 matrix path, 16x below the fp16
rate for a third of the products.  These are the figures behind the choice of
split FP16; the GPU tests check the kernel itself against FP64.
"""
import numpy as np
import pytest

N = 13
NR = N * N


def tri(n):
    return n * (n + 1) // 2


def real_index(n, m, part):
    return n * n + (0 if m == 0 else 2 * m - 1 + part)


def to_real(M):
    x = np.zeros(NR)
    for n in range(N):
        for m in range(n + 1):
            x[real_index(n, m, 0)] = M[tri(n) + m].real
            if m:
                x[real_index(n, m, 1)] = M[tri(n) + m].imag
    return x


def from_real(x):
    L = np.zeros(tri(N), dtype=complex)
    for n in range(N):
        for m in range(n + 1):
            L[tri(n) + m] = x[real_index(n, m, 0)] + (1j * x[real_index(n, m, 1)] if m else 0)
    return L


def r32(v):
    return v.astype(np.float32).astype(np.float64)


def f16(a):
    return np.asarray(a).astype(np.float16).astype(np.float64)


def bf16(a):
    """Round to nearest even bfloat16 (the top 16 bits of an FP32)."""
    f = np.asarray(a, dtype=np.float32).view(np.uint32)
    r = ((f + 0x7FFF + ((f >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return r.view(np.float32).astype(np.float64)


@pytest.fixture(scope="module")
def box(lfmm, orc):
    dv = lfmm.host_op_tables(N, "dir_vec").reshape(-1, 4)
    dirs = lfmm.host_op_tables(N, "class_dirs").reshape(8, -1)[0]
    Ks = {int(d): lfmm.host_m2l_dense(N, int(d)) for d in dirs}
    # the kernel's operator scale (a power of two set by the largest entry
    # over all directions) and its block bounds
    _, kinv, kpre = lfmm.host_m2l16_pack(N)
    sk = 1.0 / kinv
    rng = np.random.default_rng(1)
    X = []
    for d in dirs:
        xyz = rng.random((64, 3)) - 0.5 + dv[d, :3]
        q = rng.uniform(-1.0, 1.0, 64)
        with orc.exact_geometry():
            X.append(to_real(orc.p2m3d(N, xyz, q, dv[d, :3].astype(float))))
    X = np.array(X)
    sx = 2.0 ** (14 - np.frexp(np.float32(np.abs(X).max()))[1])
    Lex = sum(Ks[int(d)] @ X[i] for i, d in enumerate(dirs))
    pts = rng.random((400, 3)) - 0.5
    with orc.exact_geometry():
        phi = orc.l2p3d(N, from_real(Lex), np.zeros(3), pts)[0]
    return dict(dirs=dirs, Ks=Ks, sk=sk, sx=sx, X=X, kpre=np.asarray(kpre), pts=pts, phi=phi)


def potential_error(orc, b, L):
    with orc.exact_geometry():
        ph = orc.l2p3d(N, from_real(L), np.zeros(3), b["pts"])[0]
    return np.linalg.norm(ph - b["phi"]) / np.linalg.norm(b["phi"])


def accumulate(b, kind):
    """The kernel's accumulation order: segments of 3 K chunks (chunks s,
    s + NSEG, s + 2 NSEG), row tile 0 restarted per direction."""
    RT, KC = (NR + 31) // 32, (NR + 15) // 16
    nseg = (KC + 2) // 3
    order = [k for k in (s + nseg * j for s in range(nseg) for j in range(3)) if k < KC]
    C = np.zeros(NR)
    T = np.zeros(NR)
    for i, d in enumerate(b["dirs"]):
        K = b["Ks"][int(d)] * b["sk"]
        x = b["X"][i] * b["sx"]
        if kind == "fp32":
            C = r32(C + r32(K) @ r32(x))
        else:
            rnd = f16 if kind == "fp16x2" else bf16
            Kh = rnd(K)
            Kl = rnd(K - Kh)
            xh = rnd(x)
            xl = rnd(x - xh)
            for kc in order:
                sl = slice(16 * kc, min(NR, 16 * kc + 16))
                for rt in range(RT):
                    rows = slice(32 * rt, min(NR, 32 * rt + 32))
                    k3, k1 = (b["kpre"][d, rt] if kind == "fp16x2" else (KC, KC))
                    if kc < k1:
                        C[rows] = r32(C[rows] + Kh[rows, sl] @ xh[sl])
                    if kc < k3:
                        C[rows] = r32(C[rows] + Kh[rows, sl] @ xl[sl])
                        C[rows] = r32(C[rows] + Kl[rows, sl] @ xh[sl])
        T[:32] = r32(T[:32] + C[:32])
        C[:32] = 0.0
    return (T + C) / (b["sk"] * b["sx"])


def test_split_fp16_meets_its_bound_and_beats_split_bf16(orc, box):
    e16 = potential_error(orc, box, accumulate(box, "fp16x2"))
    ebf = potential_error(orc, box, accumulate(box, "bf16x2"))
    e32 = potential_error(orc, box, accumulate(box, "fp32"))
    print("13 terms, potential rel error: split fp16 %.2e, split bf16 %.2e, fp32 %.2e" % (e16, ebf, e32))
    assert e16 < 3e-7
    assert ebf > 2.0 * e16
    assert e32 < e16
