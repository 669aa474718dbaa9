"""GPU: many independent systems in one call (lfmm_fmm3d_batch).

Pinned by:
  per-system identity  every system's outputs equal lfmm_fmm3d on that system
                       alone, bit for bit (same kernels, same tree, same
                       tables in each worker context), for uneven sizes, an
                       empty system, host and device memory, FP64 and mixed
  accuracy             sampled systems vs the oracle's FP64 direct sum
  errors               invalid offsets / worker counts are rejected before
                       any system starts
"""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def case_name(request):
    print("\n[case] %s" % request.node.name, flush=True)


def systems(sizes, seed):
    rng = np.random.default_rng(seed)
    xs, qs = [], []
    for i, n in enumerate(sizes):
        if i % 2:  # mixture of 4 Gaussian blobs
            c = rng.uniform(0.2, 0.8, (4, 3))
            x = c[rng.integers(0, 4, n)] + rng.normal(0, 0.05, (n, 3))
        else:
            x = rng.random((n, 3))
        xs.append(x)
        qs.append(rng.uniform(-1, 1, n))
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    return np.concatenate(xs), np.concatenate(qs), off


@pytest.mark.parametrize("mixed", [False, True])
def test_batch_equals_single_solves(lfmm, request, mixed):
    case_name(request)
    sizes = [30000, 1, 0, 2048, 777, 12000, 5, 40000, 3000, 64, 9000, 20000]
    x, q, off = systems(sizes, 3)
    pot, field, reps = lfmm.fmm3d_batch(x, q, off, terms=9, leaf_max=32, mixed=mixed, workers=4, reports=True)
    for s, n in enumerate(sizes):
        a, b = off[s], off[s + 1]
        if n == 0:
            assert reps[s]["terms"] == 0
            continue
        p1, f1, r1 = lfmm.fmm3d(x[a:b], q[a:b], terms=9, leaf_max=32, mixed=mixed)
        assert np.array_equal(pot[a:b], p1) and np.array_equal(field[a:b], f1), "system %d (n=%d)" % (s, n)
        assert reps[s]["tree_depth"] == r1["tree_depth"] and reps[s]["terms"] == 9


def test_batch_device_matches_host_and_direct(lfmm, orc, request):
    case_name(request)
    import torch
    rng = np.random.default_rng(5)
    sizes = np.rint(np.exp(rng.uniform(np.log(1024), np.log(65536), 40))).astype(np.int64)
    x, q, off = systems(sizes, 7)
    ph, fh = lfmm.fmm3d_batch(x, q, off, terms=9, leaf_max=32, exclude_self=True)
    xd, qd = torch.from_numpy(x).cuda(), torch.from_numpy(q).cuda()
    pd, fd = lfmm.fmm3d_batch(xd, qd, off, terms=9, leaf_max=32, exclude_self=True, workers=16)
    torch.cuda.synchronize()
    assert np.array_equal(pd.cpu().numpy(), ph) and np.array_equal(fd.cpu().numpy(), fh)
    for s in (0, 1, 17, 38):
        a, b = off[s], off[s + 1]
        idx = rng.choice(b - a, 200, replace=False)
        dp, df = orc.direct3d(x[a:b], q[a:b], tgt_index=idx, include_self=False)
        # p = 8: the oracle gives <= 2e-6 potential, <= 3e-4 field on these systems
        assert rel_l2(ph[a:b][idx], dp) < 2e-5 and rel_l2(fh[a:b][idx], df) < 1e-3


def test_batch_rejects_bad_arguments(lfmm, request):
    case_name(request)
    x, q, off = systems([100, 200], 1)
    for bad in ([0, 200, 100], [5, 100, 300]):
        with pytest.raises(ValueError):
            lfmm.fmm3d_batch(x, q, np.array(bad))
    with pytest.raises(lfmm.LfmmError):
        lfmm.fmm3d_batch(x, q, off, workers=17)
    L = lfmm.lib()
    bad = np.array([0, 200, 100], dtype=np.int64)  # decreasing: rejected by the library itself
    assert L.lfmm_fmm3d_batch(lfmm.default_context().handle, 2, bad.ctypes.data, x.ctypes.data, q.ctypes.data, 1e-8,
                              9, lfmm.FMM_AUTO, 32, lfmm.MEM_HOST, 0, 0, None, None, None) == 1


@pytest.mark.parametrize("mixed", [False, True])
def test_forest_equals_single_adaptive_solves(lfmm, request, mixed):
    """adaptive=True: all systems as one forest.  Each system's tree is its
    own adaptive octree in its own units, so every system equals the
    single-system adaptive solve up to the rounding of the mapped positions
    (2a + u carries 4 bits less of u): rel-L2 <= 1e-10 in FP64.  The
    matrix-core M2L scales its sources per level, and a forest level holds the
    boxes of every system, so there the splits round differently: <= 1e-7,
    the size of the mixed path's own error.  Uneven sizes, a one-particle
    system, an empty one and a system smaller than leaf_max."""
    case_name(request)
    sizes = [30000, 1, 0, 2048, 20, 12000, 5, 40000, 3000, 64, 9000, 20000]
    x, q, off = systems(sizes, 13)
    pot, field, reps = lfmm.fmm3d_batch(x, q, off, terms=9, leaf_max=32, mixed=mixed, adaptive=True, reports=True)
    print("forest depth %d leaves %d" % (reps[0]["tree_depth"], reps[0]["num_leaves"]), flush=True)
    for s, n in enumerate(sizes):
        a, b = off[s], off[s + 1]
        if n == 0:
            continue
        p1, f1, _ = lfmm.fmm3d(x[a:b], q[a:b], terms=9, leaf_max=32, adaptive=True, mixed=mixed)
        ep, ef = rel_l2(pot[a:b], p1), rel_l2(field[a:b], f1)
        bar = 1e-7 if mixed else 1e-10
        assert ep <= bar and (n == 1 or ef <= bar), "system %d (n=%d): %.3g %.3g" % (s, n, ep, ef)


def test_forest_512_systems_vs_direct(lfmm, orc, request):
    """The batch workload at full size: 512 systems of 1,024 - 65,536
    particles (log-uniform; half uniform, half four Gaussian blobs), p = 8,
    leaf_max 32, as one forest on the device; 8 sampled systems against the
    oracle's FP64 direct sum (the self term q_i / 1e-4 included on both
    sides, as the reference), and the same systems' single adaptive solves."""
    case_name(request)
    import torch
    rng = np.random.default_rng(21)
    sizes = np.rint(np.exp(rng.uniform(np.log(1024), np.log(65536), 512))).astype(np.int64)
    x, q, off = systems(sizes, 22)
    dev = torch.device("cuda:0")
    xd, qd = torch.from_numpy(x).to(dev), torch.from_numpy(q).to(dev)
    pd, fd, reps = lfmm.fmm3d_batch(xd, qd, off, terms=9, leaf_max=32, adaptive=True, reports=True)
    torch.cuda.synchronize()
    pot, field = pd.cpu().numpy(), fd.cpu().numpy()
    print("512 systems, %d particles: forest depth %d, %d leaves" % (off[-1], reps[0]["tree_depth"],
                                                                      reps[0]["num_leaves"]), flush=True)
    worst_p = worst_f = 0.0
    for s in np.random.default_rng(23).choice(512, 8, replace=False):
        a, b = int(off[s]), int(off[s + 1])
        ix = np.random.default_rng(int(s)).choice(b - a, min(256, b - a), replace=False)
        dp, df = orc.direct3d(x[a:b], q[a:b], tgt_index=ix, include_self=True)
        # compare without the self term, which dominates the potential
        ep = rel_l2(pot[a:b][ix] - q[a:b][ix] * 1e4, dp - q[a:b][ix] * 1e4)
        ef = rel_l2(field[a:b][ix], df)
        worst_p, worst_f = max(worst_p, ep), max(worst_f, ef)
        p1, f1, _ = lfmm.fmm3d(x[a:b], q[a:b], terms=9, leaf_max=32, adaptive=True)
        assert rel_l2(pot[a:b], p1) <= 1e-10 and rel_l2(field[a:b], f1) <= 1e-10, s
    print("worst of 8 sampled systems vs direct: potential %.3e field %.3e" % (worst_p, worst_f), flush=True)
    assert worst_p <= 1e-4 and worst_f <= 1e-3
