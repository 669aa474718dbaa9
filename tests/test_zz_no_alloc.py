"""GPU: the allocation-free direct sum lfmm_direct3d_noalloc against
lfmm_direct3d (same kernels, packed inputs in the fixed per-GPU scratch
instead of the context's workspace), called through the C ABI.

Two calls in a row, the second with a smaller input (single set, then with
targets), must equal the existing entry point's results bit for bit; an
input beyond the fixed 1,048,576 records is refused with
LFMM_ERR_TOO_LARGE before any device work.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_direct3d_noalloc_matches_direct3d(lfmm):
    import torch
    L = lfmm.lib()
    ctx = lfmm.Context(0)
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(41)
    for n, nt in ((300000, 0), (20000, 7000)):
        print("\n[case] noalloc n=%d targets=%d" % (n, nt), flush=True)
        x = torch.from_numpy(rng.random((n, 3))).to(dev)
        q = torch.from_numpy(rng.uniform(-1, 1, n)).to(dev)
        t = torch.from_numpy(rng.random((nt, 3))).to(dev) if nt else None
        m = nt if nt else n
        torch.cuda.synchronize()
        ref_p, ref_f = lfmm.direct3d(x, q, targets=t, exclude_self=(nt == 0), ctx=ctx)
        torch.cuda.synchronize()
        pot = torch.full((m,), float("nan"), dtype=torch.float64, device=dev)
        field = torch.full((m, 3), float("nan"), dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        src = lfmm.Points3d(n, x.data_ptr(), 3, q.data_ptr(), 1)
        tg = ctypes.byref(lfmm.Points3d(nt, t.data_ptr(), 3, None, 1)) if nt else None
        rc = L.lfmm_direct3d_noalloc(ctx.handle, ctypes.byref(src), tg, lfmm.FLAG_EXCLUDE_SELF if nt == 0 else 0,
                                     ctypes.c_void_p(pot.data_ptr()), ctypes.c_void_p(field.data_ptr()))
        assert rc == 0
        torch.cuda.synchronize()
        assert torch.equal(pot, ref_p) and torch.equal(field, ref_f)
    big = 1048577
    xb = torch.zeros((big, 3), dtype=torch.float64, device=dev)
    qb = torch.zeros(big, dtype=torch.float64, device=dev)
    ob = torch.zeros(big, dtype=torch.float64, device=dev)
    src = lfmm.Points3d(big, xb.data_ptr(), 3, qb.data_ptr(), 1)
    assert L.lfmm_direct3d_noalloc(ctx.handle, ctypes.byref(src), None, 0, ctypes.c_void_p(ob.data_ptr()),
                                   None) == 9
    ctx.close()
