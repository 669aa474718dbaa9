"""GPU: device calls of the Python binding run on torch's current stream, so
outputs are consumed and inputs produced by torch work with no device-wide
synchronisation in between.

Pinned by: results read back on torch's stream right after the call (no
torch.cuda.synchronize) equal the host-memory path; inputs written by a
torch kernel queued just before the call are the ones used; a side stream
bound with torch.cuda.stream() is followed as well.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def case_name(request):
    print("\n[case] %s" % request.node.name, flush=True)


def test_outputs_follow_torch_stream(lfmm, request):
    case_name(request)
    import torch
    rng = np.random.default_rng(21)
    x = rng.random((200000, 3))
    q = rng.uniform(-1, 1, 200000)
    ph, fh, _ = lfmm.fmm3d(x, q, terms=9, tree_depth=5)
    dev = torch.device("cuda:0")
    xd = torch.from_numpy(x).to(dev)
    # the charges come out of a torch kernel still queued when the solve starts
    qd = torch.from_numpy(q).to(dev) * 2.0
    qd.mul_(0.5)
    pd, fd, _ = lfmm.fmm3d(xd, qd, terms=9, tree_depth=5)
    got_p = pd.cpu().numpy()  # torch copy on its current stream, no explicit sync
    got_f = fd.cpu().numpy()
    assert np.array_equal(got_p, ph) and np.array_equal(got_f, fh)


def test_side_stream_is_followed(lfmm, request):
    case_name(request)
    import torch
    rng = np.random.default_rng(22)
    x = rng.random((100000, 3))
    q = rng.uniform(-1, 1, 100000)
    ph, fh = lfmm.direct3d(x[:2000], q[:2000])
    pf, ff, _ = lfmm.fmm3d(x, q, terms=8, tree_depth=5)
    dev = torch.device("cuda:0")
    s = torch.cuda.Stream(device=dev)
    with torch.cuda.stream(s):
        xd = torch.from_numpy(x).to(dev)
        qd = torch.from_numpy(q).to(dev)
        pd, fd, _ = lfmm.fmm3d(xd, qd, terms=8, tree_depth=5)
        dp, df = lfmm.direct3d(xd[:2000], qd[:2000])
        out = (pd + 0.0, dp + 0.0)  # consumed by torch kernels on the side stream
    s.synchronize()
    assert np.array_equal(out[0].cpu().numpy(), pf)
    assert np.array_equal(out[1].cpu().numpy(), ph)
