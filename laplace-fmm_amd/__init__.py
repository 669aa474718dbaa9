"""Python binding of the MI355X Laplace FMM C ABI (include/lfmm.h).

The product is the HIP library laplace-fmm_amd/_build/liblfmm.so; this module
only marshals arguments to it through ctypes (host numpy arrays, or device
buffers given as torch tensors).  It mirrors the reference's drop-in class
fmm::gpu::FastMultipole3d<double> (src/cuda/multipole3d.h:7-42): construct
with particles, eps, N, tree_depth and read `.potentials` and `.forces`.

There is no CPU fallback: if the library is missing or no GPU is present the
calls raise.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liblfmm.so")
CSRC = os.path.join(_HERE, "csrc")

FMM_AUTO = 2147483647
MAX_TERMS_3D = 18
MEM_HOST, MEM_DEVICE = 0, 1
FLAG_EXCLUDE_SELF, FLAG_TIMINGS, FLAG_COUNT_OPS, FLAG_MIXED, FLAG_ADAPTIVE, FLAG_LEARNED_ORDER = 1, 2, 4, 8, 16, 32
ORDER_FEATURES, ORDER_CLASSES, ORDER_MIN_TERMS, ORDER_MODEL_SIZE = 8, 13, 5, 5597
STREAM_DEFAULT = 1  # LFMM_STREAM_DEFAULT: the null stream

STATUS = {0: "ok", 1: "invalid argument", 2: "expansion terms out of range", 3: "tree depth out of range",
          4: "out of device memory", 5: "HIP runtime error", 6: "non-finite input", 7: "no usable GPU",
          8: "unsupported option", 9: "input larger than a fixed scratch"}

# Every extern "C" entry point declared in include/lfmm.h.
EXPORTS = ("lfmm_create", "lfmm_destroy", "lfmm_strerror", "lfmm_set_stream", "lfmm_trim", "lfmm_fmm3d",
           "lfmm_tree3d", "lfmm_fmm2d", "lfmm_tree2d", "lfmm_version", "lfmm_host_wigner_d", "lfmm_host_op_tables",
           "lfmm_host_m2l_dense", "lfmm_host_m2m_dense", "lfmm_host_m2l16_pack", "lfmm_direct3d", "lfmm_direct2d", "lfmm_exact_filename",
           "lfmm_write_exact", "lfmm_read_exact", "lfmm_error_stats", "lfmm_fmm3d_vec", "lfmm_direct3d_vec",
           "lfmm_fmm3d_batch", "lfmm_tree3d_adaptive", "lfmm_fmm3d_orders", "lfmm_set_order_model",
           "lfmm_order_features", "lfmm_order_predict", "lfmm_order_loss_grad", "lfmm_order_train",
           "lfmm_direct3d_noalloc")


class LfmmError(RuntimeError):
    def __init__(self, code, where=""):
        self.code = code
        super().__init__("%s: lfmm status %d (%s)" % (where, code, STATUS.get(code, "?")))


class Points3d(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("pos", ctypes.c_void_p), ("pos_stride", ctypes.c_int64),
                ("q", ctypes.c_void_p), ("q_stride", ctypes.c_int64)]


class Report(ctypes.Structure):
    _fields_ = [("terms", ctypes.c_int32), ("tree_depth", ctypes.c_int32), ("num_leaves", ctypes.c_int64),
                ("max_leaf", ctypes.c_int64), ("m2l_ops", ctypes.c_int64), ("p2p_pairs", ctypes.c_int64),
                ("t_tree_ms", ctypes.c_double), ("t_upward_ms", ctypes.c_double),
                ("t_downward_ms", ctypes.c_double), ("t_leaf_ms", ctypes.c_double),
                ("t_total_ms", ctypes.c_double), ("t_m2l_kernels_ms", ctypes.c_double),
                ("m2l_launches", ctypes.c_int32), ("m2l_mixed_launches", ctypes.c_int32),
                ("t_leaf_kernel_ms", ctypes.c_double), ("t_m2l_mixed_ms", ctypes.c_double),
                ("m2l_ops_mixed", ctypes.c_int64), ("num_boxes", ctypes.c_int64), ("xw_pairs", ctypes.c_int64),
                ("terms_min", ctypes.c_int32), ("fallback", ctypes.c_int32), ("check_rel_l2", ctypes.c_double),
                ("check_rel_l2_field", ctypes.c_double), ("m2l_flops", ctypes.c_double),
                ("t_p2m_ms", ctypes.c_double), ("t_m2m_ms", ctypes.c_double), ("t_l2l_ms", ctypes.c_double),
                ("t_p2l_ms", ctypes.c_double), ("t_l2p_ms", ctypes.c_double), ("t_p2p_ms", ctypes.c_double),
                ("p2p_source_reads", ctypes.c_int64), ("m2m_bytes", ctypes.c_double), ("l2l_bytes", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def build(jobs=8, force=False):
    """Compile liblfmm.so for gfx950 with hipcc (works without a GPU)."""
    cmd = ["make", "-C", CSRC, "-j%d" % jobs]
    if force:
        cmd.append("-B")
    subprocess.run(cmd, check=True)


def _share_torch_hip_runtime():
    """torch's ROCm build ships its own libamdhip64 (loaded under the name
    libamdhip64.so); liblfmm.so asks for libamdhip64.so.7.  Loading torch first
    makes the dynamic loader bind liblfmm.so to that same runtime, so the
    process holds one HIP runtime and torch tensors and lfmm share devices."""
    try:
        import torch  # noqa: F401
    except Exception:
        pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError("liblfmm.so is not built (%s); run __graft_entry__.build()" % LIB_PATH)
        _share_torch_hip_runtime()
        L = ctypes.CDLL(LIB_PATH)
        L.lfmm_create.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int]
        L.lfmm_create.restype = ctypes.c_int
        L.lfmm_destroy.argtypes = [ctypes.c_void_p]
        L.lfmm_destroy.restype = None
        L.lfmm_strerror.argtypes = [ctypes.c_int]
        L.lfmm_strerror.restype = ctypes.c_char_p
        L.lfmm_set_stream.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        L.lfmm_trim.argtypes = [ctypes.c_void_p]
        L.lfmm_fmm3d.argtypes = [ctypes.c_void_p, ctypes.POINTER(Points3d), ctypes.POINTER(Points3d),
                                 ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_uint, ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(Report)]
        L.lfmm_fmm3d.restype = ctypes.c_int
        L.lfmm_tree3d.argtypes = [ctypes.c_void_p, ctypes.POINTER(Points3d), ctypes.c_int, ctypes.c_int,
                                  ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                  ctypes.c_void_p, ctypes.c_void_p]
        L.lfmm_tree3d.restype = ctypes.c_int
        L.lfmm_fmm2d.argtypes = L.lfmm_fmm3d.argtypes
        L.lfmm_fmm2d.restype = ctypes.c_int
        L.lfmm_tree2d.argtypes = L.lfmm_tree3d.argtypes
        L.lfmm_tree2d.restype = ctypes.c_int
        L.lfmm_version.restype = ctypes.c_char_p
        L.lfmm_host_wigner_d.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_void_p]
        L.lfmm_host_wigner_d.restype = ctypes.c_int
        L.lfmm_host_op_tables.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int64]
        L.lfmm_host_op_tables.restype = ctypes.c_int64
        L.lfmm_host_m2l_dense.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
        L.lfmm_host_m2l_dense.restype = ctypes.c_int
        L.lfmm_host_m2m_dense.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
        L.lfmm_host_m2m_dense.restype = ctypes.c_int
        L.lfmm_host_m2l16_pack.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                                           ctypes.c_void_p]
        L.lfmm_host_m2l16_pack.restype = ctypes.c_int64
        L.lfmm_direct3d.argtypes = [ctypes.c_void_p, ctypes.POINTER(Points3d), ctypes.POINTER(Points3d), ctypes.c_int,
                                    ctypes.c_uint, ctypes.c_void_p, ctypes.c_void_p]
        L.lfmm_direct3d.restype = ctypes.c_int
        L.lfmm_direct2d.argtypes = L.lfmm_direct3d.argtypes
        L.lfmm_direct2d.restype = ctypes.c_int
        L.lfmm_exact_filename.argtypes = [ctypes.c_int]
        L.lfmm_exact_filename.restype = ctypes.c_char_p
        L.lfmm_write_exact.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int64, ctypes.c_void_p]
        L.lfmm_write_exact.restype = ctypes.c_int
        L.lfmm_read_exact.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int64, ctypes.c_void_p]
        L.lfmm_read_exact.restype = ctypes.c_int64
        L.lfmm_error_stats.argtypes = [ctypes.c_int, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                       ctypes.POINTER(ctypes.c_double)]
        L.lfmm_error_stats.restype = ctypes.c_int
        L.lfmm_fmm3d_vec.argtypes = [ctypes.c_void_p, ctypes.POINTER(Points3d), ctypes.POINTER(Points3d),
                                     ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_uint, ctypes.c_void_p, ctypes.POINTER(Report)]
        L.lfmm_fmm3d_vec.restype = ctypes.c_int
        L.lfmm_fmm3d_batch.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_int, ctypes.c_uint, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_void_p]
        L.lfmm_fmm3d_batch.restype = ctypes.c_int
        L.lfmm_direct3d_vec.argtypes = [ctypes.c_void_p, ctypes.POINTER(Points3d), ctypes.POINTER(Points3d),
                                        ctypes.c_int, ctypes.c_void_p]
        L.lfmm_direct3d_vec.restype = ctypes.c_int
        L.lfmm_direct3d_noalloc.argtypes = [ctypes.c_void_p, ctypes.POINTER(Points3d), ctypes.POINTER(Points3d),
                                            ctypes.c_uint, ctypes.c_void_p, ctypes.c_void_p]
        L.lfmm_direct3d_noalloc.restype = ctypes.c_int
        L.lfmm_tree3d_adaptive.argtypes = [ctypes.c_void_p, ctypes.POINTER(Points3d), ctypes.c_int, ctypes.c_int] + \
            [ctypes.c_void_p] * 13
        L.lfmm_tree3d_adaptive.restype = ctypes.c_int
        L.lfmm_fmm3d_orders.argtypes = [ctypes.c_void_p, ctypes.POINTER(Points3d), ctypes.POINTER(Points3d),
                                        ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint,
                                        ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(Report)]
        L.lfmm_fmm3d_orders.restype = ctypes.c_int
        L.lfmm_set_order_model.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_double,
                                           ctypes.c_int]
        L.lfmm_set_order_model.restype = ctypes.c_int
        L.lfmm_order_features.argtypes = [ctypes.c_void_p, ctypes.POINTER(Points3d), ctypes.c_int, ctypes.c_int,
                                          ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.lfmm_order_features.restype = ctypes.c_int
        L.lfmm_order_predict.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                         ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p]
        L.lfmm_order_predict.restype = ctypes.c_int
        L.lfmm_order_loss_grad.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                           ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p]
        L.lfmm_order_loss_grad.restype = ctypes.c_int
        L.lfmm_order_train.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_uint64,
                                       ctypes.c_void_p]
        L.lfmm_order_train.restype = ctypes.c_int
        _lib = L
    return _lib


def _check(rc, where):
    if rc != 0:
        raise LfmmError(rc, where)


class Context:
    """One lfmm_ctx (a HIP stream plus cached workspace) on one GPU."""

    def __init__(self, device=0):
        self._h = ctypes.c_void_p()
        self.device = int(device)
        _check(lib().lfmm_create(ctypes.byref(self._h), int(device)), "lfmm_create")

    @property
    def handle(self):
        return self._h

    def set_stream(self, stream_ptr):
        _check(lib().lfmm_set_stream(self._h, ctypes.c_void_p(stream_ptr)), "lfmm_set_stream")

    def close(self):
        if self._h:
            lib().lfmm_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx = None


def default_context():
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


def _is_torch_cuda(x):
    return hasattr(x, "is_cuda") and x.is_cuda


def _bind_stream(ctx, t):
    """Runs the context's next calls on torch's current stream of t's device,
    so they follow the torch work that produced the inputs and precede the
    torch work that consumes the outputs (a device solve returns before its
    kernels finish)."""
    import torch
    if t.device.index is not None and t.device.index != ctx.device:
        raise ValueError("tensors are on %s but the context is bound to GPU %d" % (t.device, ctx.device))
    handle = torch.cuda.current_stream(t.device).cuda_stream
    # torch's default stream is the null stream (handle 0); a NULL argument
    # would select the context's own stream instead
    ctx.set_stream(handle if handle else STREAM_DEFAULT)


def _dev_points(name, t, dim, device):
    """Checks a CUDA float64 (n, dim) point tensor, contiguous per point."""
    import torch
    if not _is_torch_cuda(t):
        raise TypeError("%s must be a CUDA tensor like the other device inputs" % name)
    if t.dtype != torch.float64:
        raise TypeError("%s must be float64" % name)
    if t.dim() != 2 or t.shape[1] != dim:
        raise ValueError("%s must have shape (n, %d), got %s" % (name, dim, tuple(t.shape)))
    if t.shape[0] > 0 and t.stride(1) != 1:
        raise ValueError("%s must be contiguous per point" % name)
    if t.device != device:
        raise ValueError("%s is on %s, the positions on %s" % (name, t.device, device))
    return t.shape[0]


def _dev_vector(name, t, n, device, width=None):
    """Checks a CUDA float64 tensor of n rows (1-D, or (n, width) contiguous per row)."""
    import torch
    if not _is_torch_cuda(t):
        raise TypeError("%s must be a CUDA tensor like the other device inputs" % name)
    if t.dtype != torch.float64:
        raise TypeError("%s must be float64" % name)
    if t.device != device:
        raise ValueError("%s is on %s, the positions on %s" % (name, t.device, device))
    if width is None:
        if t.dim() != 1 or t.shape[0] != n:
            raise ValueError("%s must have shape (%d,), got %s" % (name, n, tuple(t.shape)))
        if n > 0 and t.stride(0) < 1:
            raise ValueError("%s must have a positive stride" % name)
    else:
        if t.dim() != 2 or t.shape[0] != n or t.shape[1] != width:
            raise ValueError("%s must have shape (%d, %d), got %s" % (name, n, width, tuple(t.shape)))
        if n > 0 and t.stride(1) != 1:
            raise ValueError("%s must be contiguous per row" % name)


def _dev_out(name, t, n, device, width=None):
    """Checks an output tensor: CUDA float64, n rows, densely packed."""
    _dev_vector(name, t, n, device, width)
    if n > 0 and not t.is_contiguous():
        raise ValueError("%s must be contiguous" % name)


def _host_points(name, a, dim):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if a.ndim != 2 or a.shape[1] != dim:
        if a.ndim == 1 and a.size % dim == 0:
            a = a.reshape(-1, dim)
        else:
            raise ValueError("%s must have shape (n, %d), got %s" % (name, dim, a.shape))
    return a


def _host_charges(name, a, n, width=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    want = (n,) if width is None else (n, width)
    if a.shape != want:
        if width is not None and a.ndim == 1 and a.size == n * width:
            return a.reshape(n, width)
        raise ValueError("%s must have shape %s to match the positions, got %s" % (name, want, a.shape))
    return a


def _solve(dim, xyz, q, eps, terms, tree_depth, leaf_max, exclude_self, timings, count_ops, ctx, out, targets,
           mixed, adaptive=False, learned_order=False, particle_terms=None):
    fn = lib().lfmm_fmm3d if dim == 3 else lib().lfmm_fmm2d
    name = "lfmm_fmm3d" if dim == 3 else "lfmm_fmm2d"
    flags = ((FLAG_EXCLUDE_SELF if exclude_self else 0) | (FLAG_TIMINGS if timings else 0) |
             (FLAG_COUNT_OPS if count_ops else 0) | (FLAG_MIXED if mixed else 0) |
             (FLAG_ADAPTIVE if adaptive else 0) | (FLAG_LEARNED_ORDER if learned_order else 0))
    rep = Report()
    req_ptr = None
    if particle_terms is not None:
        if dim != 3:
            raise ValueError("per-particle orders are a 3D option")
        def fn(c, pts, tg, eps_, terms_, depth_, leaf_, mem, flags_, pot_, field_, rep_):
            return lib().lfmm_fmm3d_orders(c, pts, tg, ctypes.c_void_p(req_ptr), depth_, leaf_, mem, flags_, pot_,
                                           field_, rep_)
        name = "lfmm_fmm3d_orders"
    if _is_torch_cuda(xyz):
        import torch
        n = _dev_points("positions", xyz, dim, xyz.device)
        _dev_vector("charges", q, n, xyz.device)
        pts = Points3d(n, xyz.data_ptr(), xyz.stride(0), q.data_ptr(), q.stride(0))
        tg = None
        m = n
        if targets is not None:
            m = _dev_points("targets", targets, dim, xyz.device)
            tg = ctypes.byref(Points3d(m, targets.data_ptr(), targets.stride(0), None, 1))
        if out is None:
            pot = torch.empty(m, dtype=torch.float64, device=xyz.device)
            field = torch.empty((m, dim), dtype=torch.float64, device=xyz.device)
        else:
            pot, field = out
            _dev_out("out potential", pot, m, xyz.device)
            _dev_out("out field", field, m, xyz.device, dim)
        if particle_terms is not None:
            if not _is_torch_cuda(particle_terms) or particle_terms.dtype != torch.int32:
                raise TypeError("particle_terms must be a CUDA int32 tensor like the other device inputs")
            if particle_terms.device != xyz.device or tuple(particle_terms.shape) != (m,) or \
                    (m > 0 and not particle_terms.is_contiguous()):
                raise ValueError("particle_terms must be contiguous with shape (%d,) on %s" % (m, xyz.device))
            req_ptr = particle_terms.data_ptr()
        ctx = ctx or default_context()
        _bind_stream(ctx, xyz)
        rc = fn(ctx.handle, ctypes.byref(pts), tg, float(eps), int(terms), int(tree_depth), int(leaf_max), MEM_DEVICE,
                flags, ctypes.c_void_p(pot.data_ptr()), ctypes.c_void_p(field.data_ptr()), ctypes.byref(rep))
        _check(rc, name)
        return pot, field, rep.as_dict()
    if out is not None:
        raise ValueError("out= takes CUDA tensors (device inputs); host calls return new arrays")
    xyz = _host_points("positions", xyz, dim)
    n = len(xyz)
    q = _host_charges("charges", q, n)
    pts = Points3d(n, xyz.ctypes.data, dim, q.ctypes.data, 1)
    tg = None
    m = n
    if targets is not None:
        if _is_torch_cuda(targets):
            raise TypeError("targets are a CUDA tensor but the sources are host arrays")
        targets = _host_points("targets", targets, dim)
        m = len(targets)
        tg = ctypes.byref(Points3d(m, targets.ctypes.data, dim, None, 1))
    pot = np.empty(m)
    field = np.empty((m, dim))
    if particle_terms is not None:
        if _is_torch_cuda(particle_terms):
            raise TypeError("particle_terms is a CUDA tensor but the sources are host arrays")
        pt_ = np.ascontiguousarray(particle_terms, dtype=np.int32)
        if pt_.shape != (m,):
            raise ValueError("particle_terms must have shape (%d,), got %s" % (m, pt_.shape))
        req_ptr = pt_.ctypes.data
    ctx = ctx or default_context()
    rc = fn(ctx.handle, ctypes.byref(pts), tg, float(eps), int(terms), int(tree_depth), int(leaf_max), MEM_HOST, flags,
            pot.ctypes.data, field.ctypes.data, ctypes.byref(rep))
    _check(rc, name)
    return pot, field, rep.as_dict()


def fmm3d(xyz, q, eps=1e-8, terms=FMM_AUTO, tree_depth=FMM_AUTO, leaf_max=0, exclude_self=False,
          timings=False, count_ops=False, ctx=None, out=None, targets=None, mixed=False, adaptive=False,
          learned_order=False, particle_terms=None):
    """3D FMM: potential (m,) and field (m,3) at every target.

    xyz (n,3) and q (n,) are the sources; targets (m,3) defaults to the
    sources themselves (m = n).  float64 numpy arrays take the host path,
    CUDA torch tensors the device path (outputs are then CUDA tensors, or
    written into `out=(pot, field)`).  Returns (pot, field, report_dict).
    mixed=True runs the M2L of levels >= 4 on the matrix cores in split FP16
    with FP32 accumulation (LFMM_FLAG_MIXED); everything else stays FP64.
    adaptive=True builds the adaptive octree (LFMM_FLAG_ADAPTIVE): boxes of
    more than leaf_max particles (64 if 0) are split, leaves on any level.
    learned_order=True lets the context's order network (set_order_model)
    pick the order of every leaf (LFMM_FLAG_LEARNED_ORDER; terms and eps are
    ignored; a sampled direct sum checks the run, report['fallback'] = 1 when
    it was rerun at p = 16).  particle_terms (int32, one per target) asks for
    at least that many terms at each target (lfmm_fmm3d_orders).
    """
    return _solve(3, xyz, q, eps, terms, tree_depth, leaf_max, exclude_self, timings, count_ops, ctx, out, targets,
                  mixed, adaptive, learned_order, particle_terms)


def fmm2d(xy, q, eps=1e-8, terms=FMM_AUTO, tree_depth=FMM_AUTO, leaf_max=0, exclude_self=False,
          timings=False, count_ops=False, ctx=None, out=None, targets=None, mixed=False):
    """2D FMM (complex series): potential (m,) = sum q log|z - z_j| and the
    complex force (m,2) = (re, im) of sum q / (z - z_j) at every target.
    Arguments as fmm3d with 2 coordinates per point; mixed=True forms the
    M2L sums of levels >= 4 in FP32 (LFMM_FLAG_MIXED)."""
    return _solve(2, xy, q, eps, terms, tree_depth, leaf_max, exclude_self, timings, count_ops, ctx, out, targets,
                  mixed)


# result kinds (the reference's fmm::InteractionType order)
POTENTIAL2D, FORCE2D, POTENTIAL3D, FORCE3D = 0, 1, 2, 3
_KIND_WIDTH = {POTENTIAL2D: 1, FORCE2D: 2, POTENTIAL3D: 1, FORCE3D: 3}


def _direct(dim, xyz, q, targets, exclude_self, ctx):
    fn = lib().lfmm_direct3d if dim == 3 else lib().lfmm_direct2d
    flags = FLAG_EXCLUDE_SELF if exclude_self else 0
    if _is_torch_cuda(xyz):
        import torch
        n = _dev_points("positions", xyz, dim, xyz.device)
        _dev_vector("charges", q, n, xyz.device)
        pts = Points3d(n, xyz.data_ptr(), xyz.stride(0), q.data_ptr(), q.stride(0))
        tg, m = None, n
        if targets is not None:
            m = _dev_points("targets", targets, dim, xyz.device)
            tg = ctypes.byref(Points3d(m, targets.data_ptr(), targets.stride(0), None, 1))
        pot = torch.empty(m, dtype=torch.float64, device=xyz.device)
        field = torch.empty((m, dim), dtype=torch.float64, device=xyz.device)
        ctx = ctx or default_context()
        _bind_stream(ctx, xyz)
        _check(fn(ctx.handle, ctypes.byref(pts), tg, MEM_DEVICE, flags, ctypes.c_void_p(pot.data_ptr()),
                  ctypes.c_void_p(field.data_ptr())), fn.__name__)
        return pot, field
    xyz = _host_points("positions", xyz, dim)
    q = _host_charges("charges", q, len(xyz))
    pts = Points3d(len(q), xyz.ctypes.data, dim, q.ctypes.data, 1)
    tg, m = None, len(q)
    if targets is not None:
        targets = _host_points("targets", targets, dim)
        m = len(targets)
        tg = ctypes.byref(Points3d(m, targets.ctypes.data, dim, None, 1))
    pot = np.empty(m)
    field = np.empty((m, dim))
    ctx = ctx or default_context()
    _check(fn(ctx.handle, ctypes.byref(pts), tg, MEM_HOST, flags, pot.ctypes.data, field.ctypes.data), fn.__name__)
    return pot, field


def _vec_call(fn, name, xyz, qvec, targets, ctx, extra):
    if _is_torch_cuda(xyz):
        import torch
        n = _dev_points("positions", xyz, 3, xyz.device)
        _dev_vector("charge vectors", qvec, n, xyz.device, 3)
        pts = Points3d(n, xyz.data_ptr(), xyz.stride(0), qvec.data_ptr(), qvec.stride(0))
        tg, m = None, n
        if targets is not None:
            m = _dev_points("targets", targets, 3, xyz.device)
            tg = ctypes.byref(Points3d(m, targets.data_ptr(), targets.stride(0), None, 1))
        force = torch.empty((m, 3), dtype=torch.float64, device=xyz.device)
        ctx = ctx or default_context()
        _bind_stream(ctx, xyz)
        _check(fn(ctx.handle, ctypes.byref(pts), tg, *extra(MEM_DEVICE, ctypes.c_void_p(force.data_ptr()))), name)
        return force
    xyz = _host_points("positions", xyz, 3)
    qvec = _host_charges("charge vectors", qvec, len(xyz), 3)
    pts = Points3d(len(qvec), xyz.ctypes.data, 3, qvec.ctypes.data, 3)
    tg, m = None, len(qvec)
    if targets is not None:
        targets = _host_points("targets", targets, 3)
        m = len(targets)
        tg = ctypes.byref(Points3d(m, targets.ctypes.data, 3, None, 1))
    force = np.empty((m, 3))
    ctx = ctx or default_context()
    _check(fn(ctx.handle, ctypes.byref(pts), tg, *extra(MEM_HOST, force.ctypes.data)), name)
    return force


def fmm3d_vec(xyz, qvec, eps=1e-8, terms=FMM_AUTO, tree_depth=FMM_AUTO, leaf_max=0, ctx=None, targets=None,
              mixed=False):
    """FMM with vector charges (reference FastMultipole3d<Vector3d>): the
    force sum q_j x (r_i - r_j) / |r_i - r_j|^3 (m,3) at every target."""
    rep = Report()
    flags = FLAG_MIXED if mixed else 0
    force = _vec_call(lib().lfmm_fmm3d_vec, "lfmm_fmm3d_vec", xyz, qvec, targets, ctx,
                      lambda mem, out: (float(eps), int(terms), int(tree_depth), int(leaf_max), mem, flags, out,
                                        ctypes.byref(rep)))
    return force, rep.as_dict()


def fmm3d_batch(xyz, q, offsets, eps=1e-8, terms=FMM_AUTO, tree_depth=FMM_AUTO, leaf_max=0, exclude_self=False,
                mixed=False, workers=0, ctx=None, out=None, reports=False, adaptive=False, timings=False,
                count_ops=False):
    """Many independent single-set 3D systems (lfmm_fmm3d_batch).

    xyz (total,3) and q (total,) hold the systems packed one after another;
    system s is rows offsets[s] .. offsets[s+1]-1 (offsets: n_sys + 1 ints).
    Returns (pot (total,), field (total,3)[, list of report dicts]) in the same
    rows.  float64 numpy arrays take the host path, CUDA torch tensors the
    device path (outputs are then CUDA tensors, or written into out).
    adaptive=True solves every system on its own adaptive octree, all of them
    as one forest in one solve (tree_depth and workers are then ignored and
    every report is the forest's); otherwise worker contexts run one
    uniform-tree solve per system."""
    off = np.ascontiguousarray(offsets, dtype=np.int64)
    n_sys = len(off) - 1
    if n_sys < 0:
        raise ValueError("offsets needs n_sys + 1 entries")
    flags = ((FLAG_EXCLUDE_SELF if exclude_self else 0) | (FLAG_MIXED if mixed else 0) |
             (FLAG_ADAPTIVE if adaptive else 0) | (FLAG_TIMINGS if timings else 0) |
             (FLAG_COUNT_OPS if count_ops else 0))
    reps = (Report * max(n_sys, 1))() if reports else None
    if _is_torch_cuda(xyz):
        import torch
        total = _dev_points("positions", xyz, 3, xyz.device)
        _dev_out("charges", q, total, xyz.device)
        if total > 0 and not xyz.is_contiguous():
            raise ValueError("positions must be contiguous (packed systems)")
        if out is None:
            pot = torch.empty(total, dtype=torch.float64, device=xyz.device)
            field = torch.empty((total, 3), dtype=torch.float64, device=xyz.device)
        else:
            pot, field = out
            _dev_out("out potential", pot, total, xyz.device)
            _dev_out("out field", field, total, xyz.device, 3)
        ctx = ctx or default_context()
        _bind_stream(ctx, xyz)
        args = (xyz.data_ptr(), q.data_ptr(), MEM_DEVICE, pot.data_ptr(), field.data_ptr())
    else:
        if out is not None:
            raise ValueError("out= takes CUDA tensors (device inputs); host calls return new arrays")
        xyz = _host_points("positions", xyz, 3)
        total = len(xyz)
        q = _host_charges("charges", q, total)
        pot = np.empty(total)
        field = np.empty((total, 3))
        args = (xyz.ctypes.data, q.ctypes.data, MEM_HOST, pot.ctypes.data, field.ctypes.data)
    if n_sys > 0 and (off[0] != 0 or off[-1] != total):
        raise ValueError("offsets must run from 0 to the particle count")
    if n_sys > 0 and np.any(np.diff(off) < 0):
        raise ValueError("offsets must be non-decreasing")
    ctx = ctx or default_context()
    rc = lib().lfmm_fmm3d_batch(ctx.handle, n_sys, off.ctypes.data, ctypes.c_void_p(args[0]), ctypes.c_void_p(args[1]),
                                float(eps), int(terms), int(tree_depth), int(leaf_max), args[2], flags, int(workers),
                                ctypes.c_void_p(args[3]), ctypes.c_void_p(args[4]),
                                ctypes.cast(reps, ctypes.c_void_p) if reps is not None else None)
    _check(rc, "lfmm_fmm3d_batch")
    if reports:
        return pot, field, [reps[s].as_dict() for s in range(n_sys)]
    return pot, field


def direct3d_vec(xyz, qvec, targets=None, ctx=None):
    """FP64 direct sums of the vector-charge force (reference ComputeExact on particle3d3)."""
    return _vec_call(lib().lfmm_direct3d_vec, "lfmm_direct3d_vec", xyz, qvec, targets, ctx,
                     lambda mem, out: (mem, out))


def direct3d(xyz, q, targets=None, exclude_self=False, ctx=None):
    """FP64 direct sums on the GPU (reference fmm::gpu::ComputeExact with
    Potential3d / Force3d): potential (m,) and field (m,3) at every target,
    the self pair included unless exclude_self (single-set calls)."""
    return _direct(3, xyz, q, targets, exclude_self, ctx)


def direct2d(xy, q, targets=None, exclude_self=False, ctx=None):
    """FP64 direct sums on the GPU (Potential2d / Force2d): potential (m,)
    and the complex force (m,2) as (re, im)."""
    return _direct(2, xy, q, targets, exclude_self, ctx)


def exact_filename(kind):
    """The reference's exact-result file name for a result kind."""
    return lib().lfmm_exact_filename(int(kind)).decode()


def write_exact(path, kind, values):
    """Writes values in the reference's exact-result text format."""
    v = np.ascontiguousarray(values, dtype=np.float64).reshape(-1, _KIND_WIDTH[kind])
    _check(lib().lfmm_write_exact(str(path).encode(), int(kind), len(v), v.ctypes.data), "lfmm_write_exact")


def read_exact(path, kind, n):
    """Reads n values of a reference exact-result file; returns (count read, values)."""
    v = np.zeros((int(n), _KIND_WIDTH[kind]))
    got = lib().lfmm_read_exact(str(path).encode(), int(kind), int(n), v.ctypes.data)
    if got < 0:
        raise LfmmError(int(-got), "lfmm_read_exact")
    return int(got), (v[:, 0] if _KIND_WIDTH[kind] == 1 else v)


def error_stats(kind, exact, result):
    """The reference's ReadError figures: (max error, l2 error, relative error)."""
    w = _KIND_WIDTH[kind]
    a = np.ascontiguousarray(exact, dtype=np.float64).reshape(-1, w)
    b = np.ascontiguousarray(result, dtype=np.float64).reshape(-1, w)
    if a.shape != b.shape:
        raise ValueError("shape mismatch")
    mx, l2, rel = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    _check(lib().lfmm_error_stats(int(kind), len(a), a.ctypes.data, b.ctypes.data, ctypes.byref(mx), ctypes.byref(l2),
                                  ctypes.byref(rel)), "lfmm_error_stats")
    return mx.value, l2.value, rel.value


def _tree(dim, xyz, q, tree_depth, leaf_max, ctx):
    ctx = ctx or default_context()
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, dim)
    n = len(xyz)
    if q is not None:
        q = np.ascontiguousarray(q, dtype=np.float64)
    pts = Points3d(n, xyz.ctypes.data, dim, q.ctypes.data if q is not None else None, 1)
    sizes = np.zeros(32, dtype=np.int64)
    geom = np.zeros(4)
    pos_map = np.zeros(n, dtype=np.int64)
    depth = np.zeros(1, dtype=np.int32)
    # total boxes <= n * depth
    cap = n * (max(int(tree_depth), 1) if tree_depth != FMM_AUTO else 32) + 64
    keys = np.zeros(cap, dtype=np.uint64)
    ranges = np.zeros(cap + 64, dtype=np.int64)
    fn = lib().lfmm_tree3d if dim == 3 else lib().lfmm_tree2d
    rc = fn(ctx.handle, ctypes.byref(pts), int(tree_depth), int(leaf_max), depth.ctypes.data, sizes.ctypes.data,
            geom.ctypes.data, pos_map.ctypes.data, keys.ctypes.data, ranges.ctypes.data)
    _check(rc, "lfmm_tree3d" if dim == 3 else "lfmm_tree2d")
    d = int(depth[0])
    out = {"depth": d, "level_sizes": sizes[:d].copy(), "shift": geom[:dim].copy(), "scale": float(geom[3]),
           "pos_map": pos_map, "keys": [], "begin": []}
    o = b = 0
    for l in range(d):
        s = int(sizes[l])
        out["keys"].append(keys[o:o + s].copy())
        out["begin"].append(ranges[b:b + s + 1].copy())
        o += s
        b += s + 1
    return out


def _f32(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if shape is not None and a.shape != shape:
        raise ValueError("expected shape %s, got %s" % (shape, a.shape))
    return a


def set_order_model(model, confidence=0.95, check_tol=1e-6, check_samples=256, ctx=None):
    """Installs the order network (ORDER_MODEL_SIZE floats) on the context for
    learned_order solves; None removes it (lfmm_set_order_model)."""
    ctx = ctx or default_context()
    if model is None:
        _check(lib().lfmm_set_order_model(ctx.handle, None, 0.0, 0.0, 0), "lfmm_set_order_model")
        return
    m = _f32(model, (ORDER_MODEL_SIZE,))
    _check(lib().lfmm_set_order_model(ctx.handle, m.ctypes.data, float(confidence), float(check_tol),
                                      int(check_samples)), "lfmm_set_order_model")


def order_features(xyz, q, leaf_max=64, tree_depth=FMM_AUTO, ctx=None):
    """Per-leaf features (nleaf, 8) of the adaptive octree of a solve with the
    same leaf_max and tree_depth, and the leaf of every particle (n,)."""
    xyz = _host_points("positions", xyz, 3)
    q = _host_charges("charges", q, len(xyz))
    ctx = ctx or default_context()
    pts = Points3d(len(xyz), xyz.ctypes.data, 3, q.ctypes.data, 1)
    nl = np.zeros(1, np.int64)
    _check(lib().lfmm_order_features(ctx.handle, ctypes.byref(pts), int(tree_depth), int(leaf_max), nl.ctypes.data,
                                     None, None), "lfmm_order_features")
    feat = np.zeros((int(nl[0]), ORDER_FEATURES), np.float32)
    leaf = np.zeros(len(xyz), np.int32)
    _check(lib().lfmm_order_features(ctx.handle, ctypes.byref(pts), int(tree_depth), int(leaf_max), nl.ctypes.data,
                                     feat.ctypes.data, leaf.ctypes.data), "lfmm_order_features")
    return feat, leaf


def order_predict(model, features, confidence=0.95, ctx=None):
    """Class probabilities (n, 13) and chosen terms (n,) of feature rows."""
    ctx = ctx or default_context()
    m = _f32(model, (ORDER_MODEL_SIZE,))
    f = _f32(features)
    n = len(f)
    probs = np.zeros((n, ORDER_CLASSES), np.float32)
    terms = np.zeros(n, np.int32)
    _check(lib().lfmm_order_predict(ctx.handle, m.ctypes.data, f.ctypes.data, n, float(confidence), probs.ctypes.data,
                                    terms.ctypes.data), "lfmm_order_predict")
    return probs, terms


def order_loss_grad(model, features, labels, ctx=None):
    """Mean cross-entropy of labelled feature rows and its gradient (GPU)."""
    ctx = ctx or default_context()
    m = _f32(model, (ORDER_MODEL_SIZE,))
    f = _f32(features)
    y = np.ascontiguousarray(labels, dtype=np.int32)
    loss = np.zeros(1, np.float32)
    grad = np.zeros(ORDER_MODEL_SIZE, np.float32)
    _check(lib().lfmm_order_loss_grad(ctx.handle, m.ctypes.data, f.ctypes.data, y.ctypes.data, len(y), loss.ctypes.data,
                                      grad.ctypes.data), "lfmm_order_loss_grad")
    return float(loss[0]), grad


def order_train(model, features, labels, epochs=30, batch=1024, lr=3e-3, seed=1, ctx=None):
    """Adam on the GPU; returns (trained model, mean loss per epoch)."""
    ctx = ctx or default_context()
    m = _f32(model, (ORDER_MODEL_SIZE,)).copy()
    f = _f32(features)
    y = np.ascontiguousarray(labels, dtype=np.int32)
    hist = np.zeros(int(epochs), np.float32)
    _check(lib().lfmm_order_train(ctx.handle, m.ctypes.data, f.ctypes.data, y.ctypes.data, len(y), int(epochs),
                                  int(batch), float(lr), int(seed), hist.ctypes.data), "lfmm_order_train")
    return m, hist


def tree3d_adaptive(xyz, leaf_max=64, tree_depth=FMM_AUTO, ctx=None):
    """The adaptive octree and its U, X, W lists (lfmm_tree3d_adaptive) as a
    dict of numpy arrays: depth, geom, box_level, box_key, box_range (sorted
    particle ranges), leaf_box, u_off / u_leaf, x_off / x_leaf, w_off / w_box."""
    xyz = _host_points("positions", xyz, 3)
    ctx = ctx or default_context()
    pts = Points3d(len(xyz), xyz.ctypes.data, 3, None, 1)
    counts = np.zeros(4, np.int64)
    nul = [None] * 13
    nul[0] = counts.ctypes.data
    _check(lib().lfmm_tree3d_adaptive(ctx.handle, ctypes.byref(pts), int(tree_depth), int(leaf_max), *nul),
           "lfmm_tree3d_adaptive")
    nb, nl, nu, nx = (int(v) for v in counts)
    out = dict(depth=np.zeros(1, np.int32), geom=np.zeros(4), box_level=np.zeros(nb, np.int32),
               box_key=np.zeros(nb, np.uint64), box_range=np.zeros((nb, 2), np.int32), leaf_box=np.zeros(nl, np.int32),
               u_off=np.zeros(nl + 1, np.int64), u_leaf=np.zeros(max(nu, 1), np.int32),
               x_off=np.zeros(nb + 1, np.int64), x_leaf=np.zeros(max(nx, 1), np.int32),
               w_off=np.zeros(nl + 1, np.int64), w_box=np.zeros(max(nx, 1), np.int32))
    order = ("depth", "geom", "box_level", "box_key", "box_range", "leaf_box", "u_off", "u_leaf", "x_off", "x_leaf",
             "w_off", "w_box")
    _check(lib().lfmm_tree3d_adaptive(ctx.handle, ctypes.byref(pts), int(tree_depth), int(leaf_max),
                                      counts.ctypes.data, *[out[k].ctypes.data for k in order]),
           "lfmm_tree3d_adaptive")
    out["depth"] = int(out["depth"][0])
    out["u_leaf"] = out["u_leaf"][:nu]
    out["x_leaf"] = out["x_leaf"][:nx]
    out["w_box"] = out["w_box"][:nx]
    return out


def tree3d(xyz, q, tree_depth, leaf_max=0, ctx=None):
    """The GPU Morton octree: dict with depth, level sizes, geometry (shift,
    scale), pos_map (input -> sorted position), per-level keys and ranges.
    With charges q, zero-charge particles follow the others inside each leaf."""
    return _tree(3, xyz, q, tree_depth, leaf_max, ctx)


def tree2d(xy, q, tree_depth, leaf_max=0, ctx=None):
    """The GPU Morton quadtree (keys x << 1 | y), as tree3d."""
    return _tree(2, xy, q, tree_depth, leaf_max, ctx)


def host_wigner_d(j, beta):
    """Host table builder's small Wigner d^j(beta), [(2j+1) x (2j+1)] indexed [m'+j, m+j]."""
    out = np.zeros((2 * j + 1) * (2 * j + 1))
    _check(lib().lfmm_host_wigner_d(int(j), float(beta), out.ctypes.data), "lfmm_host_wigner_d")
    return out.reshape(2 * j + 1, 2 * j + 1)


HOST_TABLES = {"rot": 0, "ephi": 1, "trans": 2, "dir_theta": 3, "dir_trans": 4, "dir_vec": 5, "class_dirs": 6,
               "knm": 7}


def host_op_tables(N, which):
    """Copy of one host-built operator table for N terms (float64 or int32 numpy array)."""
    w = HOST_TABLES[which]
    count = lib().lfmm_host_op_tables(int(N), w, None, 0)
    if count < 0:
        raise LfmmError(int(-count), "lfmm_host_op_tables")
    isint = which in ("dir_theta", "dir_trans", "dir_vec", "class_dirs")
    out = np.zeros(count, dtype=np.int32 if isint else np.float64)
    got = lib().lfmm_host_op_tables(int(N), w, out.ctypes.data, count)
    if got != count:
        raise LfmmError(1, "lfmm_host_op_tables")
    return out


def host_m2l_dense(N, direction):
    """Dense real M2L operator (N^2, N^2) of one direction (index into dir_vec)."""
    out = np.zeros((N * N, N * N))
    _check(lib().lfmm_host_m2l_dense(int(N), int(direction), out.ctypes.data), "lfmm_host_m2l_dense")
    return out


def host_m2m_dense(N, cls):
    """Dense real M2M operator (N^2, N^2) from the child of class cls (0..7)
    to its parent, on normalised coefficients."""
    out = np.zeros((N * N, N * N))
    _check(lib().lfmm_host_m2m_dense(int(N), int(cls), out.ctypes.data), "lfmm_host_m2m_dense")
    return out


def host_m2l16_pack(N):
    """Split-fp16 operator fragments of all directions: (uint16 array, kinv,
    block prefixes int32 (316, ceil(N^2/32), 2) = (k3, k1))."""
    kinv = ctypes.c_double(0.0)
    count = lib().lfmm_host_m2l16_pack(int(N), None, 0, ctypes.byref(kinv), None)
    if count < 0:
        raise LfmmError(int(-count), "lfmm_host_m2l16_pack")
    out = np.zeros(count, dtype=np.uint16)
    rt = (N * N + 31) // 32
    kpre = np.zeros((316, rt, 2), dtype=np.int32)
    if lib().lfmm_host_m2l16_pack(int(N), out.ctypes.data, count, ctypes.byref(kinv), kpre.ctypes.data) != count:
        raise LfmmError(1, "lfmm_host_m2l16_pack")
    return out, kinv.value, kpre


class FastMultipole3d:
    """Drop-in for fmm::gpu::FastMultipole3d<double> (src/cuda/multipole3d.h:7-42).

    particles: (n,4) array [x, y, z, q] (or a tuple (xyz, q)).  Computes on
    construction; results in `.potentials` (n,) and `.forces` (n,3), in input
    order.  eps / N / tree_depth follow the reference's FMM_AUTO rules.
    """

    def __init__(self, particles, eps=1e-8, N=FMM_AUTO, tree_depth=FMM_AUTO, leaf_max=0, ctx=None):
        if isinstance(particles, tuple):
            xyz, q = particles
        else:
            p = np.asarray(particles, dtype=np.float64)
            xyz, q = p[:, :3], p[:, 3]
        self.potentials, self.forces, self.report = fmm3d(xyz, q, eps=eps, terms=N, tree_depth=tree_depth,
                                                          leaf_max=leaf_max, ctx=ctx)


from . import order  # noqa: E402  (the order network's training data and driver)
