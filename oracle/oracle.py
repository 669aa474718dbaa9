"""ctypes binding of the CPU oracle (oracle/fmm_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker.  The product library
never loads it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liblfmm_oracle.so")
AUTO = 2147483647

_lib = None

_d = ctypes.POINTER(ctypes.c_double)
_i64 = ctypes.POINTER(ctypes.c_int64)
_u64 = ctypes.POINTER(ctypes.c_uint64)


def build(force=False):
    """Compile the oracle with its Makefile (gcc, no GPU needed)."""
    subprocess.run(["make", "-C", _HERE] + (["-B"] if force else []), check=True,
                   stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.orc_m2lcoef.restype = ctypes.c_double
        L.orc_m2lcoef.argtypes = [ctypes.c_int] * 3
        L.orc_tree3d_build.restype = ctypes.c_void_p
        L.orc_tree3d_build.argtypes = [ctypes.c_int64, _d, _d, ctypes.c_int]
        L.orc_tree2d_build.restype = ctypes.c_void_p
        L.orc_tree2d_build.argtypes = [ctypes.c_int64, _d, _d, ctypes.c_int]
        L.orc_tree3d_free.argtypes = [ctypes.c_void_p]
        L.orc_tree3d_depth.argtypes = [ctypes.c_void_p]
        L.orc_tree3d_info.argtypes = [ctypes.c_void_p, _i64, _d, _i64]
        L.orc_tree3d_particles.argtypes = [ctypes.c_void_p, _u64, _i64, _d]
        L.orc_tree3d_level.argtypes = [ctypes.c_void_p, ctypes.c_int, _u64, _i64, _i64, _i64, _d]
        L.orc_tree3d_list_size.restype = ctypes.c_int64
        L.orc_tree3d_list_size.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        L.orc_tree3d_lists.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _i64, _i64]
        for f in ("orc_fmm3d", "orc_fmm2d"):
            fn = getattr(L, f)
            fn.restype = ctypes.c_int
            fn.argtypes = [ctypes.c_int64, _d, _d, ctypes.c_int64, _d, ctypes.c_double,
                           ctypes.c_int, ctypes.c_int, _d, _d, _d]
        L.orc_fmm3d_orders.restype = ctypes.c_int
        L.orc_fmm3d_orders.argtypes = [ctypes.c_int64, _d, _d, ctypes.c_int64, _d, ctypes.c_void_p, ctypes.c_int,
                                       _d, _d]
        for f in ("orc_direct3d", "orc_direct2d"):
            fn = getattr(L, f)
            fn.argtypes = [ctypes.c_int64, _d, _d, ctypes.c_int64, _i64, _d, ctypes.c_int, _d, _d]
        L.orc_read_error.argtypes = [ctypes.c_int64, ctypes.c_int, _d, _d, _d]
        L.orc_pnm.argtypes = [ctypes.c_int, ctypes.c_double, _d]
        L.orc_wignerd.argtypes = [ctypes.c_int, ctypes.c_double, _d]
        L.orc_dmatrix_lookup.argtypes = [ctypes.c_double, ctypes.c_int, _d]
        L.orc_dec_to_sph.argtypes = [_d, _d]
        L.orc_knm.argtypes = [ctypes.c_int, _d]
        L.orc_anm.argtypes = [ctypes.c_int, _d]
        L.orc_p2m3d.argtypes = [ctypes.c_int, ctypes.c_int64, _d, _d, _d, _d]
        for f in ("orc_m2m3d", "orc_m2l3d", "orc_l2l3d"):
            getattr(L, f).argtypes = [ctypes.c_int, _d, _d, _d]
        L.orc_l2p3d.argtypes = [ctypes.c_int, _d, _d, ctypes.c_int64, _d, _d, _d]
        L.orc_num_threads.restype = ctypes.c_int
        L.orc_set_num_threads.argtypes = [ctypes.c_int]
        L.orc_set_r_eps.argtypes = [ctypes.c_double]
        L.orc_get_r_eps.restype = ctypes.c_double
        L.orc_init()
        _lib = L
    return _lib


def _p(a, t=ctypes.c_double):
    return a.ctypes.data_as(ctypes.POINTER(t)) if a is not None else None


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def tri(N):
    return N * (N + 1) // 2


# ---- tables -------------------------------------------------------------
def knm(N):
    out = np.zeros(tri(N))
    lib().orc_knm(N, _p(out))
    return out


def anm(N):
    out = np.zeros(tri(2 * N))
    lib().orc_anm(N, _p(out))
    return out


def m2lcoef(j, k, n):
    return lib().orc_m2lcoef(j, k, n)


def pnm(N, theta):
    out = np.zeros(tri(N))
    lib().orc_pnm(N, float(theta), _p(out))
    return out


def wignerd(N, theta):
    out = np.zeros(N * (N + 1) * (4 * N - 1) // 6)
    lib().orc_wignerd(N, float(theta), _p(out))
    return out


def dmatrix_lookup(theta, N):
    out = np.zeros(N * (N + 1) * (4 * N - 1) // 6)
    rc = lib().orc_dmatrix_lookup(float(theta), N, _p(out))
    return out if rc == 0 else None


def dec_to_sph(v):
    v = _f64(v)
    out = np.zeros(3)
    lib().orc_dec_to_sph(_p(v), _p(out))
    return out


# ---- single operators ---------------------------------------------------
def p2m3d(N, xyz, q, centre):
    xyz, q, c = _f64(xyz), _f64(q), _f64(centre)
    M = np.zeros(2 * tri(N))
    lib().orc_p2m3d(N, len(q), _p(xyz), _p(q), _p(c), _p(M))
    return M.view(np.complex128)


def _op(fn, N, a, p0, b=None):
    a = np.ascontiguousarray(a, dtype=np.complex128).view(np.float64)
    p0 = _f64(p0)
    out = np.zeros(2 * tri(N)) if b is None else np.ascontiguousarray(b, dtype=np.complex128).view(np.float64).copy()
    fn(N, _p(a), _p(p0), _p(out))
    return out.view(np.complex128)


def m2m3d(N, a, p0, b=None):
    return _op(lib().orc_m2m3d, N, a, p0, b)


def m2l3d(N, a, p0, b=None):
    return _op(lib().orc_m2l3d, N, a, p0, b)


def l2l3d(N, a, p0, b=None):
    return _op(lib().orc_l2l3d, N, a, p0, b)


def l2p3d(N, L, centre, xyz):
    L = np.ascontiguousarray(L, dtype=np.complex128).view(np.float64)
    xyz = _f64(xyz).reshape(-1, 3)
    pot = np.zeros(len(xyz))
    field = np.zeros((len(xyz), 3))
    lib().orc_l2p3d(N, _p(L), _p(_f64(centre)), len(xyz), _p(xyz), _p(pot), _p(field))
    return pot, field


# ---- tree ---------------------------------------------------------------
class Tree3d:
    """The reference MortonTree<3> (stable order), as numpy arrays."""
    dim = 3

    def __init__(self, xyz, q, depth):
        D = self.dim
        xyz = _f64(xyz).reshape(-1, D)
        q = _f64(q)
        self.n = len(q)
        build = lib().orc_tree3d_build if D == 3 else lib().orc_tree2d_build
        h = build(self.n, _p(xyz), _p(q), depth)
        if not h:
            raise ValueError("oracle tree build failed")
        try:
            self.depth = depth
            sizes = np.zeros(depth, dtype=np.int64)
            ss = np.zeros(4)
            tn = np.zeros(1, dtype=np.int64)
            lib().orc_tree3d_info(h, _p(sizes, ctypes.c_int64), _p(ss), _p(tn, ctypes.c_int64))
            self.level_sizes = sizes
            self.shift = ss[:D].copy()
            self.scale = float(ss[D])
            self.targets_num = int(tn[0])
            self.keys = np.zeros(self.n, dtype=np.uint64)
            self.pos_map = np.zeros(self.n, dtype=np.int64)
            self.xyzq = np.zeros((self.n, D + 1))
            lib().orc_tree3d_particles(h, _p(self.keys, ctypes.c_uint64), _p(self.pos_map, ctypes.c_int64),
                                       _p(self.xyzq))
            self.levels = []
            for l in range(depth):
                s = int(sizes[l])
                lv = dict(keys=np.zeros(s, dtype=np.uint64), src=np.zeros((s, 2), dtype=np.int64),
                          tgt=np.zeros((s, 2), dtype=np.int64), parent=np.zeros(s, dtype=np.int64),
                          centre=np.zeros((s, D)))
                lib().orc_tree3d_level(h, l, _p(lv["keys"], ctypes.c_uint64), _p(lv["src"], ctypes.c_int64),
                                       _p(lv["tgt"], ctypes.c_int64), _p(lv["parent"], ctypes.c_int64),
                                       _p(lv["centre"]))
                for far, name in ((0, "near"), (1, "far")):
                    cnt = lib().orc_tree3d_list_size(h, l, far)
                    off = np.zeros(s + 1, dtype=np.int64)
                    idx = np.zeros(max(cnt, 1), dtype=np.int64)
                    lib().orc_tree3d_lists(h, l, far, _p(off, ctypes.c_int64), _p(idx, ctypes.c_int64))
                    lv[name + "_off"] = off
                    lv[name + "_idx"] = idx[:cnt]
                self.levels.append(lv)
        finally:
            lib().orc_tree3d_free(h)


# ---- solves -------------------------------------------------------------
def fmm3d(src_xyz, src_q, N, depth, tgt_xyz=None, eps=1e-8, timings=False):
    src_xyz = _f64(src_xyz).reshape(-1, 3)
    src_q = _f64(src_q)
    nt = 0 if tgt_xyz is None else len(tgt_xyz)
    tgt = None if tgt_xyz is None else _f64(tgt_xyz).reshape(-1, 3)
    out_n = nt if nt else len(src_q)
    pot = np.zeros(out_n)
    field = np.zeros((out_n, 3))
    tm = np.zeros(6)
    rc = lib().orc_fmm3d(len(src_q), _p(src_xyz), _p(src_q), nt, _p(tgt), float(eps), int(N), int(depth),
                         _p(pot), _p(field), _p(tm))
    if rc != 0:
        raise RuntimeError("oracle fmm3d failed: %d" % rc)
    return (pot, field, tm) if timings else (pot, field)


def fmm3d_orders(src_xyz, src_q, pterms, depth):
    """The 3D solve with per-box truncation orders requested per particle
    (fmm3d_impl in fmm_oracle.c)."""
    src_xyz = _f64(src_xyz).reshape(-1, 3)
    src_q = _f64(src_q)
    pt = np.ascontiguousarray(pterms, dtype=np.int32)
    if pt.shape != (len(src_q),):
        raise ValueError("one order request per particle")
    pot = np.zeros(len(src_q))
    field = np.zeros((len(src_q), 3))
    rc = lib().orc_fmm3d_orders(len(src_q), _p(src_xyz), _p(src_q), 0, None, pt.ctypes.data, int(depth), _p(pot),
                                _p(field))
    if rc != 0:
        raise RuntimeError("oracle fmm3d_orders failed: %d" % rc)
    return pot, field


class Tree2d(Tree3d):
    """The reference MortonTree<2> (stable order), as numpy arrays."""
    dim = 2


def fmm2d(src_xy, src_q, N, depth, tgt_xy=None, eps=1e-8, timings=False):
    src_xy = _f64(src_xy).reshape(-1, 2)
    src_q = _f64(src_q)
    nt = 0 if tgt_xy is None else len(tgt_xy)
    tgt = None if tgt_xy is None else _f64(tgt_xy).reshape(-1, 2)
    out_n = nt if nt else len(src_q)
    pot = np.zeros(out_n)
    field = np.zeros((out_n, 2))
    tm = np.zeros(6)
    rc = lib().orc_fmm2d(len(src_q), _p(src_xy), _p(src_q), nt, _p(tgt), float(eps), int(N), int(depth),
                         _p(pot), _p(field), _p(tm))
    if rc != 0:
        raise RuntimeError("oracle fmm2d failed: %d" % rc)
    return (pot, field, tm) if timings else (pot, field)


def direct3d(src_xyz, src_q, tgt_index=None, tgt_xyz=None, include_self=True):
    """FP64 direct sum at the sources listed in tgt_index (or at tgt_xyz)."""
    src_xyz = _f64(src_xyz).reshape(-1, 3)
    src_q = _f64(src_q)
    if tgt_index is not None:
        ti = np.ascontiguousarray(tgt_index, dtype=np.int64)
        nt, tg = len(ti), None
    else:
        ti = None
        tg = _f64(tgt_xyz).reshape(-1, 3)
        nt = len(tg)
    pot = np.zeros(nt)
    field = np.zeros((nt, 3))
    lib().orc_direct3d(len(src_q), _p(src_xyz), _p(src_q), nt, _p(ti, ctypes.c_int64), _p(tg),
                       1 if include_self else 0, _p(pot), _p(field))
    return pot, field


def direct2d(src_xy, src_q, tgt_index=None, tgt_xy=None, include_self=True):
    src_xy = _f64(src_xy).reshape(-1, 2)
    src_q = _f64(src_q)
    if tgt_index is not None:
        ti = np.ascontiguousarray(tgt_index, dtype=np.int64)
        nt, tg = len(ti), None
    else:
        ti = None
        tg = _f64(tgt_xy).reshape(-1, 2)
        nt = len(tg)
    pot = np.zeros(nt)
    field = np.zeros((nt, 2))
    lib().orc_direct2d(len(src_q), _p(src_xy), _p(src_q), nt, _p(ti, ctypes.c_int64), _p(tg),
                       1 if include_self else 0, _p(pot), _p(field))
    return pot, field


def read_error(exact, res):
    """The reference ReadError statistics: (max |e|, rms relative, sum|e|/sum|x|)."""
    exact = _f64(exact)
    res = _f64(res)
    dim = 1 if exact.ndim == 1 else exact.shape[1]
    out = np.zeros(3)
    lib().orc_read_error(exact.shape[0], dim, _p(exact), _p(res), _p(out))
    return tuple(out)


class exact_geometry:
    """Context manager: DecToSph without the reference's +1e-12 radius offset."""

    def __enter__(self):
        self._old = lib().orc_get_r_eps()
        lib().orc_set_r_eps(0.0)
        return self

    def __exit__(self, *a):
        lib().orc_set_r_eps(self._old)


def num_threads():
    return lib().orc_num_threads()


class threads:
    """Context: the oracle's parallel loops on n threads (1: every sum in a
    fixed order, bit-reproducible results)."""

    def __init__(self, n):
        self.n = n

    def __enter__(self):
        self._old = lib().orc_num_threads()
        lib().orc_set_num_threads(int(self.n))
        return self

    def __exit__(self, *a):
        lib().orc_set_num_threads(self._old)
