"""ctypes binding of the host build of the GPU per-item logic (test only)."""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "_build", "liblsgemul.so")
_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        build()  # make is incremental: rebuilds after a change to the device headers
        L = ctypes.CDLL(_LIB)
        P = ctypes.c_void_p
        i32p = ctypes.POINTER(ctypes.c_int32)
        dp = ctypes.POINTER(ctypes.c_double)
        L.emu_domain_new.restype = P
        L.emu_domain_new.argtypes = [ctypes.c_int, i32p, i32p, ctypes.POINTER(ctypes.c_int), ctypes.c_double]
        L.emu_domain_free.argtypes = [P]
        L.emu_domain_set.argtypes = [P, ctypes.c_int64, i32p, dp, ctypes.c_int]
        L.emu_domain_size.restype = ctypes.c_int64
        L.emu_domain_size.argtypes = [P]
        L.emu_domain_width.argtypes = [P]
        L.emu_domain_get.argtypes = [P, i32p, dp]
        for f in ("emu_expand", "emu_reduce", "emu_prune"):
            getattr(L, f).argtypes = [P, ctypes.c_int]
        L.emu_rebuild.argtypes = [P]
        L.emu_boolean.argtypes = [P, P, ctypes.c_int, ctypes.c_double, ctypes.c_double]
        L.emu_advect.argtypes = [P, ctypes.c_void_p, dp]
        u32p = ctypes.POINTER(ctypes.c_uint32)
        L.emu_fastdiv.argtypes = [ctypes.c_uint32, u32p, u32p, ctypes.c_int64]
        u64p = ctypes.POINTER(ctypes.c_uint64)
        L.emu_deposit.argtypes = [u64p, u64p, u64p, ctypes.c_int64]
        L.emu_select.argtypes = [u64p, i32p, i32p, ctypes.c_int64]
        L.emu_set_tiled.argtypes = [ctypes.c_int]
        L.emu_set_expand_pairs.argtypes = [ctypes.c_int]
        L.emu_geometry.restype = ctypes.c_int64
        L.emu_geometry.argtypes = [ctypes.c_int, ctypes.c_int, i32p, i32p, ctypes.POINTER(ctypes.c_int),
                                   ctypes.c_double, dp, dp, ctypes.c_int, i32p, dp, ctypes.c_int64]
        _lib = L
    return _lib


class Domain:
    def __init__(self, D, gmin, gmax, bc, delta):
        self.D = D
        g0 = np.ascontiguousarray(gmin, dtype=np.int32)
        g1 = np.ascontiguousarray(gmax, dtype=np.int32)
        c = (ctypes.c_int * D)(*bc)
        self.h = lib().emu_domain_new(D, g0.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                      g1.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), c, float(delta))

    def __del__(self):
        if getattr(self, "h", None):
            lib().emu_domain_free(self.h)
            self.h = None

    def set(self, idx, val, width):
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        val = np.ascontiguousarray(val, dtype=np.float64)
        r = lib().emu_domain_set(self.h, len(val), idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                 val.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), int(width))
        assert r == 0

    def get(self):
        n = lib().emu_domain_size(self.h)
        idx = np.zeros((n, self.D), dtype=np.int32)
        val = np.zeros(n, dtype=np.float64)
        lib().emu_domain_get(self.h, idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                             val.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return idx, val

    @property
    def width(self):
        return lib().emu_domain_width(self.h)

    def expand(self, w):
        assert lib().emu_expand(self.h, int(w)) == 0

    def reduce(self, w):
        assert lib().emu_reduce(self.h, int(w)) == 0

    def prune(self, rsz=False):
        assert lib().emu_prune(self.h, int(bool(rsz))) == 0

    def rebuild(self):
        assert lib().emu_rebuild(self.h) == 0

    def boolean(self, other, kind, wt=0.0, ws=0.0):
        assert lib().emu_boolean(self.h, other.h, int(kind), float(wt), float(ws)) == 0


def advect(dom, params):
    """params: the oracle's Params structure (same layout)."""
    out = (ctypes.c_double * 4)()
    st = lib().emu_advect(dom.h, ctypes.byref(params), out)
    return st, dict(advected_time=out[0], steps=int(out[2]), point_updates=out[3])


def fastdiv(d, n):
    """The frame's run-time-constant division (lsg_device.hpp FastDiv)."""
    n = np.ascontiguousarray(n, dtype=np.uint32)
    q = np.zeros_like(n)
    u32p = ctypes.POINTER(ctypes.c_uint32)
    lib().emu_fastdiv(int(d), n.ctypes.data_as(u32p), q.ctypes.data_as(u32p), len(n))
    return q


def deposit(x, m):
    x = np.ascontiguousarray(x, dtype=np.uint64)
    m = np.ascontiguousarray(m, dtype=np.uint64)
    out = np.zeros_like(x)
    u64p = ctypes.POINTER(ctypes.c_uint64)
    lib().emu_deposit(x.ctypes.data_as(u64p), m.ctypes.data_as(u64p), out.ctypes.data_as(u64p), len(x))
    return out


def select_bit(w, k):
    w = np.ascontiguousarray(w, dtype=np.uint64)
    k = np.ascontiguousarray(k, dtype=np.int32)
    out = np.zeros_like(k)
    i32p = ctypes.POINTER(ctypes.c_int32)
    lib().emu_select(w.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), k.ctypes.data_as(i32p),
                     out.ctypes.data_as(i32p), len(w))
    return out


def set_expand_pairs(on):
    """Expand two cycles per band operation and fuse each stage's update
    with its rebuild (candidate-list pipeline)."""
    lib().emu_set_expand_pairs(int(bool(on)))


def set_tiled(on):
    """Band operations through the single-pass tile pipeline (True) or the
    candidate-list pipeline (False)."""
    lib().emu_set_tiled(int(bool(on)))


def geometry(kind, D, gmin, gmax, bc, dx, a, b, mask=0):
    """The library's MakeGeometry plane (kind 0) or box (kind 1) on the
    host: (idx, val) of the defined points, or the negative error code."""
    g0 = np.ascontiguousarray(gmin, dtype=np.int32)
    g1 = np.ascontiguousarray(gmax, dtype=np.int32)
    c = (ctypes.c_int * D)(*bc)
    pa = np.ascontiguousarray(list(a) + [0.0] * (3 - len(a)), dtype=np.float64)
    pb = np.ascontiguousarray(list(b) + [0.0] * (3 - len(b)), dtype=np.float64)
    ip = ctypes.POINTER(ctypes.c_int32)
    dp = ctypes.POINTER(ctypes.c_double)
    args = (int(kind), D, g0.ctypes.data_as(ip), g1.ctypes.data_as(ip), c, float(dx), pa.ctypes.data_as(dp),
            pb.ctypes.data_as(dp), int(mask))
    n = lib().emu_geometry(*args, None, None, 0)
    if n < 0:
        return n
    idx = np.zeros((n, D), dtype=np.int32)
    val = np.zeros(n, dtype=np.float64)
    lib().emu_geometry(*args, idx.ctypes.data_as(ip), val.ctypes.data_as(dp), n)
    return idx, val
