"""ctypes binding of the CPU oracle (test infrastructure only).

Imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg,
never by the viennals_amd package.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# LSO_ORACLE_LIB: another build of the oracle (the pin tests load builds
# with planted mistakes through it)
_LIB_PATH = os.environ.get("LSO_ORACLE_LIB") or os.path.join(_HERE, "_build", "liblsoracle.so")
_lib = None

POS = np.finfo(np.float64).max
NEG = -POS


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


class Vel(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int), ("dir", ctypes.c_int), ("k", ctypes.c_int),
        ("b", ctypes.c_double), ("aPos", ctypes.c_double), ("aNeg", ctypes.c_double),
        ("vec", ctypes.c_double * 3),
        ("alpha", ctypes.c_double * 3),
    ]


class Params(ctypes.Structure):
    _fields_ = [
        ("spatial", ctypes.c_int), ("temporal", ctypes.c_int),
        ("cfl", ctypes.c_double), ("advectionTime", ctypes.c_double),
        ("singleStep", ctypes.c_int), ("dissipationAlpha", ctypes.c_double),
        ("calcNormals", ctypes.c_int), ("checkDissipation", ctypes.c_int),
        ("velKind", ctypes.c_int), ("velDir", ctypes.c_int), ("velK", ctypes.c_int),
        ("velB", ctypes.c_double), ("velAPos", ctypes.c_double), ("velANeg", ctypes.c_double),
        ("vec", ctypes.c_double * 3),
        ("mlp", ctypes.c_void_p),
        ("nMatVel", ctypes.c_int),
        ("matVel", Vel * 8),
        ("adaptive", ctypes.c_int),
        ("subdivisions", ctypes.c_int),
        ("velAlpha", ctypes.c_double * 3),
        ("updatePointData", ctypes.c_int),
        ("saveVelocities", ctypes.c_int),
        ("ignoreVoids", ctypes.c_int),
        ("velUpdate", ctypes.c_void_p),
        ("velUpdateUser", ctypes.c_void_p),
    ]


VELOCITY_UPDATE_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(Params), ctypes.c_void_p)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.c_void_p
        i32p = ctypes.POINTER(ctypes.c_int32)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        L.lso_domain_new.restype = P
        L.lso_domain_new.argtypes = [ctypes.c_int, i32p, i32p, ip, ctypes.c_double]
        L.lso_domain_free.argtypes = [P]
        L.lso_domain_copy.restype = P
        L.lso_domain_copy.argtypes = [P]
        L.lso_domain_size.restype = ctypes.c_int64
        L.lso_domain_size.argtypes = [P]
        L.lso_domain_width.restype = ctypes.c_int
        L.lso_domain_width.argtypes = [P]
        L.lso_domain_dim.restype = ctypes.c_int
        L.lso_domain_dim.argtypes = [P]
        L.lso_domain_get.argtypes = [P, i32p, dp]
        L.lso_domain_set.argtypes = [P, ctypes.c_int64, i32p, dp, ctypes.c_int]
        L.lso_domain_from_points.argtypes = [P, ctypes.c_int64, i32p, dp, ctypes.c_int]
        L.lso_make_sphere.argtypes = [P, dp, ctypes.c_double]
        L.lso_make_plane.argtypes = [P, dp, dp]
        L.lso_make_plane.restype = ctypes.c_int
        L.lso_make_box.argtypes = [P, dp, dp, ctypes.c_int]
        L.lso_expand.argtypes = [P, ctypes.c_int]
        L.lso_reduce.argtypes = [P, ctypes.c_int]
        L.lso_prune.argtypes = [P, ctypes.c_int]
        L.lso_rebuild.argtypes = [P]
        L.lso_boolean.argtypes = [P, P, ctypes.c_int]
        L.lso_boolean.restype = ctypes.c_int
        L.lso_combine.argtypes = [P, P, ctypes.c_double, ctypes.c_double]
        L.lso_check.argtypes = [P, ctypes.c_char_p, ctypes.c_int]
        L.lso_check.restype = ctypes.c_int
        L.lso_compare_volume.argtypes = [P, P]
        L.lso_compare_volume.restype = ctypes.c_uint64
        L.lso_weno5.argtypes = [dp, ctypes.c_double, ctypes.c_int]
        L.lso_weno5.restype = ctypes.c_double
        L.lso_weno3.argtypes = [dp, ctypes.c_double, ctypes.c_int]
        L.lso_weno3.restype = ctypes.c_double
        L.lso_advect.argtypes = [P, ctypes.POINTER(Params), dp]
        L.lso_advect.restype = ctypes.c_int
        L.lso_advect_multi.argtypes = [ctypes.POINTER(P), ctypes.c_int, ctypes.POINTER(Params), dp]
        L.lso_advect_multi.restype = ctypes.c_int
        L.lso_compute_rates.argtypes = [P, ctypes.POINTER(Params), ctypes.c_double, dp, ctypes.c_int64]
        L.lso_compute_rates.restype = ctypes.c_double
        L.lso_compute_rates_diss.argtypes = [P, ctypes.POINTER(Params), ctypes.c_double, dp, dp, ctypes.c_int64]
        L.lso_compute_rates_diss.restype = ctypes.c_double
        fp = ctypes.POINTER(ctypes.c_float)
        L.lso_mlp_forward.argtypes = [fp, fp, fp]
        L.lso_weno5_learned.argtypes = [dp, ctypes.c_double, ctypes.c_int, fp]
        L.lso_weno5_learned.restype = ctypes.c_double
        L.lso_mlp_loss_grad.argtypes = [fp, ctypes.c_int64, fp, fp, dp, ctypes.c_int, ctypes.c_int]
        L.lso_mlp_loss_grad.restype = ctypes.c_double
        L.lso_bf16_round.argtypes = [ctypes.c_float]
        L.lso_bf16_round.restype = ctypes.c_float
        L.lso_pd_insert.argtypes = [P, ctypes.c_int, ctypes.c_char_p, dp, ctypes.c_int64]
        L.lso_pd_count.argtypes = [P, ctypes.c_int]
        L.lso_pd_count.restype = ctypes.c_int
        L.lso_pd_label.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
        L.lso_pd_label.restype = ctypes.c_int
        L.lso_pd_size.argtypes = [P, ctypes.c_int, ctypes.c_int]
        L.lso_pd_size.restype = ctypes.c_int64
        L.lso_pd_get.argtypes = [P, ctypes.c_int, ctypes.c_int, dp]
        L.lso_pd_clear.argtypes = [P]
        L.lso_mark_void_points.argtypes = [P]
        L.lso_num_threads.restype = ctypes.c_int
        L.lso_set_num_threads.argtypes = [ctypes.c_int]
        _lib = L
    return _lib


def _i32(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


def _f64(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Domain:
    """Oracle level set: canonical run-start list (see ls_oracle.hpp)."""

    def __init__(self, D, gmin, gmax, bc, delta, _handle=None):
        L = lib()
        self.D = D
        self.gmin = list(gmin)
        self.gmax = list(gmax)
        self.bc = list(bc)
        self.delta = float(delta)
        if _handle is not None:
            self.h = _handle
        else:
            a, pa = _i32(gmin)
            b, pb = _i32(gmax)
            c = (ctypes.c_int * D)(*bc)
            self.h = L.lso_domain_new(D, pa, pb, c, float(delta))

    def __del__(self):
        if getattr(self, "h", None):
            lib().lso_domain_free(self.h)
            self.h = None

    def copy(self):
        return Domain(self.D, self.gmin, self.gmax, self.bc, self.delta, _handle=lib().lso_domain_copy(self.h))

    @property
    def width(self):
        return lib().lso_domain_width(self.h)

    def get(self):
        n = lib().lso_domain_size(self.h)
        idx = np.zeros((n, self.D), dtype=np.int32)
        val = np.zeros(n, dtype=np.float64)
        lib().lso_domain_get(self.h, idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                             val.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return idx, val

    def set(self, idx, val, width):
        a, pa = _i32(idx)
        b, pb = _f64(val)
        lib().lso_domain_set(self.h, len(b), pa, pb, int(width))

    def from_points(self, idx, val, width=2):
        a, pa = _i32(idx)
        b, pb = _f64(val)
        lib().lso_domain_from_points(self.h, len(b), pa, pb, int(width))

    def make_plane(self, origin, normal):
        """MakeGeometry with a Plane (lsMakeGeometry.hpp:245-360 through
        FromSurfaceMesh); returns 0 or the negative error code."""
        o, po_ = _f64(list(origin) + [0.0] * (3 - len(origin)))
        n, pn = _f64(list(normal) + [0.0] * (3 - len(normal)))
        return lib().lso_make_plane(self.h, po_, pn)

    def make_box(self, min_corner, max_corner, ignore_mask=0):
        """MakeGeometry with a Box (lsMakeGeometry.hpp:363-419 through
        FromSurfaceMesh); ignore_mask: setIgnoreBoundaryConditions per axis."""
        a, pa = _f64(list(min_corner) + [0.0] * (3 - len(min_corner)))
        b, pb = _f64(list(max_corner) + [0.0] * (3 - len(max_corner)))
        lib().lso_make_box(self.h, pa, pb, int(ignore_mask))

    def make_sphere(self, origin, radius):
        o, po = _f64(list(origin) + [0.0] * (3 - len(origin)))
        lib().lso_make_sphere(self.h, po, float(radius))

    def expand(self, w):
        lib().lso_expand(self.h, int(w))

    def reduce(self, w):
        lib().lso_reduce(self.h, int(w))

    def prune(self, remove_stray_zeros=False):
        lib().lso_prune(self.h, int(bool(remove_stray_zeros)))

    def rebuild(self):
        lib().lso_rebuild(self.h)

    def boolean(self, other, op):
        return lib().lso_boolean(self.h, other.h, int(op))

    def combine(self, other, w_target, w_source):
        lib().lso_combine(self.h, other.h, float(w_target), float(w_source))

    def check(self):
        buf = ctypes.create_string_buffer(4096)
        n = lib().lso_check(self.h, buf, 4096)
        return n, buf.value.decode()

    def defined(self):
        idx, val = self.get()
        m = np.abs(val) != POS
        return idx[m], val[m]

    # point data: kind 0 = scalar arrays, 1 = vector arrays (n x 3)
    def insert_data(self, label, values, vector=False):
        a = np.ascontiguousarray(values, dtype=np.float64)
        n = len(a)
        a = a.reshape(-1)
        lib().lso_pd_insert(self.h, int(vector), label.encode(), a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n)

    def data_labels(self, vector=False):
        k = int(vector)
        out = []
        for i in range(lib().lso_pd_count(self.h, k)):
            buf = ctypes.create_string_buffer(256)
            lib().lso_pd_label(self.h, k, i, buf, 256)
            out.append(buf.value.decode())
        return out

    def data(self, label, vector=False):
        k = int(vector)
        i = self.data_labels(vector).index(label)
        n = lib().lso_pd_size(self.h, k, i)
        a = np.zeros(n * (3 if vector else 1), dtype=np.float64)
        lib().lso_pd_get(self.h, k, i, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return a.reshape(n, 3) if vector else a

    def clear_data(self):
        lib().lso_pd_clear(self.h)

    def mark_void_points(self):
        lib().lso_mark_void_points(self.h)


class _View:
    """Read access to a level set the oracle is advecting (not owned)."""

    def __init__(self, h):
        self.h = h

    def width(self):
        return lib().lso_domain_width(self.h)

    def get(self):
        D = lib().lso_domain_dim(self.h)
        n = lib().lso_domain_size(self.h)
        idx = np.zeros((n, D), dtype=np.int32)
        val = np.zeros(n, dtype=np.float64)
        lib().lso_domain_get(self.h, idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                             val.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return idx, val


def make_params(spatial=0, temporal=0, cfl=0.4999, time=0.0, single_step=False,
                dissipation_alpha=1.0, calc_normals=True, check_dissipation=True,
                vel=None, mlp=None, mat_vels=None, adaptive=False, subdivisions=20,
                update_point_data=True, save_velocities=False, ignore_voids=False, velocity_update=None):
    """vel: dict(kind, dir, k, b, a_pos, a_neg, vec) of the built-in velocity.
    mlp: float32 array of the learned-weno5 network (spatial scheme 12).
    mat_vels: list of such dicts, the velocity of materials 0, 1, ...
    (multi-material advection; other materials use vel).
    adaptive, subdivisions: setAdaptiveTimeStepping.
    velocity_update: callback(top, params) between RK stages
    (setVelocityUpdateCallback); top is a view of the advected level set,
    params the parameters whose velocity fields (velB, matVel, ...) it may
    change for the rest of the apply."""
    p = Params()
    if velocity_update is not None:
        def trampoline(h, params, user):
            velocity_update(_View(h), params.contents)
            return 0
        p._velUpdateKeep = VELOCITY_UPDATE_FN(trampoline)
        p.velUpdate = ctypes.cast(p._velUpdateKeep, ctypes.c_void_p)
    p.updatePointData = int(bool(update_point_data))
    p.saveVelocities = int(bool(save_velocities))
    p.ignoreVoids = int(bool(ignore_voids))
    p.adaptive = int(bool(adaptive))
    p.subdivisions = int(subdivisions)
    if mlp is not None:
        p._mlp_keep = np.ascontiguousarray(mlp, dtype=np.float32)
        p.mlp = p._mlp_keep.ctypes.data
    p.spatial = spatial
    p.temporal = temporal
    p.cfl = cfl
    p.advectionTime = time
    p.singleStep = int(single_step)
    p.dissipationAlpha = dissipation_alpha
    p.calcNormals = int(calc_normals)
    p.checkDissipation = int(check_dissipation)
    vel = vel or {}
    p.velKind = vel.get("kind", 0)
    p.velDir = vel.get("dir", 2)
    p.velK = vel.get("k", 1)
    p.velB = vel.get("b", 0.0)
    p.velAPos = vel.get("a_pos", 0.0)
    p.velANeg = vel.get("a_neg", 0.0)
    v = vel.get("vec", (0.0, 0.0, 0.0))
    for i in range(3):
        p.vec[i] = v[i]
    al = vel.get("alpha", (0.0, 0.0, 0.0))
    for i in range(3):
        p.velAlpha[i] = al[i]
    mat_vels = mat_vels or []
    if len(mat_vels) > 8:
        raise ValueError("at most 8 material velocities")
    p.nMatVel = len(mat_vels)
    for m, mv in enumerate(mat_vels):
        t = p.matVel[m]
        t.kind, t.dir, t.k = mv.get("kind", 0), mv.get("dir", 2), mv.get("k", 1)
        t.b, t.aPos, t.aNeg = mv.get("b", 0.0), mv.get("a_pos", 0.0), mv.get("a_neg", 0.0)
        vv = mv.get("vec", (0.0, 0.0, 0.0))
        al = mv.get("alpha", (0.0, 0.0, 0.0))
        for i in range(3):
            t.vec[i] = vv[i]
            t.alpha[i] = al[i]
    return p


def advect(dom, params):
    out = (ctypes.c_double * 4)()
    st = lib().lso_advect(dom.h, ctypes.byref(params), out)
    return st, dict(advected_time=out[0], time_step=out[1], steps=int(out[2]), point_updates=out[3])


def advect_multi(doms, params):
    """Advect::apply with several level sets (doms[0] bottom, doms[-1] the
    advected top)."""
    arr = (ctypes.c_void_p * len(doms))(*[d.h for d in doms])
    out = (ctypes.c_double * 4)()
    st = lib().lso_advect_multi(arr, len(doms), ctypes.byref(params), out)
    return st, dict(advected_time=out[0], time_step=out[1], steps=int(out[2]), point_updates=out[3])


def compute_rates(dom, params, max_dt=np.finfo(np.float64).max):
    cap = lib().lso_domain_size(dom.h) * 32 + 4096
    buf = np.zeros(cap, dtype=np.float64)
    dt = lib().lso_compute_rates(dom.h, ctypes.byref(params), float(max_dt),
                                 buf.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), cap)
    n = lib().lso_domain_size(dom.h)
    return dt, buf[:n]


def compute_rates_diss(dom, params, max_dt=np.finfo(np.float64).max):
    """(dt, rates, dissipation) of one prepareLS + integrateTime pass."""
    cap = lib().lso_domain_size(dom.h) * 32 + 4096
    r = np.zeros(cap, dtype=np.float64)
    q = np.zeros(cap, dtype=np.float64)
    dpp = ctypes.POINTER(ctypes.c_double)
    dt = lib().lso_compute_rates_diss(dom.h, ctypes.byref(params), float(max_dt), r.ctypes.data_as(dpp),
                                      q.ctypes.data_as(dpp), cap)
    n = lib().lso_domain_size(dom.h)
    return dt, r[:n], q[:n]


def weno5(x, dx, plus):
    a, pa = _f64(x)
    return lib().lso_weno5(pa, float(dx), int(plus))


MLP_NPARAM = 1347
MLP_OFFSETS = dict(W1=(0, (32, 5)), b1=(160, (32,)), W2=(192, (32, 32)), b2=(1216, (32,)),
                   W3=(1248, (3, 32)), b3=(1344, (3,)))


def _f32(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def mlp_forward(params, x):
    P, pp = _f32(params)
    X, px = _f32(x)
    w = np.zeros(3, np.float32)
    lib().lso_mlp_forward(pp, px, w.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
    return w


def weno5_learned(x, dx, plus, params):
    a, pa = _f64(x)
    P, pp = _f32(params)
    return lib().lso_weno5_learned(pa, float(dx), int(plus), pp)


def mlp_loss_grad(params, X, T, emulate=False, use_double=False):
    """Loss and gradient (float64 array of 1347) over samples X (n x 8), T (n)."""
    P, pp = _f32(params)
    Xa, px = _f32(X)
    Ta, pt = _f32(T)
    g = np.zeros(MLP_NPARAM, np.float64)
    loss = lib().lso_mlp_loss_grad(pp, len(Ta), px, pt, g.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                   int(emulate), int(use_double))
    return loss, g


def bf16_round(x):
    return np.array([lib().lso_bf16_round(float(v)) for v in np.ravel(x)], np.float32).reshape(np.shape(x))


def weno3(x, dx, plus):
    a, pa = _f64(x)
    return lib().lso_weno3(pa, float(dx), int(plus))


def set_threads(n):
    lib().lso_set_num_threads(int(n))


def num_threads():
    return lib().lso_num_threads()


def compare_volume(target, sample):
    """CompareVolume (lsCompareVolume.hpp:176-289): (cells that differ,
    mismatch volume = cells * gridDelta^D)."""
    n = int(lib().lso_compare_volume(target.h, sample.h))
    return n, n * target.delta ** target.D
