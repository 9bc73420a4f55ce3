"""MI355X-native level-set advection with the ViennaLS Advect API.

Python host layer over the C ABI of libviennals_gpu.so, mirroring the
reference's Python bindings (viennals.d2 / viennals.d3: Domain, MakeGeometry,
Sphere, Plane, Box, Expand, Reduce, Prune, BooleanOperation, Advect;
python/pyWrap.hpp:324-401).  All level-set data lives in GPU memory.
"""
import ctypes
import warnings

import numpy as np

from ._lib import LsgError, check, lib

__all__ = [
    "BoundaryConditionEnum", "SpatialSchemeEnum", "TemporalSchemeEnum", "BooleanOperationEnum",
    "Domain", "Sphere", "Plane", "Box", "MakeGeometry", "Expand", "Reduce", "Prune", "BooleanOperation",
    "VelocityField", "HostVelocityField", "Advect", "BatchAdvect", "LsgError", "weno", "PointData",
    "MarkVoidPoints", "CompareChamfer", "surface_points", "PrepareStencilLocalLaxFriedrichs",
    "FinalizeStencilLocalLaxFriedrichs", "Writer", "Reader",
]

POS_VALUE = float(np.finfo(np.float64).max)
NEG_VALUE = -POS_VALUE


class BoundaryConditionEnum:
    """viennahrle::BoundaryType (python/viennals/_core.pyi:26-30)."""
    REFLECTIVE_BOUNDARY = 0
    INFINITE_BOUNDARY = 1
    PERIODIC_BOUNDARY = 2
    POS_INFINITE_BOUNDARY = 3
    NEG_INFINITE_BOUNDARY = 4


class SpatialSchemeEnum:
    """lsAdvectIntegrationSchemes.hpp:10-23."""
    ENGQUIST_OSHER_1ST_ORDER = 0
    ENGQUIST_OSHER_2ND_ORDER = 1
    LAX_FRIEDRICHS_1ST_ORDER = 2
    LAX_FRIEDRICHS_2ND_ORDER = 3
    LOCAL_LAX_FRIEDRICHS_ANALYTICAL_1ST_ORDER = 4
    LOCAL_LOCAL_LAX_FRIEDRICHS_1ST_ORDER = 5
    LOCAL_LOCAL_LAX_FRIEDRICHS_2ND_ORDER = 6
    LOCAL_LAX_FRIEDRICHS_1ST_ORDER = 7
    LOCAL_LAX_FRIEDRICHS_2ND_ORDER = 8
    STENCIL_LOCAL_LAX_FRIEDRICHS_1ST_ORDER = 9
    WENO_3RD_ORDER = 10
    WENO_5TH_ORDER = 11
    WENO_5TH_ORDER_LEARNED = 12  # extension: learned weno5 weights (viennals_amd.learned)


class TemporalSchemeEnum:
    """lsAdvectIntegrationSchemes.hpp:31-35."""
    FORWARD_EULER = 0
    RUNGE_KUTTA_2ND_ORDER = 1
    RUNGE_KUTTA_3RD_ORDER = 2


class BooleanOperationEnum:
    """lsBooleanOperation.hpp:27-33 (CUSTOM is internal here)."""
    INTERSECT = 0
    UNION = 1
    RELATIVE_COMPLEMENT = 2
    INVERT = 3


def _i32(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


def _f64(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class PointData:
    """Point data of a Domain (ViennaCore PointData as bound in
    python/viennals/_core.pyi:819-845): named scalar arrays (one value per
    defined point, point-id order) and vector arrays (three values per
    point), kept on the device."""

    def __init__(self, domain):
        self._d = domain

    def _insert(self, values, label, comps, replace):
        a = np.ascontiguousarray(values, dtype=np.float64)
        if comps == 3:
            a = a.reshape(-1, 3)
        n = a.shape[0] if a.ndim else 0
        check(lib().lsg_domain_point_data_insert(self._d._h, label.encode(), comps, a.ctypes.data, n, int(replace)))

    def insertNextScalarData(self, scalars, label="Scalars"):
        self._insert(scalars, label, 1, False)

    def insertNextVectorData(self, vectors, label="Vectors"):
        self._insert(vectors, label, 3, False)

    def insertReplaceScalarData(self, scalars, label="Scalars"):
        self._insert(scalars, label, 1, True)

    def insertReplaceVectorData(self, vectors, label="Vectors"):
        self._insert(vectors, label, 3, True)

    def getScalarDataSize(self):
        return lib().lsg_domain_point_data_count(self._d._h, 1)

    def getVectorDataSize(self):
        return lib().lsg_domain_point_data_count(self._d._h, 3)

    def _label(self, comps, i):
        buf = ctypes.create_string_buffer(1024)
        n = lib().lsg_domain_point_data_label(self._d._h, comps, int(i), buf, 1024)
        if n < 0:
            check(n)
        return buf.value.decode()

    def getScalarDataLabel(self, i):
        return self._label(1, i)

    def getVectorDataLabel(self, i):
        return self._label(3, i)

    def _get(self, comps, key, noWarning):
        count = lib().lsg_domain_point_data_count(self._d._h, comps)
        if isinstance(key, str):
            idx = [i for i in range(count) if self._label(comps, i) == key]
            if not idx:
                if not noWarning:
                    import warnings
                    warnings.warn(f"PointData: no data with label {key!r}")
                return None
            i = idx[0]
        else:
            i = int(key)
        n = lib().lsg_domain_point_data_size(self._d._h, comps, i)
        if n < 0:
            check(int(n))
        out = np.zeros(n * comps, dtype=np.float64)
        check(lib().lsg_domain_point_data_get(self._d._h, comps, i, out.ctypes.data, n))
        return out.reshape(n, 3) if comps == 3 else out

    def getScalarData(self, key, noWarning=False):
        return self._get(1, key, noWarning)

    def getVectorData(self, key, noWarning=False):
        return self._get(3, key, noWarning)

    def clear(self):
        check(lib().lsg_domain_point_data_clear(self._d._h))

    def empty(self):
        return self.getScalarDataSize() == 0 and self.getVectorDataSize() == 0


class Domain:
    """Device-resident level set (Domain<T,D>, lsDomain.hpp:27-269).

    Domain(bounds, boundaryConditions, gridDelta) with bounds
    [x0, x1, y0, y1(, z0, z1)] in physical units (lsDomain.hpp:68-78), or
    Domain(gridDelta, dim=3) for an all-infinite domain (lsDomain.hpp:55-66),
    or Domain(other) for a deep copy of another level set (lsDomain.hpp:108).
    """

    def __init__(self, bounds=None, boundaryConditions=None, gridDelta=1.0, dim=None):
        if isinstance(bounds, Domain):
            src = bounds
            Domain.__init__(self, None, None, src.gridDelta, dim=src.D)
            self.gmin, self.gmax, self.bc = list(src.gmin), list(src.gmax), list(src.bc)
            self.deepCopy(src)
            return
        if bounds is None:
            D = dim or 3
            gmin = [0] * D
            gmax = [0] * D
            bc = [BoundaryConditionEnum.INFINITE_BOUNDARY] * D
        else:
            D = len(bounds) // 2
            gmin = [int(np.floor(bounds[2 * i] / gridDelta)) for i in range(D)]
            gmax = [int(np.ceil(bounds[2 * i + 1] / gridDelta)) for i in range(D)]
            bc = list(boundaryConditions) if boundaryConditions is not None else \
                [BoundaryConditionEnum.INFINITE_BOUNDARY] * D
        self.D = D
        self.gridDelta = float(gridDelta)
        self.gmin, self.gmax, self.bc = gmin, gmax, [int(b) for b in bc]
        a, pa = _i32(gmin)
        b, pb = _i32(gmax)
        c = (ctypes.c_int * D)(*self.bc)
        h = ctypes.c_void_p()
        check(lib().lsg_domain_create(D, pa, pb, c, self.gridDelta, ctypes.byref(h)))
        self._h = h

    @classmethod
    def from_index_bounds(cls, D, gmin, gmax, bc, gridDelta):
        """Domain from index bounds (floor(b0/dx), ceil(b1/dx)) directly."""
        obj = cls.__new__(cls)
        obj.D = D
        obj.gridDelta = float(gridDelta)
        obj.gmin, obj.gmax, obj.bc = list(gmin), list(gmax), [int(x) for x in bc]
        a, pa = _i32(gmin)
        b, pb = _i32(gmax)
        c = (ctypes.c_int * D)(*obj.bc)
        h = ctypes.c_void_p()
        check(lib().lsg_domain_create(D, pa, pb, c, obj.gridDelta, ctypes.byref(h)))
        obj._h = h
        return obj

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                lib().lsg_domain_destroy(h)
            except Exception:
                pass
            self._h = None

    # Domain API
    def deepCopy(self, other):
        check(lib().lsg_domain_deep_copy(self._h, other._h))

    def insertPoints(self, indices, values, width=2):
        """Domain::insertPoints: defined points (index, value) (lsDomain.hpp:135-138)."""
        a, pa = _i32(indices)
        b, pb = _f64(values)
        check(lib().lsg_domain_insert_points(self._h, len(b), pa, pb, int(width)))

    def getNumberOfPoints(self):
        n = ctypes.c_int64()
        check(lib().lsg_domain_num_points(self._h, ctypes.byref(n)))
        return n.value

    def getLevelSetWidth(self):
        w = ctypes.c_int()
        check(lib().lsg_domain_width(self._h, ctypes.byref(w)))
        return w.value

    def setLevelSetWidth(self, w):
        check(lib().lsg_domain_set_width(self._h, int(w)))

    def getNumberOfSegments(self):
        return 1  # one device band (the reference splits by OpenMP thread)

    def clearMetaData(self):
        pass  # no metadata is kept on the device band

    def getPointData(self):
        """Domain::getPointData (lsDomain.hpp:44): the device-resident
        arrays carried through every band operation and Advect."""
        return PointData(self)

    def print(self):
        print(f"Domain: D={self.D} gridDelta={self.gridDelta} index bounds {self.gmin}..{self.gmax} "
              f"bc {self.bc}; {self.getNumberOfPoints()} defined points, width {self.getLevelSetWidth()}")

    # run-start list exchange
    def set_runs(self, indices, values, width):
        a, pa = _i32(indices)
        b, pb = _f64(values)
        check(lib().lsg_domain_set_runs(self._h, len(b), pa, pb, int(width)))

    def get_runs(self):
        n = ctypes.c_int64()
        check(lib().lsg_domain_num_runs(self._h, ctypes.byref(n)))
        idx = np.zeros((n.value, self.D), dtype=np.int32)
        val = np.zeros(n.value, dtype=np.float64)
        check(lib().lsg_domain_get_runs(self._h, idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                        val.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n.value))
        return idx, val

    def serialize(self):
        """Domain::serialize (lsDomain.hpp:188-215): the level set as bytes."""
        n = ctypes.c_int64()
        check(lib().lsg_domain_serialize(self._h, None, 0, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value)
        check(lib().lsg_domain_serialize(self._h, buf, n.value, ctypes.byref(n)))
        return buf.raw[:n.value]

    def deserialize(self, data):
        """Domain::deserialize (lsDomain.hpp:218-268): grid, band, width and
        point data from bytes written by serialize."""
        data = bytes(data)
        check(lib().lsg_domain_deserialize(self._h, data, len(data)))
        self._read_grid()

    def _read_grid(self):
        D = ctypes.c_int()
        gmin = (ctypes.c_int32 * 3)()
        gmax = (ctypes.c_int32 * 3)()
        bc = (ctypes.c_int * 3)()
        dx = ctypes.c_double()
        check(lib().lsg_domain_get_grid(self._h, ctypes.byref(D), gmin, gmax, bc, ctypes.byref(dx)))
        self.D = D.value
        self.gmin = [gmin[k] for k in range(self.D)]
        self.gmax = [gmax[k] for k in range(self.D)]
        self.bc = [bc[k] for k in range(self.D)]
        self.gridDelta = dx.value

    def defined_points(self):
        idx, val = self.get_runs()
        m = np.abs(val) != POS_VALUE
        return idx[m], val[m]


class Sphere:
    def __init__(self, origin, radius):
        self.origin, self.radius = list(origin), float(radius)


class Plane:
    def __init__(self, origin, normal):
        self.origin, self.normal = list(origin), list(normal)


class Box:
    def __init__(self, minPoint, maxPoint):
        self.minPoint, self.maxPoint = list(minPoint), list(maxPoint)


class MakeGeometry:
    """MakeGeometry(domain, geometry).apply() (lsMakeGeometry.hpp); also
    MakeGeometry() / MakeGeometry(domain) with setLevelSet / setGeometry."""

    def __init__(self, domain=None, geometry=None):
        self.domain, self.geometry = domain, geometry

    def setLevelSet(self, domain):
        self.domain = domain

    def setGeometry(self, geometry):
        self.geometry = geometry

    def setIgnoreBoundaryConditions(self, ignore):
        """One flag for every dimension or a list of flags
        (lsMakeGeometry.hpp:125-141): a periodic dimension whose flag is set
        keeps a sphere's indices unmapped (:227-235) and drops a box's mesh
        elements on its boundary (:415-417); planes ignore it."""
        flags = list(ignore) if hasattr(ignore, "__len__") else [bool(ignore)] * 3
        self.ignoreMask = sum(1 << i for i, f in enumerate(flags[:3]) if f)

    def apply(self):
        g, d = self.geometry, self.domain
        if d is None:
            raise LsgError("No level set was passed to MakeGeometry.")
        pad = lambda v: list(v) + [0.0] * (3 - len(v))
        mask = getattr(self, "ignoreMask", 0)
        if isinstance(g, Sphere):
            a, pa = _f64(pad(g.origin))
            check(lib().lsg_make_sphere_ignore_bc(d._h, pa, g.radius, int(mask)))
        elif isinstance(g, Plane):
            a, pa = _f64(pad(g.origin))
            b, pb = _f64(pad(g.normal))
            check(lib().lsg_make_plane(d._h, pa, pb))
        elif isinstance(g, Box):
            a, pa = _f64(pad(g.minPoint))
            b, pb = _f64(pad(g.maxPoint))
            check(lib().lsg_make_box_ignore_bc(d._h, pa, pb, int(mask)))
        else:
            raise LsgError("Invalid geometry type was specified for MakeGeometry.")


def _lvst_name(fileName):
    # Writer::apply / Reader::apply append ".lvst" to other names
    # (lsWriter.hpp:46-50, lsReader.hpp:44-48)
    name = str(fileName)
    if not name.endswith(".lvst"):
        warnings.warn("File name does not end in '.lvst', appending it.")
        name += ".lvst"
    return name


class Writer:
    """Writer<T,D> (lsWriter.hpp:12-60): Domain::serialize into a .lvst file."""

    def __init__(self, levelSet=None, fileName=""):
        self.levelSet, self.fileName = levelSet, fileName

    def setLevelSet(self, levelSet):
        self.levelSet = levelSet

    def setFileName(self, fileName):
        self.fileName = fileName

    def apply(self):
        if self.levelSet is None:
            raise LsgError("No level-set was passed to Writer.")
        if not self.fileName:
            raise LsgError("No file name specified for Writer.")
        self.fileName = _lvst_name(self.fileName)
        with open(self.fileName, "wb") as f:
            f.write(self.levelSet.serialize())


class Reader:
    """Reader<T,D> (lsReader.hpp:12-60): Domain::deserialize from a .lvst file."""

    def __init__(self, levelSet=None, fileName=""):
        self.levelSet, self.fileName = levelSet, fileName

    def setLevelSet(self, levelSet):
        self.levelSet = levelSet

    def setFileName(self, fileName):
        self.fileName = fileName

    def apply(self):
        if self.levelSet is None:
            raise LsgError("No level-set was passed to Reader.")
        if not self.fileName:
            raise LsgError("No file name specified for Reader.")
        self.fileName = _lvst_name(self.fileName)
        with open(self.fileName, "rb") as f:
            self.levelSet.deserialize(f.read())


class Expand:
    """Expand(domain, width) (lsExpand.hpp)."""

    def __init__(self, domain=None, width=0):
        self.domain, self.width = domain, width

    def setLevelSet(self, domain):
        self.domain = domain

    def setWidth(self, w):
        self.width = w

    def apply(self):
        check(lib().lsg_expand(self.domain._h, int(self.width)))


class Reduce:
    """Reduce(domain, width, noNewSegment) (lsReduce.hpp); the GPU band has
    one segment, so noNewSegment has no effect."""

    def __init__(self, domain=None, width=0, noNewSegment=False):
        self.domain, self.width = domain, width

    def setLevelSet(self, domain):
        self.domain = domain

    def setWidth(self, w):
        self.width = w

    def setNoNewSegment(self, v):
        pass

    def apply(self):
        check(lib().lsg_reduce(self.domain._h, int(self.width)))


class Prune:
    """Prune(domain) (lsPrune.hpp)."""

    def __init__(self, domain=None):
        self.domain, self.rsz = domain, False

    def setLevelSet(self, domain):
        self.domain = domain

    def setRemoveStrayZeros(self, v):
        self.rsz = bool(v)

    def apply(self):
        check(lib().lsg_prune(self.domain._h, int(self.rsz)))


class MarkVoidPoints:
    """MarkVoidPoints(domain).apply() with the default options
    (lsMarkVoidPoints.hpp:156-331): point data "VoidPointMarkers"."""

    voidPointLabel = "VoidPointMarkers"

    def __init__(self, domain=None):
        self.domain = domain

    def setLevelSet(self, domain):
        self.domain = domain

    def apply(self):
        check(lib().lsg_mark_void_points(self.domain._h))


def surface_points(domain):
    """Zero crossings of a level set on grid edges between face-adjacent
    defined points (physical coordinates, n x D), extracted on the GPU."""
    n = ctypes.c_int64(0)
    check(lib().lsg_surface_points(domain._h, None, 0, ctypes.byref(n)))
    out = np.zeros((n.value, domain.D), dtype=np.float64)
    if n.value:
        check(lib().lsg_surface_points(domain._h, out.ctypes.data, n.value, ctypes.byref(n)))
    return out


class CompareChamfer:
    """CompareChamfer (lsCompareChamfer.hpp:104-330) on the GPU surface points
    (see surface_points): forward / backward / Chamfer / RMS / max distance."""

    def __init__(self, target=None, sample=None):
        self.target, self.sample = target, sample
        self._r = [float("nan")] * 5
        self._n = [0, 0]

    def setLevelSetTarget(self, d):
        self.target = d

    def setLevelSetSample(self, d):
        self.sample = d

    def apply(self):
        out = np.zeros(5)
        cnt = np.zeros(2, dtype=np.int64)
        check(lib().lsg_compare_chamfer(self.target._h, self.sample._h, out.ctypes.data, cnt.ctypes.data))
        self._r = [float(x) for x in out]
        self._n = [int(x) for x in cnt]

    def getForwardDistance(self):
        return self._r[0]

    def getBackwardDistance(self):
        return self._r[1]

    def getChamferDistance(self):
        return self._r[2]

    def getRMSChamferDistance(self):
        return self._r[3]

    def getMaxDistance(self):
        return self._r[4]

    def getNumTargetPoints(self):
        return self._n[0]

    def getNumSamplePoints(self):
        return self._n[1]


class CompareVolume:
    """CompareVolume (lsCompareVolume.hpp:176-289) on the GPU: the cells whose
    corner average is inside (<= 0) in exactly one of the two level sets
    (narrower level sets are expanded to width 3 on copies); both must share
    their grid.  CompareArea is the same class in 2D."""

    def __init__(self, target=None, sample=None):
        self.target, self.sample = target, sample
        self._cells = 0
        self._volume = float("nan")

    def setLevelSetTarget(self, d):
        self.target = d

    def setLevelSetSample(self, d):
        self.sample = d

    def apply(self):
        c = ctypes.c_uint64()
        v = ctypes.c_double()
        check(lib().lsg_compare_volume(self.target._h, self.sample._h, ctypes.byref(c), ctypes.byref(v)))
        self._cells, self._volume = int(c.value), float(v.value)

    def getVolumeMismatch(self):
        return self._volume

    def getAreaMismatch(self):
        return self._volume

    def getCellCount(self):
        return self._cells


CompareArea = CompareVolume


class BooleanOperation:
    """BooleanOperation(a, b, op) (lsBooleanOperation.hpp); also the
    reference's (), (a), (a, b), (a, op) forms and setters."""

    def __init__(self, a=None, b=None, op=BooleanOperationEnum.INTERSECT):
        if isinstance(b, (int, np.integer)) and not isinstance(b, Domain):
            b, op = None, b
        self.a, self.b, self.op = a, b, op

    def setLevelSet(self, a):
        self.a = a

    setLevelset = setLevelSet  # the reference's Python spelling

    def setSecondLevelSet(self, b):
        self.b = b

    def setBooleanOperation(self, op):
        self.op = op

    def apply(self):
        if int(self.op) == BooleanOperationEnum.INVERT:
            check(lib().lsg_boolean(self.a._h, None, 3))
            return
        if self.a is None or self.b is None:
            raise LsgError("BooleanOperation needs two level sets")
        check(lib().lsg_boolean(self.a._h, self.b._h, int(self.op)))


def _empty_like(d):
    """Domain::New(grid): the same grid, no defined points (width 1)."""
    e = type(d).from_index_bounds(d.D, d.gmin, d.gmax, d.bc, d.gridDelta)
    idx, _ = d.get_runs()  # the first entry is the grid's minimum point
    e.set_runs(idx[:1].astype(np.int32), np.array([POS_VALUE]), 1)
    return e


def _copy_of(d):
    c = type(d).from_index_bounds(d.D, d.gmin, d.gmax, d.bc, d.gridDelta)
    c.deepCopy(d)
    return c


def PrepareStencilLocalLaxFriedrichs(levelSets, isDepo):
    """The layer wrapping for the SLLF scheme
    (lsStencilLocalLaxFriedrichsScalar.hpp:423-479): the top level set
    becomes the inverted union of every deposition layer minus the masking
    layer below it.  The condition reads isDepo[isDepo[k]] for the k-th
    level set from the top, as the reference's loop does (:446-447)."""
    n = len(levelSets)
    isDepo = [bool(x) for x in isDepo]
    if len(isDepo) < n:
        import warnings
        warnings.warn("PrepareStencilLocalLaxFriedrichs: isDepo does not have enough elements. "
                      "Assuming all higher layers are not depo layers.")
    isDepo = (isDepo + [False] * n)[:n]
    layer = n - 1
    finalTop = _empty_like(levelSets[0])
    layerAboveIsDepo = False
    for k in range(n):
        mask = levelSets[n - 1 - k]
        if isDepo[int(isDepo[n - 1 - k])]:
            if not layerAboveIsDepo:
                layer = n - 1 - k
                layerAboveIsDepo = True
        else:
            if layerAboveIsDepo:
                layerAbove = _copy_of(levelSets[layer])
                BooleanOperation(layerAbove, mask, BooleanOperationEnum.RELATIVE_COMPLEMENT).apply()
                BooleanOperation(finalTop, layerAbove, BooleanOperationEnum.UNION).apply()
                Prune(finalTop).apply()
            layerAboveIsDepo = False
    if layerAboveIsDepo:
        BooleanOperation(finalTop, levelSets[layer], BooleanOperationEnum.UNION).apply()
    BooleanOperation(finalTop, None, BooleanOperationEnum.INVERT).apply()
    levelSets[-1].deepCopy(finalTop)


def FinalizeStencilLocalLaxFriedrichs(levelSets):
    """Restore the usual layer wrapping after SLLF advection
    (lsStencilLocalLaxFriedrichsScalar.hpp:483-492)."""
    BooleanOperation(levelSets[-1], None, BooleanOperationEnum.INVERT).apply()
    BooleanOperation(levelSets[-1], levelSets[-2], BooleanOperationEnum.UNION).apply()


class VelocityField:
    """Built-in GPU velocity functors replacing VelocityField<T>
    (lsVelocityField.hpp:11-41):

    scalar = b                                            (constant, kind 0)
           = b + (n[dir] > 0 ? a_pos : a_neg) * |n[dir]|^k  (normal power, kind 1)
           = b + a_pos * (|n_0| + |n_1| + |n_2|)           (normal sum, kind 2)
    vector = vec (constant)
    """

    def __init__(self, kind=0, dir=2, k=1, b=0.0, a_pos=0.0, a_neg=0.0, vec=(0.0, 0.0, 0.0), alpha=(0.0, 0.0, 0.0)):
        self.kind, self.dir, self.k = int(kind), int(dir), int(k)
        self.b, self.a_pos, self.a_neg = float(b), float(a_pos), float(a_neg)
        self.vec = tuple(float(v) for v in vec) + (0.0,) * (3 - len(vec))
        # getDissipationAlpha (LOCAL_LAX_FRIEDRICHS_ANALYTICAL): constant per direction
        self.alpha = tuple(float(v) for v in alpha) + (0.0,) * (3 - len(alpha))

    @classmethod
    def constant(cls, scalar=0.0, vector=(0.0, 0.0, 0.0)):
        return cls(kind=0, b=scalar, vec=vector)

    @classmethod
    def normal_power(cls, dir, k, b, a_pos, a_neg=0.0, vector=(0.0, 0.0, 0.0)):
        return cls(kind=1, dir=dir, k=k, b=b, a_pos=a_pos, a_neg=a_neg, vec=vector)

    @classmethod
    def normal_sum(cls, a, b=0.0, vector=(0.0, 0.0, 0.0)):
        return cls(kind=2, b=b, a_pos=a, vec=vector)

    def as_dict(self):
        return dict(kind=self.kind, dir=self.dir, k=self.k, b=self.b, a_pos=self.a_pos, a_neg=self.a_neg,
                    vec=self.vec, alpha=self.alpha)

    # VelocityField<T> interface (lsVelocityField.hpp:17-38).  A subclass
    # that overrides these is evaluated on the host for every active point
    # (like HostVelocityField); the functor itself runs on the GPU.
    def getScalarVelocity(self, coordinate, material, normal, pointId):
        if self.kind == 0:
            return self.b
        if self.kind == 2:
            s = 0.0
            for i in range(3):
                s += abs(normal[i])
            return self.b + self.a_pos * s
        c = normal[self.dir]
        m = 1.0
        if self.k >= 1:
            m = abs(c)
            for _ in range(1, self.k):
                m = m * abs(c)
        return self.b + (self.a_pos if c > 0 else self.a_neg) * m

    def getVectorVelocity(self, coordinate, material, normal, pointId):
        return tuple(self.vec)

    def getDissipationAlpha(self, direction, material, centralDifferences):
        return getattr(self, "alpha", (0.0, 0.0, 0.0))[direction]

    def _overridden(self):
        t = type(self)
        return (t.getScalarVelocity is not VelocityField.getScalarVelocity or
                t.getVectorVelocity is not VelocityField.getVectorVelocity)


class MaterialVelocityField:
    """A VelocityField whose built-in functor depends on the material
    (several level sets: the material of a point is the index of the level
    set below or at it, lsAdvect.hpp:513-527).  `per_material[m]` is the
    VelocityField of material m; materials beyond the list use `default`
    (or the last entry)."""

    def __init__(self, per_material, default=None):
        self.per_material = list(per_material)
        if not self.per_material:
            raise ValueError("MaterialVelocityField needs at least one functor")
        if len(self.per_material) > 8:
            raise ValueError("at most 8 materials")
        self.default = default if default is not None else self.per_material[-1]

    def as_dicts(self):
        return [v.as_dict() for v in self.per_material]


class HostVelocityField:
    """VelocityField<T> (lsVelocityField.hpp:11-41) implemented in Python:
    override getScalarVelocity / getVectorVelocity.  Evaluated on the host
    for every active point of every stage (the slow path; the built-in
    VelocityField functors run on the GPU)."""

    def getScalarVelocity(self, coordinate, material, normal, pointId):
        return 0.0

    def getVectorVelocity(self, coordinate, material, normal, pointId):
        return (0.0, 0.0, 0.0)

    def getDissipationAlpha(self, direction, material, centralDifferences):
        return 0.0


class Advect:
    """Advect<T,D> (lsAdvect.hpp:54-1044) running on the GPU."""

    def __init__(self, domain=None, velocities=None):
        h = ctypes.c_void_p()
        check(lib().lsg_advect_create(ctypes.byref(h)))
        self._h = h
        self._domains = []
        if domain is not None:
            self.insertNextLevelSet(domain)
        if velocities is not None:
            self.setVelocityField(velocities)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                lib().lsg_advect_destroy(h)
            except Exception:
                pass
            self._h = None

    def insertNextLevelSet(self, d):
        self._domains.append(d)
        check(lib().lsg_advect_insert_next_level_set(self._h, d._h))

    def clearLevelSets(self):
        self._domains = []
        check(lib().lsg_advect_clear_level_sets(self._h))

    def setVelocityField(self, v):
        if isinstance(v, HostVelocityField) or (isinstance(v, VelocityField) and v._overridden()):
            from ._lib import DISSIPATION_ALPHA_FN, SCALAR_VELOCITY_FN, VECTOR_VELOCITY_FN

            def scalar(c, m, n, pid, user):
                return float(v.getScalarVelocity((c[0], c[1], c[2]), m, (n[0], n[1], n[2]), pid))

            def vector(c, m, n, pid, user, out):
                r = v.getVectorVelocity((c[0], c[1], c[2]), m, (n[0], n[1], n[2]), pid)
                out[0], out[1], out[2] = float(r[0]), float(r[1]), float(r[2])

            def alpha(direction, m, cd, user):
                return float(v.getDissipationAlpha(direction, m, (cd[0], cd[1], cd[2])))

            self._callbacks = (SCALAR_VELOCITY_FN(scalar), VECTOR_VELOCITY_FN(vector), DISSIPATION_ALPHA_FN(alpha))
            check(lib().lsg_advect_set_velocity_host(self._h, ctypes.cast(self._callbacks[0], ctypes.c_void_p),
                                                     ctypes.cast(self._callbacks[1], ctypes.c_void_p), None))
            check(lib().lsg_advect_set_dissipation_alpha_host(self._h, ctypes.cast(self._callbacks[2],
                                                                                   ctypes.c_void_p), None))
            return
        self._callbacks = None
        if isinstance(v, MaterialVelocityField):
            d = v.default
            vec = (ctypes.c_double * 3)(*d.vec)
            check(lib().lsg_advect_set_velocity_builtin(self._h, d.kind, d.dir, d.k, d.b, d.a_pos, d.a_neg, vec))
            check(lib().lsg_advect_set_dissipation_alpha_builtin(self._h, -1, (ctypes.c_double * 3)(*d.alpha)))
            for m, f in enumerate(v.per_material):
                vec = (ctypes.c_double * 3)(*f.vec)
                check(lib().lsg_advect_set_velocity_builtin_material(self._h, m, f.kind, f.dir, f.k, f.b, f.a_pos,
                                                                     f.a_neg, vec))
                check(lib().lsg_advect_set_dissipation_alpha_builtin(self._h, m, (ctypes.c_double * 3)(*f.alpha)))
            return
        vec = (ctypes.c_double * 3)(*v.vec)
        check(lib().lsg_advect_set_velocity_builtin(self._h, v.kind, v.dir, v.k, v.b, v.a_pos, v.a_neg, vec))
        check(lib().lsg_advect_set_dissipation_alpha_builtin(self._h, -1, (ctypes.c_double * 3)(*v.alpha)))

    def setSpatialScheme(self, s):
        check(lib().lsg_advect_set_spatial_scheme(self._h, int(s)))

    def setIntegrationScheme(self, s):
        self.setSpatialScheme(s)

    def setTemporalScheme(self, s):
        check(lib().lsg_advect_set_temporal_scheme(self._h, int(s)))

    def setAdvectionTime(self, t):
        check(lib().lsg_advect_set_advection_time(self._h, float(t)))

    def setTimeStepRatio(self, c):
        self._cfl = float(c)
        check(lib().lsg_advect_set_time_step_ratio(self._h, float(c)))

    def getTimeStepRatio(self):
        return getattr(self, "_cfl", 0.4999)  # lsAdvect.hpp:66

    def setSingleStep(self, s):
        check(lib().lsg_advect_set_single_step(self._h, int(bool(s))))

    def setDissipationAlpha(self, a):
        check(lib().lsg_advect_set_dissipation_alpha(self._h, float(a)))

    def setCheckDissipation(self, c):
        check(lib().lsg_advect_set_check_dissipation(self._h, int(bool(c))))

    def setCalculateNormalVectors(self, c):
        self._calcNormals = bool(c)
        check(lib().lsg_advect_set_calculate_normal_vectors(self._h, int(bool(c))))

    def getCalculateNormalVectors(self):
        return getattr(self, "_calcNormals", True)  # lsAdvect.hpp:68

    def setSaveAdvectionVelocities(self, on):
        """Store "AdvectionVelocities" and "Dissipation" as point data of the
        advected level set (lsAdvect.hpp:762-796, 916)."""
        check(lib().lsg_advect_set_save_advection_velocities(self._h, int(bool(on))))

    def setVelocityUpdateCallback(self, callback):
        """callback(domain) runs after the first stage of RK2 and after the
        first and second stages of RK3 with the advected (last inserted)
        level set (lsAdvect.hpp:966-969, lsAdvectIntegrationSchemes.hpp:87-89,
        128-130,148-150); its return value is ignored.  It may read level sets
        and change the velocity field (setVelocityField or a host
        VelocityField's state); an exception raised in it stops apply() at
        that stage (the level set keeps the stages done) and is raised again
        by apply(), as pybind11 carries it out of the reference's apply.
        None removes it."""
        from ._lib import VELOCITY_UPDATE_FN
        if callback is None:
            self._updateCallback = None
            check(lib().lsg_advect_set_velocity_update_callback(self._h, None, None))
            return
        import weakref
        ref = weakref.ref(self)

        def trampoline(top, user):
            me = ref()
            try:
                callback(me._domains[-1])
            except BaseException as e:  # raised again by apply(), which stops at this stage
                me._updateError = e
                return -1
            return 0

        self._updateCallback = VELOCITY_UPDATE_FN(trampoline)
        check(lib().lsg_advect_set_velocity_update_callback(
            self._h, ctypes.cast(self._updateCallback, ctypes.c_void_p), None))

    def setIgnoreVoids(self, v):
        check(lib().lsg_advect_set_ignore_voids(self._h, int(bool(v))))

    def setAdaptiveTimeStepping(self, on=True, subdivisions=20):
        check(lib().lsg_advect_set_adaptive_time_stepping(self._h, int(bool(on)), int(subdivisions)))

    def setUpdatePointData(self, u):
        check(lib().lsg_advect_set_update_point_data(self._h, int(bool(u))))

    def prepareLS(self):
        check(lib().lsg_advect_prepare_ls(self._h))

    def apply(self):
        self._updateError = None
        st = lib().lsg_advect_apply(self._h)
        err, self._updateError = self._updateError, None
        if err is not None:
            raise err
        check(st)

    def getAdvectedTime(self):
        t = ctypes.c_double()
        check(lib().lsg_advect_get_advected_time(self._h, ctypes.byref(t)))
        return t.value

    def getCurrentTimeStep(self):
        t = ctypes.c_double()
        check(lib().lsg_advect_get_current_time_step(self._h, ctypes.byref(t)))
        return t.value

    def getNumberOfTimeSteps(self):
        n = ctypes.c_uint()
        check(lib().lsg_advect_get_number_of_time_steps(self._h, ctypes.byref(n)))
        return n.value

    # no counterpart in the reference
    def getPointUpdates(self):
        n = ctypes.c_double()
        check(lib().lsg_advect_get_point_updates(self._h, ctypes.byref(n)))
        return n.value

    def setLearnedWeights(self, params):
        """Network parameters (1347 floats) of WENO_5TH_ORDER_LEARNED."""
        p = np.ascontiguousarray(params, dtype=np.float32)
        check(lib().lsg_advect_set_learned_weights(self._h, p.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                                   int(p.size)))

    def setProfiling(self, on):
        check(lib().lsg_advect_set_profiling(self._h, int(bool(on))))

    def getRateKernelStats(self):
        ms = ctypes.c_double()
        n = ctypes.c_int64()
        b = ctypes.c_double()
        check(lib().lsg_advect_get_rate_kernel_stats(self._h, ctypes.byref(ms), ctypes.byref(n), ctypes.byref(b)))
        return ms.value, n.value, b.value

    def getStarKernelStats(self):
        """(ms, launches, algorithmic bytes) of the star-operation evaluation
        kernel over the steps since setProfiling(True)."""
        ms = ctypes.c_double()
        n = ctypes.c_int64()
        b = ctypes.c_double()
        check(lib().lsg_advect_get_star_kernel_stats(self._h, ctypes.byref(ms), ctypes.byref(n), ctypes.byref(b)))
        return ms.value, n.value, b.value

    def getExpand2KernelStats(self):
        """(ms, launches, algorithmic bytes) of the kernel that evaluates two
        Expand cycles at once, over the steps since setProfiling(True)."""
        ms = ctypes.c_double()
        n = ctypes.c_int64()
        b = ctypes.c_double()
        check(lib().lsg_advect_get_expand2_kernel_stats(self._h, ctypes.byref(ms), ctypes.byref(n), ctypes.byref(b)))
        return ms.value, n.value, b.value


class BatchAdvect:
    """Many independent Advect instances (one level set each) stepped
    together on the GPU: one launch per kernel for all level sets, the step
    captured into a HIP graph, a time step per level set.  apply() is
    Advect::apply with advectionTime 0 for every level set; writeBack()
    stores the results in the domains."""

    def __init__(self, domains, velocities, spatial=SpatialSchemeEnum.WENO_5TH_ORDER,
                 temporal=TemporalSchemeEnum.RUNGE_KUTTA_3RD_ORDER, cfl=0.4):
        self._domains = list(domains)
        arr = (ctypes.c_void_p * len(self._domains))(*[d._h.value for d in self._domains])
        h = ctypes.c_void_p()
        check(lib().lsg_batch_create(ctypes.byref(h), arr, len(self._domains)))
        self._h = h
        for k, v in enumerate(velocities):
            vec = (ctypes.c_double * 3)(*v.vec)
            check(lib().lsg_batch_set_velocity_builtin(h, k, v.kind, v.dir, v.k, v.b, v.a_pos, v.a_neg, vec))
        check(lib().lsg_batch_set_schemes(h, int(spatial), int(temporal), float(cfl)))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                lib().lsg_batch_destroy(h)
            except Exception:
                pass
            self._h = None

    def apply(self):
        check(lib().lsg_batch_apply(self._h))

    def advectedTimes(self):
        t = np.zeros(len(self._domains), np.float64)
        check(lib().lsg_batch_get_advected_times(self._h, t.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return t

    def pointUpdates(self):
        n = ctypes.c_double()
        s = ctypes.c_uint()
        check(lib().lsg_batch_get_point_updates(self._h, ctypes.byref(n), ctypes.byref(s)))
        return n.value, s.value

    def writeBack(self):
        check(lib().lsg_batch_write_back(self._h))

    def setScanBlocks(self, blocks):
        check(lib().lsg_batch_set_scan_blocks(self._h, int(blocks)))

    def autotune(self):
        """Times one step per candidate scan grid (band restored); returns (blocks, ms)."""
        b = ctypes.c_int()
        ms = ctypes.c_double()
        check(lib().lsg_batch_autotune(self._h, ctypes.byref(b), ctypes.byref(ms)))
        return b.value, ms.value


def weno(order, stencils, dx, plus):
    """FiniteDifferences::weno3/weno5 evaluated on the GPU for rows of stencils."""
    x = np.ascontiguousarray(stencils, dtype=np.float64)
    n = x.shape[0]
    out = np.zeros(n, dtype=np.float64)
    check(lib().lsg_weno_eval(int(order), n, x.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), float(dx),
                              int(bool(plus)), out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    return out


def weno5_learned(stencils, dx, plus, params):
    """Learned weno5 (network weights, plain weno5 at kinks) on the GPU."""
    x = np.ascontiguousarray(stencils, dtype=np.float64)
    p = np.ascontiguousarray(params, dtype=np.float32)
    out = np.zeros(x.shape[0], dtype=np.float64)
    check(lib().lsg_weno5_learned_eval(x.shape[0], x.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), float(dx),
                                       int(bool(plus)), p.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                       out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    return out


def rebuild(domain):
    """Advect::rebuildLS (lsAdvect.hpp:238-429) applied to a level set."""
    check(lib().lsg_rebuild(domain._h))


def synchronize():
    check(lib().lsg_device_synchronize())
