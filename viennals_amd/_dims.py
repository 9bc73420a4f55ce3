"""Dimension-bound module contents for viennals_amd.d2 / viennals_amd.d3,
the drop-in counterparts of the reference's Python modules viennals.d2 and
viennals.d3 (python/viennals/d2.pyi, d3.pyi) for the classes of the
advection path: Domain, Advect, MakeGeometry, Sphere, Plane, Box,
BooleanOperation, Expand, Reduce, Prune, VelocityField and the enums."""
from . import (Advect, BooleanOperation, BooleanOperationEnum, BoundaryConditionEnum, Box, CompareArea, CompareChamfer,
               CompareVolume, Expand,
               FinalizeStencilLocalLaxFriedrichs, HostVelocityField, LsgError, MakeGeometry, MarkVoidPoints,
               MaterialVelocityField, Plane, PointData, PrepareStencilLocalLaxFriedrichs, Prune, Reduce, Sphere,
               SpatialSchemeEnum, TemporalSchemeEnum, VelocityField, Reader, Writer)
from . import Domain as _Domain


def domain_class(D):
    class Domain(_Domain):
        """Domain<double, D> with the reference's Python constructors:
        Domain(), Domain(gridDelta), Domain(bounds, boundaryConditions,
        gridDelta=1.0), Domain(other) (a deep copy)."""

        def __init__(self, *args, **kw):
            if len(args) == 1 and isinstance(args[0], _Domain):
                self._init_from(args[0])
                return
            if not args and "bounds" not in kw:
                super().__init__(gridDelta=float(kw.get("gridDelta", 1.0)), dim=D)
                return
            if len(args) == 1 and not hasattr(args[0], "__len__"):
                super().__init__(gridDelta=float(args[0]), dim=D)
                return
            bounds = args[0] if args else kw["bounds"]
            bcs = args[1] if len(args) > 1 else kw.get("boundaryConditions")
            dx = args[2] if len(args) > 2 else kw.get("gridDelta", 1.0)
            if len(bounds) != 2 * D:
                raise LsgError(f"Domain: {2 * D} bounds expected in {D}D")
            super().__init__(list(bounds), list(bcs) if bcs is not None else None, float(dx))

        def _init_from(self, o):
            if o.D != D:
                raise LsgError("Domain: dimension mismatch")
            fresh = _Domain.from_index_bounds(o.D, o.gmin, o.gmax, o.bc, o.gridDelta)
            self.__dict__.update(fresh.__dict__)
            fresh._h = None
            self.deepCopy(o)

        def getNumberOfSegments(self):
            return 1

    Domain.__qualname__ = Domain.__name__ = "Domain"
    return Domain


__all__ = ["Advect", "BooleanOperation", "BooleanOperationEnum", "BoundaryConditionEnum", "Box", "Expand",
           "HostVelocityField", "MakeGeometry", "MaterialVelocityField", "Plane", "Prune", "Reduce", "Sphere",
           "SpatialSchemeEnum", "TemporalSchemeEnum", "VelocityField", "domain_class"]
