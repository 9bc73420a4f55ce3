"""Drop-in counterpart of the reference's Python module viennals.d2
(python/viennals/d2.pyi) for the advection path: the same class names and
call forms, bound to 2 dimensions.  `from viennals_amd import d2 as vls`."""
from ._dims import (Advect, BooleanOperation, BooleanOperationEnum, BoundaryConditionEnum, Box, CompareArea,
                    CompareChamfer, CompareVolume,
                    Expand, FinalizeStencilLocalLaxFriedrichs, HostVelocityField, MakeGeometry, MarkVoidPoints,
                    MaterialVelocityField, Plane, PointData, PrepareStencilLocalLaxFriedrichs, Prune, Reduce, Sphere,
                    SpatialSchemeEnum, TemporalSchemeEnum, VelocityField, Reader, Writer, domain_class)

D = 2
Domain = domain_class(2)

__all__ = ["Advect", "BooleanOperation", "BooleanOperationEnum", "BoundaryConditionEnum", "Box", "CompareArea", "CompareChamfer", "CompareVolume",
           "Domain", "Expand", "FinalizeStencilLocalLaxFriedrichs", "HostVelocityField", "MakeGeometry", "MarkVoidPoints", "MaterialVelocityField", "Plane",
           "PointData", "PrepareStencilLocalLaxFriedrichs", "Prune", "Reduce", "Sphere", "SpatialSchemeEnum", "TemporalSchemeEnum", "VelocityField", "Reader", "Writer"]
