"""Drop-in counterpart of the reference's Python module viennals.d3
(python/viennals/d3.pyi) for the advection path: the same class names and
call forms, bound to 3 dimensions.  `from viennals_amd import d3 as vls`."""
from ._dims import (Advect, BooleanOperation, BooleanOperationEnum, BoundaryConditionEnum, Box, CompareArea,
                    CompareChamfer, CompareVolume,
                    Expand, FinalizeStencilLocalLaxFriedrichs, HostVelocityField, MakeGeometry, MarkVoidPoints,
                    MaterialVelocityField, Plane, PointData, PrepareStencilLocalLaxFriedrichs, Prune, Reduce, Sphere,
                    SpatialSchemeEnum, TemporalSchemeEnum, VelocityField, Reader, Writer, domain_class)

D = 3
Domain = domain_class(3)

__all__ = ["Advect", "BooleanOperation", "BooleanOperationEnum", "BoundaryConditionEnum", "Box", "CompareArea", "CompareChamfer", "CompareVolume",
           "Domain", "Expand", "FinalizeStencilLocalLaxFriedrichs", "HostVelocityField", "MakeGeometry", "MarkVoidPoints", "MaterialVelocityField", "Plane",
           "PointData", "PrepareStencilLocalLaxFriedrichs", "Prune", "Reduce", "Sphere", "SpatialSchemeEnum", "TemporalSchemeEnum", "VelocityField", "Reader", "Writer"]
