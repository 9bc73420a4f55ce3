"""The Python drop-in modules viennals_amd.d2 / d3 offer the methods that
the reference's viennals.d2 / d3 modules (python/viennals/d3.pyi) have on
the advection path's classes (CPU: names and call forms only)."""
import pytest

REFERENCE_METHODS = {
    "Advect": ["apply", "clearLevelSets", "getAdvectedTime", "getCalculateNormalVectors", "getCurrentTimeStep",
               "getNumberOfTimeSteps", "getTimeStepRatio", "insertNextLevelSet", "prepareLS",
               "setAdaptiveTimeStepping", "setAdvectionTime", "setCalculateNormalVectors", "setCheckDissipation",
               "setDissipationAlpha", "setIgnoreVoids", "setIntegrationScheme", "setSaveAdvectionVelocities",
               "setSingleStep", "setSpatialScheme", "setTemporalScheme", "setTimeStepRatio", "setUpdatePointData",
               "setVelocityField", "setVelocityUpdateCallback"],
    "BooleanOperation": ["apply", "setBooleanOperation", "setLevelset", "setSecondLevelSet"],
    "Domain": ["clearMetaData", "deepCopy", "getLevelSetWidth", "getNumberOfPoints", "getNumberOfSegments", "print",
               "setLevelSetWidth"],
    "Expand": ["apply", "setLevelSet", "setWidth"],
    "MakeGeometry": ["apply", "setGeometry", "setIgnoreBoundaryConditions", "setLevelSet"],
    "Prune": ["apply", "setLevelSet"],
    "Reduce": ["apply", "setLevelSet", "setNoNewSegment", "setWidth"],
}


@pytest.mark.parametrize("dim", [2, 3])
def test_reference_methods_present(dim):
    import importlib
    m = importlib.import_module(f"viennals_amd.d{dim}")
    for cls, names in REFERENCE_METHODS.items():
        c = getattr(m, cls)
        missing = [n for n in names if not hasattr(c, n)]
        assert not missing, (cls, missing)
    for name in ("Sphere", "Plane", "Box", "VelocityField", "SpatialSchemeEnum", "TemporalSchemeEnum",
                 "BoundaryConditionEnum", "BooleanOperationEnum"):
        assert hasattr(m, name)
    assert m.Domain.__name__ == "Domain" and m.D == dim


def test_velocity_field_subclass_goes_to_the_host_path():
    import viennals_amd as vls

    class V(vls.VelocityField):
        def getScalarVelocity(self, coordinate, material, normal, pointId):
            return 1.0

    assert V()._overridden() and not vls.VelocityField.constant(1.0)._overridden()
    f = vls.VelocityField.normal_power(dir=2, k=2, b=0.1, a_pos=2.0, a_neg=0.5)
    assert f.getScalarVelocity((0, 0, 0), 0, (0.0, 0.0, -0.5), 0) == 0.1 + 0.5 * 0.25
