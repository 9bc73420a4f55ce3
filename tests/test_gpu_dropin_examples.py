"""A reference example run unchanged through the drop-in module
(examples/VoidEtching/VoidEtching.py without its VTK mesh output): plane
minus four spheres, a VelocityField subclass etching isotropically (the host
callback path), setIgnoreVoids(True); every step bit-exact against the
oracle advecting the same band with the built-in constant functor."""
import numpy as np
import pytest

import pyoracle as po
from test_gpu_parity import assert_same, say

pytestmark = pytest.mark.gpu


def test_void_etching_example():
    import viennals_amd.viennals as viennals
    say("VoidEtching.py through viennals_amd.viennals (3D, host velocity, ignoreVoids)")
    viennals.setDimension(3)

    class VelocityField(viennals.VelocityField):
        def getScalarVelocity(self, coord, material, normal, pointId):
            return -1.0

        def getVectorVelocity(self, coord, material, normal, pointId):
            return [0.0, 0.0, 0.0]

    extent, gridDelta = 30.0, 1.0
    bounds = [-extent, extent, -extent, extent, -extent, extent]
    bcs = [viennals.BoundaryConditionEnum.REFLECTIVE_BOUNDARY, viennals.BoundaryConditionEnum.REFLECTIVE_BOUNDARY,
           viennals.BoundaryConditionEnum.INFINITE_BOUNDARY]
    substrate = viennals.Domain(bounds, bcs, gridDelta)
    viennals.MakeGeometry(substrate, viennals.Plane([0.0, 0.0, 0.0], [0.0, 0.0, 1.0])).apply()
    sphere = viennals.Domain(bounds, bcs, gridDelta)
    boolOp = viennals.BooleanOperation(substrate, sphere, viennals.BooleanOperationEnum.RELATIVE_COMPLEMENT)
    for origin, radius in (([-12.0, -5.0, -15.0], 10.0), ([-7.0, -30.0, -20.0], 8.0), ([5.0, 15.0, -2.0], 8.0),
                           ([2.0, 8.0, -27.0], 8.0)):
        viennals.MakeGeometry(sphere, viennals.Sphere(origin, radius)).apply()
        boolOp.apply()
    # the oracle starts from the same band
    idx, val = substrate.get_runs()
    o = po.Domain(3, substrate.gmin, substrate.gmax, substrate.bc, gridDelta)
    o.set(idx, val, substrate.getLevelSetWidth())
    p = po.make_params(spatial=0, temporal=0, time=0.0, vel=dict(b=-1.0), ignore_voids=True)

    advectionKernel = viennals.Advect()
    advectionKernel.insertNextLevelSet(substrate)
    advectionKernel.setVelocityField(VelocityField())
    advectionKernel.setIgnoreVoids(True)
    passed = 0.0
    for i in range(8):
        advectionKernel.apply()
        st, r = po.advect(o, p)
        assert st == 0
        assert advectionKernel.getAdvectedTime() == r["advected_time"]
        passed += advectionKernel.getAdvectedTime()
        assert_same(o, substrate, f"step {i}")
    marks = substrate.getPointData().getScalarData("VoidPointMarkers")
    np.testing.assert_array_equal(marks, o.data("VoidPointMarkers"))
    assert marks.sum() > 0 and passed > 0
