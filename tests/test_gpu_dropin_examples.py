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


def _oracle_copy(d):
    idx, val = d.get_runs()
    o = po.Domain(d.D, d.gmin, d.gmax, d.bc, d.gridDelta)
    o.set(idx, val, d.getLevelSetWidth())
    return o


@pytest.mark.parametrize("temporal,host", [(0, True), (1, True), (2, True), (2, False)])
def test_air_gap_deposition_example(temporal, host):
    """examples/AirGapDeposition/AirGapDeposition.py without its mesh output
    and with 6 time units instead of 16.5 (the air gap closes after 4): a trench (plane minus box), a new
    layer copied from it, a VelocityField subclass returning sum |n_i| for
    every material (the host callback path with two level sets),
    setIgnoreVoids(True); both level sets bit-exact against the oracle after
    every apply, the oracle using the built-in normal-sum functor (which the
    GPU runs too when host is False)."""
    import viennals_amd.viennals as vls
    say(f"AirGapDeposition.py through viennals_amd.viennals (2D, two level sets, host velocity {host}), "
        f"temporal {temporal}")
    vls.setDimension(2)

    class velocityField(vls.VelocityField):
        def getScalarVelocity(self, coord, material, normal, pointId):
            velocity = 0.0
            for i in range(len(normal)):
                velocity += abs(normal[i])
            return velocity

        def getVectorVelocity(self, coord, material, normal, pointId):
            return (0.0, 0.0, 0.0)

    extent, gridDelta = 30.0, 0.5
    bounds = (-extent, extent, -extent, extent)
    bcs = (vls.BoundaryConditionEnum.REFLECTIVE_BOUNDARY, vls.BoundaryConditionEnum.INFINITE_BOUNDARY)
    substrate = vls.Domain(bounds, bcs, gridDelta)
    vls.MakeGeometry(substrate, vls.Plane((0.0, 0.0), (0.0, 1.0))).apply()
    trench = vls.Domain(bounds, bcs, gridDelta)
    vls.MakeGeometry(trench, vls.Box((-extent / 6.0, -25.0), (extent / 6.0, 1.0))).apply()
    vls.BooleanOperation(substrate, trench, vls.BooleanOperationEnum.RELATIVE_COMPLEMENT).apply()
    sub = vls.Domain(substrate)
    layer = vls.Domain(sub)
    osub, olayer = _oracle_copy(sub), _oracle_copy(layer)

    kernel = vls.Advect()
    kernel.insertNextLevelSet(sub)
    kernel.insertNextLevelSet(layer)
    kernel.setVelocityField(velocityField() if host else vls.VelocityField.normal_sum(1.0))
    kernel.setIgnoreVoids(True)
    kernel.setTemporalScheme(temporal)
    vel = dict(kind=2, b=0.0, a_pos=1.0)
    passed, total, interval = 0.0, 6.0, 1.0
    while passed < total:
        dt = min(interval, total - passed)
        kernel.setAdvectionTime(dt)
        kernel.apply()
        p = po.make_params(spatial=0, temporal=temporal, time=dt, vel=vel, mat_vels=[vel, vel], ignore_voids=True)
        st, r = po.advect_multi([osub, olayer], p)
        assert st == 0
        assert kernel.getAdvectedTime() == r["advected_time"]
        passed += kernel.getAdvectedTime()
        assert_same(olayer, layer, f"new layer at t={passed}")
        assert_same(osub, sub, f"substrate at t={passed}")
    assert passed == pytest.approx(total)
    marks = layer.getPointData().getScalarData("VoidPointMarkers")
    np.testing.assert_array_equal(marks, olayer.data("VoidPointMarkers"))
    assert marks.sum() > 0  # the air gap
