"""MakeGeometry.setIgnoreBoundaryConditions with spheres
(lsMakeGeometry.hpp:125-141,227-235): python/tests/run_all_tests.py's
test_make_geometry_boundary, and on a periodic grid the ignored dimension
keeps the sphere's indices unmapped, so the part that wrapped round is
missing and nothing else changes."""
import numpy as np
import pytest

from test_gpu_parity import say

pytestmark = pytest.mark.gpu


def defined(d):
    idx, val = d.get_runs()
    m = np.abs(val) < 1e300
    return {tuple(int(x) for x in i): v for i, v in zip(idx[m], val[m])}


@pytest.mark.parametrize("dim", [2, 3])
def test_reference_make_geometry_boundary(dim):
    import viennals_amd as vls
    say(f"MakeGeometry.setIgnoreBoundaryConditions, reflective bounds, {dim}D")
    bounds = [-5.0, 5.0] * dim
    bcs = [vls.BoundaryConditionEnum.REFLECTIVE_BOUNDARY] * dim
    a = vls.Domain(bounds, bcs, 1.0)
    m = vls.MakeGeometry(a, vls.Sphere([0.0] * dim, 4.0))
    m.setIgnoreBoundaryConditions(True)
    m.apply()
    assert a.getNumberOfPoints() > 0
    b = vls.Domain(bounds, bcs, 1.0)
    vls.MakeGeometry(b, vls.Sphere([0.0] * dim, 4.0)).apply()
    assert defined(a) == defined(b)  # no periodic dimension: nothing to ignore


def test_periodic_dimension_left_unmapped():
    import viennals_amd as vls
    say("MakeGeometry.setIgnoreBoundaryConditions, sphere across a periodic boundary (2D)")
    bounds = [-10.0, 10.0, -10.0, 10.0]
    bcs = [vls.BoundaryConditionEnum.PERIODIC_BOUNDARY, vls.BoundaryConditionEnum.INFINITE_BOUNDARY]
    wrapped = vls.Domain(bounds, bcs, 1.0)
    vls.MakeGeometry(wrapped, vls.Sphere([8.3, 0.4], 3.1)).apply()
    cut = vls.Domain(bounds, bcs, 1.0)
    m = vls.MakeGeometry(cut, vls.Sphere([8.3, 0.4], 3.1))
    m.setIgnoreBoundaryConditions([True, False])
    m.apply()
    w, c = defined(wrapped), defined(cut)
    assert any(k[0] < 0 for k in w)  # the sphere wraps round to the low side
    assert c and not any(k[0] < 0 for k in c)
    assert all(w[k] == v for k, v in c.items())
    assert {k for k in w if k[0] >= 0} == set(c)
    with pytest.raises(vls.LsgError, match="spheres only"):
        p = vls.MakeGeometry(vls.Domain(bounds, bcs, 1.0), vls.Plane([0.0, 0.0], [0.0, 1.0]))
        p.setIgnoreBoundaryConditions(True)
        p.apply()
