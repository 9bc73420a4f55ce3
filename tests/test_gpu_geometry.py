"""MakeGeometry on the GPU path: setIgnoreBoundaryConditions with spheres
(lsMakeGeometry.hpp:125-141,227-235; python/tests/run_all_tests.py's
test_make_geometry_boundary, and on a periodic grid the ignored dimension
keeps the sphere's indices unmapped, so the part that wrapped round is
missing and nothing else changes), and planes and boxes through
FromSurfaceMesh bit-exact against the oracle."""
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
    # planes ignore the flags (makePlane uses FromSurfaceMesh's defaults)
    p1 = vls.Domain(bounds, bcs, 1.0)
    m = vls.MakeGeometry(p1, vls.Plane([0.0, 0.0], [0.0, 1.0]))
    m.setIgnoreBoundaryConditions(True)
    m.apply()
    p2 = vls.Domain(bounds, bcs, 1.0)
    vls.MakeGeometry(p2, vls.Plane([0.0, 0.0], [0.0, 1.0])).apply()
    assert defined(p1) == defined(p2)


@pytest.mark.parametrize("case", ["trench_plane", "trench_box", "tilted_plane", "tilted_line_2d", "box_periodic",
                                  "box_periodic_ignore", "box_2d"])
def test_plane_and_box_match_oracle(case):
    """MakeGeometry with planes and boxes on the GPU path (FromSurfaceMesh,
    lsFromSurfaceMesh.hpp:243-480) against the oracle's restatement (pinned
    by the closed forms and the golden fixture in
    tests/test_oracle_geometry.py): the run lists bit for bit."""
    import viennals_amd as vls
    import cases
    import pyoracle as po
    from test_oracle_geometry import GEOMETRY_CASES
    say(f"MakeGeometry plane/box: {case}")
    kind, D, gmin, gmax, bc, dx, a, b, mask = GEOMETRY_CASES[case]
    o = po.Domain(D, gmin, gmax, bc, dx)
    g = vls.Domain.from_index_bounds(D, gmin, gmax, bc, dx)
    if kind == 0:
        assert o.make_plane(a, b) == 0
        vls.MakeGeometry(g, vls.Plane(a, b)).apply()
    else:
        o.make_box(a, b, mask)
        m = vls.MakeGeometry(g, vls.Box(a, b))
        m.setIgnoreBoundaryConditions([bool(mask & (1 << k)) for k in range(3)])
        m.apply()
    io, vo = o.get()
    ig, vg = g.get_runs()
    assert cases.bit_equal(io, vo, ig, vg), cases.first_diff(io, vo, ig, vg)
    assert g.getLevelSetWidth() == 2
