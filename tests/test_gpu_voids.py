"""Void detection and Advect::setIgnoreVoids on the GPU: the markers and the
advected level sets bit-exact against the oracle (pinned in
tests/test_oracle_voids.py, including the reference's MarkVoidPoints test)."""
import numpy as np
import pytest

import cases
import pyoracle as po
from test_gpu_parity import assert_same, say, to_gpu
from test_gpu_pointdata import assert_same_data
from test_oracle_voids import plane_with_holes

pytestmark = pytest.mark.gpu


def test_reference_mark_void_points():
    import viennals_amd as vls
    say("MarkVoidPoints: plane y=0.35 with two holes (2D)")
    o = plane_with_holes([([-8.0, -9.5], 5.1), ([8.0, -9.5], 5.1)])
    g = to_gpu(o)
    o.mark_void_points()
    vls.MarkVoidPoints(g).apply()
    assert_same_data(o, g, "markers")
    m = g.getPointData().getScalarData("VoidPointMarkers")
    assert m.sum() > 0


@pytest.mark.parametrize("temporal", [0, 2])
def test_ignore_voids_2d(temporal):
    import viennals_amd as vls
    say(f"ignoreVoids: plane with a hole (2D), temporal {temporal}")
    o = plane_with_holes([([0.0, -5.0], 3.0)], extent=10, z0=0.0)
    g = to_gpu(o)
    p = po.make_params(spatial=0, temporal=temporal, cfl=0.4, time=2.0, vel=dict(b=1.0), ignore_voids=True,
                       save_velocities=True)
    st, r = po.advect(o, p)
    assert st == 0
    a = vls.Advect(g, vls.VelocityField(b=1.0))
    a.setTemporalScheme(temporal)
    a.setTimeStepRatio(0.4)
    a.setAdvectionTime(2.0)
    a.setIgnoreVoids(True)
    a.setSaveAdvectionVelocities(True)
    a.apply()
    assert a.getAdvectedTime() == r["advected_time"]
    assert_same(o, g, "ignoreVoids 2D")
    assert_same_data(o, g, "ignoreVoids 2D")


def test_ignore_voids_3d_trench_with_bubble():
    import viennals_amd as vls
    say("ignoreVoids: 3D trench with a buried bubble, WENO5 RK3")
    L = 16
    o = cases.trench(L=L, width=8, depth=6)
    b = po.Domain(3, [-L, -L, -L], [L, L, L], [0, 0, 1], 1.0)
    b.make_sphere([7.3, 5.2, -9.6], 3.1)
    o.boolean(b, 2)
    g = to_gpu(o)
    vel = cases.VEL_ETCH_DIRECTIONAL
    p = po.make_params(spatial=11, temporal=2, cfl=0.4, time=1.5, vel=vel, ignore_voids=True)
    st, r = po.advect(o, p)
    assert st == 0
    a = vls.Advect(g, vls.VelocityField(**vel))
    a.setSpatialScheme(11)
    a.setTemporalScheme(2)
    a.setTimeStepRatio(0.4)
    a.setAdvectionTime(1.5)
    a.setIgnoreVoids(True)
    a.apply()
    assert a.getAdvectedTime() == r["advected_time"]
    assert_same(o, g, "ignoreVoids 3D")
    assert_same_data(o, g, "ignoreVoids 3D")
    assert g.getPointData().getScalarData("VoidPointMarkers").sum() > 0
