"""Advect::setVelocityUpdateCallback on the GPU: bit-exact against the oracle
(pinned in tests/test_oracle_velocity_update.py) -- the level set each call
sees, the velocity changed by the callback (a host VelocityField's state or a
new built-in functor), several level sets, and the refused misuses."""
import numpy as np
import pytest

import cases
import pyoracle as po
from test_gpu_parity import assert_same, say, to_gpu
from test_oracle_velocity_update import _hole_through_two_layers

pytestmark = pytest.mark.gpu


def same_runs(a, b):
    return np.array_equal(a[0], b[0]) and np.array_equal(a[1].view(np.int64), b[1].view(np.int64))


@pytest.mark.parametrize("temporal", [1, 2])
@pytest.mark.parametrize("host", [True, False])
def test_speed_changed_between_stages(temporal, host):
    import viennals_amd as vls
    say(f"velocity update callback: sphere, temporal {temporal}, host velocity {host}")
    dv = 0.0625
    o = cases.sphere(r=5.0, dx=0.5, origin=[0.1, 0.2, 0.0])
    g = to_gpu(o)
    seen_o = []

    def cb_o(view, p):
        seen_o.append(view.get())
        p.velB += dv
    p = po.make_params(spatial=0, temporal=temporal, cfl=0.4, time=1.0, vel=dict(b=1.0), velocity_update=cb_o)
    st, r = po.advect(o, p)
    assert st == 0

    class Field(vls.VelocityField):
        b_now = 1.0

        def getScalarVelocity(self, coord, material, normal, pointId):
            return self.b_now

        def getVectorVelocity(self, coord, material, normal, pointId):
            return (0.0, 0.0, 0.0)
    field = Field()
    a = vls.Advect(g, field if host else vls.VelocityField(b=1.0))
    seen_g = []
    state = {"b": 1.0}

    def cb_g(dom):
        assert dom is g
        seen_g.append(dom.get_runs())
        state["b"] += dv
        if host:
            field.b_now = state["b"]
        else:
            a.setVelocityField(vls.VelocityField(b=state["b"]))
        return True
    a.setVelocityUpdateCallback(cb_g)
    a.setTemporalScheme(temporal)
    a.setTimeStepRatio(0.4)
    a.setAdvectionTime(1.0)
    a.apply()
    assert a.getAdvectedTime() == r["advected_time"] and a.getNumberOfTimeSteps() == r["steps"]
    assert len(seen_g) == len(seen_o) == temporal * r["steps"]
    for k, (x, y) in enumerate(zip(seen_o, seen_g)):
        assert same_runs(x, y), f"call {k}"
    assert_same(o, g, "after the callbacks")


@pytest.mark.parametrize("temporal", [1, 2])
def test_several_level_sets(temporal):
    import viennals_amd as vls
    say(f"velocity update callback: hole through two layers (2D), temporal {temporal}")
    low, top = _hole_through_two_layers()
    gl, gt = to_gpu(low), to_gpu(top)
    mats = [dict(b=-0.8), dict(b=-1.0)]

    def cb_o(view, p):
        p.matVel[0].b -= 0.125
        p.matVel[1].b -= 0.0625
    p = po.make_params(spatial=0, temporal=temporal, cfl=0.4, time=1.5, vel=mats[-1], mat_vels=mats,
                       velocity_update=cb_o)
    st, r = po.advect_multi([low, top], p)
    assert st == 0
    a = vls.Advect()
    a.insertNextLevelSet(gl)
    a.insertNextLevelSet(gt)
    cur = [dict(m) for m in mats]
    a.setVelocityField(vls.MaterialVelocityField([vls.VelocityField(**m) for m in cur]))

    def cb_g(dom):
        cur[0]["b"] -= 0.125
        cur[1]["b"] -= 0.0625
        a.setVelocityField(vls.MaterialVelocityField([vls.VelocityField(**m) for m in cur]))
    a.setVelocityUpdateCallback(cb_g)
    a.setTemporalScheme(temporal)
    a.setTimeStepRatio(0.4)
    a.setAdvectionTime(1.5)
    a.apply()
    assert a.getAdvectedTime() == r["advected_time"]
    assert_same(top, gt, "top")
    assert_same(low, gl, "lower")


def test_refused_misuse():
    import viennals_amd as vls
    say("velocity update callback: changing the band or advecting inside it is refused")
    g = to_gpu(cases.sphere(r=4.0, dx=0.5))
    a = vls.Advect(g, vls.VelocityField(b=1.0))
    a.setTemporalScheme(1)
    a.setAdvectionTime(0.2)
    a.setVelocityUpdateCallback(lambda d: vls.Expand(d, 5).apply())
    with pytest.raises(vls.LsgError, match="must not modify"):
        a.apply()
    g2 = to_gpu(cases.sphere(r=4.0, dx=0.5))
    b = vls.Advect(g2, vls.VelocityField(b=1.0))
    b.setTemporalScheme(2)
    b.setAdvectionTime(0.2)
    b.setVelocityUpdateCallback(lambda d: b.apply())
    with pytest.raises(vls.LsgError, match="velocity update callback"):
        b.apply()
    # without the callback the same Advect runs again
    b.setVelocityUpdateCallback(None)
    b.apply()
    assert b.getAdvectedTime() == pytest.approx(0.2)


def test_exception_in_callback_stops_apply():
    """An exception raised by the callback ends apply() at that stage (as
    pybind11 carries it out of the reference's apply), so the callback is not
    called again by the remaining stages and steps."""
    import viennals_amd as vls
    say("velocity update callback: an exception stops apply at its stage")
    g = to_gpu(cases.sphere(r=4.0, dx=0.5))
    a = vls.Advect(g, vls.VelocityField(b=1.0))
    a.setTemporalScheme(2)
    a.setAdvectionTime(100.0)
    calls = []

    def cb(d):
        calls.append(d.getNumberOfPoints())
        if len(calls) == 3:
            raise ValueError("stop here")

    a.setVelocityUpdateCallback(cb)
    with pytest.raises(ValueError, match="stop here"):
        a.apply()
    assert len(calls) == 3  # RK3: two calls in step 1, the first call of step 2
