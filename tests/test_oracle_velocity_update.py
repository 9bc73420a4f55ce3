"""Advect::setVelocityUpdateCallback in the oracle
(lsAdvectIntegrationSchemes.hpp:61-168, lsAdvect.hpp:966-969).  Pinned by:

* call counts: never with forward Euler, once per RK2 step, twice per RK3
  step, and a callback that changes nothing leaves the result bit-identical;
* the level set the first call sees is the forward-Euler stage: bit-equal to
  one forward-Euler step from the same start;
* a plane whose speed grows by dv at every call moves exactly as the scalar
  RK recurrence with the speed changed between the stages predicts;
* several level sets, a hole etched through two layers: at every call and
  after the step the lower level set is its start-of-step state cut by the
  current top (INTERSECT), i.e. the first stage adjusts it and every etching
  combine restores it before the next adjustment.
"""
import numpy as np
import pytest

import cases
import pyoracle as po


def run(o, temporal, time, vel, cb=None, cfl=0.4, spatial=0, single=False):
    p = po.make_params(spatial=spatial, temporal=temporal, cfl=cfl, time=time, vel=vel, velocity_update=cb,
                       single_step=single)
    st, r = po.advect(o, p)
    assert st == 0
    return r


@pytest.mark.parametrize("temporal,per_step", [(0, 0), (1, 1), (2, 2)])
def test_call_counts_and_noop(temporal, per_step):
    a = cases.sphere(r=5.0, dx=0.5, origin=[0.1, 0.2, 0.0])
    b = a.copy()
    calls = []
    r = run(a, temporal, 1.0, cases.VEL_EGG, cb=lambda top, p: calls.append(len(top.get()[1])))
    r2 = run(b, temporal, 1.0, cases.VEL_EGG)
    assert len(calls) == per_step * r["steps"]
    assert r["advected_time"] == r2["advected_time"] and r["steps"] == r2["steps"]
    ia, va = a.get()
    ib, vb = b.get()
    assert np.array_equal(ia, ib) and np.array_equal(va.view(np.int64), vb.view(np.int64))


@pytest.mark.parametrize("temporal", [1, 2])
def test_first_call_sees_the_euler_stage(temporal):
    a = cases.sphere(r=5.0, dx=0.5, origin=[0.1, 0.2, 0.0])
    fe = a.copy()
    seen = []
    run(a, temporal, 0.3, cases.VEL_EGG, cb=lambda top, p: seen.append(top.get()), single=True)
    run(fe, 0, 0.3, cases.VEL_EGG, single=True)
    i0, v0 = seen[0]
    i1, v1 = fe.get()
    assert np.array_equal(i0, i1) and np.array_equal(v0.view(np.int64), v1.view(np.int64))


@pytest.mark.parametrize("temporal", [1, 2])
def test_plane_with_growing_speed(temporal):
    dx, T, z0, v0, dv, cfl = 0.5, 1.7, 0.2, 1.0, 0.125, 0.4
    gmin, gmax, bc = [-6, -6, -12], [6, 6, 12], [0, 0, 1]
    d = po.Domain(3, gmin, gmax, bc, dx)
    ip, vp = cases.plane_points(3, gmin, gmax, bc, dx, z0)
    d.from_points(ip, vp, 2)

    def cb(top, p):
        p.velB += dv
    r = run(d, temporal, T, dict(b=v0), cb=cb, cfl=cfl)
    # the scalar recurrence: a plane moves by dt * v per stage; dt of a later
    # stage is min(dt1, cfl * dx / v) (lsAdvect.hpp:616-682)
    z, v, t, steps = z0, v0, 0.0, 0
    while t < T:
        dt1 = min(T - t, cfl * dx / v)
        s1 = z + dt1 * v
        v += dv
        dt2 = min(dt1, cfl * dx / v)
        s2 = s1 + dt2 * v
        if temporal == 1:
            z = 0.5 * z + 0.5 * s2
            t += 0.5 * dt1 + 0.5 * dt2
        else:
            u2 = 0.75 * z + 0.25 * s2
            v += dv
            dt3 = min(dt1, cfl * dx / v)
            s3 = u2 + dt3 * v
            z = z / 3.0 + 2.0 / 3.0 * s3
            t += (dt1 + dt2 + 4.0 * dt3) / 6.0
        steps += 1
    assert r["steps"] == steps
    assert r["advected_time"] == pytest.approx(t, abs=1e-12)
    idx, val = d.defined()
    m = np.abs(val) <= 0.5
    np.testing.assert_allclose(idx[m, 2] * dx - val[m] * dx, z, atol=1e-9)
    assert abs(z - (z0 + v0 * T)) > 0.1  # the speed changes mattered


GMIN, GMAX, BC = [-16, -20], [16, 20], [0, 1]


def _plane(z0):
    d = po.Domain(2, GMIN, GMAX, BC, 1.0)
    i, v = cases.plane_points(2, GMIN, GMAX, BC, 1.0, z0)
    d.from_points(i, v, 2)
    return d


def _hole_through_two_layers():
    """A top layer over a lower one (y < -2.3), a circular hole through both."""
    low, top = _plane(-2.3), _plane(0.0)
    for d in (low, top):
        h = po.Domain(2, GMIN, GMAX, BC, 1.0)
        h.make_sphere([0.4, -0.6], 3.2)
        d.boolean(h, 2)
    return low, top


def _cut(low, idx, val, width):
    t = po.Domain(2, GMIN, GMAX, BC, 1.0)
    t.set(idx, val, width)
    e = low.copy()
    e.boolean(t, 0)  # INTERSECT, as adjustLowerLayers (lsAdvect.hpp:802-813)
    return e.get()


def _same(a, b):
    return np.array_equal(a[0], b[0]) and np.array_equal(a[1].view(np.int64), b[1].view(np.int64))


@pytest.mark.parametrize("temporal", [1, 2])
def test_several_level_sets_lower_layer_follows_the_top(temporal):
    """With a callback the lower level set is, at every call and after the
    step, its start-of-step state cut by the current top: the first stage
    adjusts it (:80-83,121-124) and an etching combine restores it before the
    next adjustment (:98-103,139-144,157-163)."""
    low, top = _hole_through_two_layers()
    start = low.copy()
    checks = []

    def cb(view, p):
        idx, val = view.get()
        now = low.get()
        checks.append((_same(now, _cut(start, idx, val, view.width())), not _same(now, start.get())))
    mats = [dict(b=-0.8), dict(b=-1.0)]
    p = po.make_params(spatial=0, temporal=temporal, cfl=0.4, time=1.0, vel=mats[-1], mat_vels=mats,
                       velocity_update=cb, single_step=True)
    st, r = po.advect_multi([low, top], p)
    assert st == 0 and r["steps"] == 1
    assert len(checks) == temporal
    assert all(ok for ok, _ in checks)
    assert checks[0][1]  # the first stage did cut the lower level set
    i, v = top.get()
    assert _same(low.get(), _cut(start, i, v, top.width))
