"""Multi-material advection in the oracle (Advect with several level sets,
lsAdvect.hpp:435-612, 686-813): pinned by cases with a closed-form result.
A lower level set the top never reaches leaves the top's advection equal to
the single-level-set one; an etch-stop layer stops the top exactly on it
(the terminal step snaps to its value, lsAdvect.hpp:575-584, 747-759); a
mask of zero velocity keeps the surface above it in place while the open
surface beside it moves by v*t."""
import numpy as np
import pytest

import pyoracle as po
import cases

REF, INF = 0, 1
GMIN, GMAX, BC = [-16, -20], [16, 20], [REF, INF]


def plane(z0):
    d = po.Domain(2, GMIN, GMAX, BC, 1.0)
    i, v = cases.plane_points(2, GMIN, GMAX, BC, 1.0, z0)
    d.from_points(i, v, 2)
    return d


def as_dict(d):
    i, v = d.get()
    return {tuple(int(x) for x in a): float(b) for a, b in zip(i, v)}


def defined(m):
    return {k: v for k, v in m.items() if abs(v) < 1e300}


@pytest.mark.parametrize("spatial,temporal", [(0, 0), (1, 1), (11, 2)])
def test_unreached_lower_layer_leaves_top_unchanged(spatial, temporal):
    top, low = plane(0.0), plane(-9.3)
    ref = plane(0.0)
    p = po.make_params(spatial=spatial, temporal=temporal, cfl=0.4, time=2.0, vel=dict(b=-1.0),
                       mat_vels=[dict(b=-5.0), dict(b=-1.0)])
    st, r = po.advect_multi([low, top], p)
    assert st == 0
    st1, r1 = po.advect(ref, po.make_params(spatial=spatial, temporal=temporal, cfl=0.4, time=2.0,
                                            vel=dict(b=-1.0)))
    assert st1 == 0 and r["advected_time"] == r1["advected_time"] and r["steps"] == r1["steps"]
    assert as_dict(top) == as_dict(ref)
    assert as_dict(low) == as_dict(plane(-9.3))  # max(lower, top) is the lower layer


@pytest.mark.parametrize("spatial", [0, 11])
def test_etch_stops_on_lower_layer(spatial):
    # forward Euler: the terminal step lands exactly on the stop layer
    top, low = plane(0.0), plane(-2.3)
    p = po.make_params(spatial=spatial, temporal=0, cfl=0.4, time=5.0, vel=dict(b=-1.0),
                       mat_vels=[dict(b=0.0), dict(b=-1.0)])
    st, r = po.advect_multi([low, top], p)
    assert st == 0 and r["advected_time"] == 5.0
    t, l = defined(as_dict(top)), defined(as_dict(low))
    assert t and all(t[k] == l[k] for k in t)  # the surface sits on the stop layer
    for x in range(GMIN[0], GMAX[0] + 1):
        assert abs(t[(x, -2)] - 0.3) < 1e-12


def test_etch_stop_rk3_never_passes_the_layer():
    # RK3 combines a snapped stage with the step's start, so it approaches
    # the layer from above without reaching it exactly, and never passes it
    top, low = plane(0.0), plane(-2.3)
    p = po.make_params(spatial=11, temporal=2, cfl=0.4, time=5.0, vel=dict(b=-1.0),
                       mat_vels=[dict(b=0.0), dict(b=-1.0)])
    st, r = po.advect_multi([low, top], p)
    assert st == 0 and r["advected_time"] == 5.0
    t, l = defined(as_dict(top)), defined(as_dict(low))
    assert all(t[k] <= l[k] for k in t if k in l)
    for x in range(GMIN[0], GMAX[0] + 1):
        assert 0.1 < t[(x, -2)] <= 0.3


def test_mask_keeps_masked_surface():
    mask = po.Domain(2, GMIN, GMAX, BC, 1.0)
    i, v = cases.box_points(2, [2.5, -1.5], [12.5, 2.0], 1.0, GMIN, GMAX)
    mask.from_points(i, v, 2)
    top = plane(0.0)
    top.boolean(mask, 1)  # the top level set wraps the mask (UNION)
    before = defined(as_dict(top))
    p = po.make_params(spatial=0, temporal=0, cfl=0.4, time=3.0, vel=dict(b=-1.0),
                       mat_vels=[dict(b=0.0), dict(b=-1.0)])
    st, r = po.advect_multi([mask, top], p)
    assert st == 0 and r["advected_time"] == 3.0
    after = defined(as_dict(top))
    # open surface far from the mask: the plane moved down by 3 exactly
    for x in range(-16, -6):
        for y in (-4, -3, -2):
            if (x, y) in after:
                assert abs(after[(x, y)] - (y + 3.0)) < 1e-12, (x, y, after[(x, y)])
        assert (x, -3) in after
    # under the mask (away from its edges) the surface stays at the mask top y = 2
    for x in range(6, 10):
        assert after[(x, 2)] == before[(x, 2)] == 0.0, x
        assert not any(k[0] == x and k[1] < 1 for k in after), x


def test_adaptive_time_stepping_near_the_stop_layer():
    # setAdaptiveTimeStepping: within one CFL budget of the layer below the
    # step is limited to cfl/subdivisions (lsAdvect.hpp:568-578) until the
    # remaining distance is below that; the surface still lands on the layer
    runs = {}
    for adaptive in (False, True):
        top, low = plane(0.0), plane(-2.3)
        p = po.make_params(spatial=0, temporal=0, cfl=0.4, time=5.0, vel=dict(b=-1.0),
                           mat_vels=[dict(b=0.0), dict(b=-1.0)], adaptive=adaptive, subdivisions=8)
        st, r = po.advect_multi([low, top], p)
        assert st == 0 and r["advected_time"] == 5.0
        t, l = defined(as_dict(top)), defined(as_dict(low))
        assert t and all(t[k] == l[k] for k in t)
        runs[adaptive] = r["steps"]
    assert runs[True] > runs[False] + 4, runs
