"""Several level sets on the GPU (Advect with level sets below the advected
top, lsAdvect.hpp:435-612, 686-813): bit-exact against the oracle's
restatement (oracle/ls_oracle.cpp, computeRatesMM / updateLevelSetMM /
adjustLowerLayers / evolveRK3), which the CPU tests pin by closed-form
cases (tests/test_oracle_multimaterial.py)."""
import numpy as np
import pytest

import cases
import pyoracle as po
from test_gpu_parity import assert_same, say, to_gpu

pytestmark = pytest.mark.gpu

REF, INF = 0, 1


def plane2(z0, gmin=(-16, -20), gmax=(16, 20)):
    gmin, gmax, bc = list(gmin), list(gmax), [REF, INF]
    d = po.Domain(2, gmin, gmax, bc, 1.0)
    i, v = cases.plane_points(2, gmin, gmax, bc, 1.0, z0)
    d.from_points(i, v, 2)
    return d


def run_both(levels, spatial, temporal, cfl, time, mat_vels, steps=1, adaptive=0):
    """Advect the oracle level sets and GPU copies of them; compare every
    level set and the advected time after each apply()."""
    import viennals_amd as vls
    gpu = [to_gpu(o) for o in levels]
    p = po.make_params(spatial=spatial, temporal=temporal, cfl=cfl, time=time, vel=mat_vels[-1],
                       mat_vels=mat_vels, adaptive=adaptive > 0, subdivisions=max(adaptive, 1))
    a = vls.Advect()
    if adaptive:
        a.setAdaptiveTimeStepping(True, adaptive)
    for g in gpu:
        a.insertNextLevelSet(g)
    a.setVelocityField(vls.MaterialVelocityField([vls.VelocityField(**v) for v in mat_vels]))
    a.setSpatialScheme(spatial)
    a.setTemporalScheme(temporal)
    a.setTimeStepRatio(cfl)
    a.setAdvectionTime(time)
    for s in range(steps):
        st, r = po.advect_multi(levels, p)
        assert st == 0
        a.apply()
        assert a.getAdvectedTime() == r["advected_time"], (s, a.getAdvectedTime(), r["advected_time"])
        for k, (o, g) in enumerate(zip(levels, gpu)):
            assert_same(o, g, f"step {s} level set {k}")
    return gpu


@pytest.mark.parametrize("spatial,temporal", [(0, 0), (11, 2)])
def test_unreached_lower_layer(spatial, temporal):
    say(f"two level sets, lower layer never reached: spatial {spatial} temporal {temporal}")
    run_both([plane2(-9.3), plane2(0.0)], spatial, temporal, 0.4, 2.0, [dict(b=-5.0), dict(b=-1.0)])


@pytest.mark.parametrize("spatial,temporal", [(0, 0), (1, 1), (11, 2)])
def test_etch_stop_layer(spatial, temporal):
    say(f"two level sets, etch stop at z=-2.3: spatial {spatial} temporal {temporal}")
    run_both([plane2(-2.3), plane2(0.0)], spatial, temporal, 0.4, 5.0, [dict(b=0.0), dict(b=-1.0)])


@pytest.mark.parametrize("spatial,temporal", [(0, 0), (11, 2)])
def test_etch_stop_adaptive_time_stepping(spatial, temporal):
    say(f"two level sets, etch stop with adaptive time stepping (8): spatial {spatial} temporal {temporal}")
    run_both([plane2(-2.3), plane2(0.0)], spatial, temporal, 0.4, 5.0, [dict(b=0.0), dict(b=-1.0)], adaptive=8)


@pytest.mark.parametrize("spatial,temporal", [(0, 0), (5, 1), (7, 2)])
def test_mask_2d(spatial, temporal):
    say(f"mask over a plane (2D): spatial {spatial} temporal {temporal}")
    gmin, gmax, bc = [-16, -20], [16, 20], [REF, INF]
    mask = po.Domain(2, gmin, gmax, bc, 1.0)
    i, v = cases.box_points(2, [2.5, -1.5], [12.5, 2.0], 1.0, gmin, gmax)
    mask.from_points(i, v, 2)
    top = plane2(0.0)
    top.boolean(mask, 1)
    run_both([mask, top], spatial, temporal, 0.4, 3.0, [dict(b=0.0), dict(b=-1.0)])


def test_trench_mask_stop_layer_3d():
    """Three level sets in 3D (stop layer, mask, wrapping top), directional
    etch with material-dependent rates, WENO5 + RK3 over several steps: the
    last combine moves the surface down, so the lower level sets are
    restored from the start of each step before they are intersected."""
    say("3D: stop layer + mask + top, WENO5/RK3, 6 steps")
    L = 20
    gmin, gmax, bc = [-L, -L, -L], [L, L, L], [REF, REF, INF]
    stop = po.Domain(3, gmin, gmax, bc, 1.0)
    i, v = cases.plane_points(3, gmin, gmax, bc, 1.0, -3.4)
    stop.from_points(i, v, 2)
    mask = po.Domain(3, gmin, gmax, bc, 1.0)
    i, v = cases.box_points(3, [-L + 0.5, -L + 0.5, -0.5], [-4.5, L - 0.5, 1.6], 1.0, gmin, gmax)
    mask.from_points(i, v, 2)
    top = po.Domain(3, gmin, gmax, bc, 1.0)
    i, v = cases.plane_points(3, gmin, gmax, bc, 1.0, 0.0)
    top.from_points(i, v, 2)
    top.boolean(mask, 1)
    mat = [dict(b=0.0), dict(kind=1, dir=2, k=1, b=-0.05, a_pos=-0.1, a_neg=0.0),
           dict(kind=1, dir=2, k=1, b=-0.1, a_pos=-1.0, a_neg=0.0)]
    run_both([stop, mask, top], 11, 2, 0.4, 0.0, mat, steps=6)


@pytest.mark.parametrize("spatial,temporal", [(2, 0), (3, 2)])
def test_lax_friedrichs_etch_stop_layer(spatial, temporal):
    """LF1/LF2 with a stop layer: findGlobalAlphas takes every active
    point's material (lsAdvect.hpp:133-171), the global alphas bound dt."""
    say(f"two level sets, etch stop, Lax-Friedrichs: spatial {spatial} temporal {temporal}")
    mats = [dict(kind=1, dir=2, k=1, b=-0.05, a_pos=-0.2, a_neg=0.0),
            dict(kind=1, dir=2, k=1, b=-0.1, a_pos=-1.0, a_neg=0.0)]
    run_both([plane2(-2.3), plane2(0.0)], spatial, temporal, 0.4, 4.0, mats)


def test_lax_friedrichs_mask_2d():
    say("mask over a plane (2D), LF2 + RK2")
    gmin, gmax, bc = [-16, -20], [16, 20], [REF, INF]
    mask = po.Domain(2, gmin, gmax, bc, 1.0)
    i, v = cases.box_points(2, [2.5, -1.5], [12.5, 2.0], 1.0, gmin, gmax)
    mask.from_points(i, v, 2)
    top = plane2(0.0)
    top.boolean(mask, 1)
    run_both([mask, top], 3, 1, 0.4, 3.0, [dict(b=0.0), dict(kind=1, dir=1, k=1, b=-0.2, a_pos=-1.0, a_neg=0.0)])


def test_too_many_level_sets_and_unsupported_scheme():
    import viennals_amd as vls
    say("multi-material argument errors")
    d = [to_gpu(plane2(-2.0)), to_gpu(plane2(0.0))]
    a = vls.Advect()
    for g in d:
        a.insertNextLevelSet(g)
    a.setVelocityField(vls.VelocityField.constant(-1.0))
    a.setSpatialScheme(12)  # learned WENO5 takes one level set
    with pytest.raises(vls.LsgError, match="several level sets"):
        a.apply()


def test_reference_style_script_through_d3():
    """A ViennaLS Python script (sphere on a plane, etched) written against
    viennals.d3 runs unchanged on viennals_amd.d3 and gives the same band as
    the core API."""
    import viennals_amd as core
    from viennals_amd import d3 as vls
    say("reference-style d3 script: plane + sphere union, directional etch, 3 steps")

    def script(mod, Domain):
        bounds = [-12.0, 12.0, -12.0, 12.0, -12.0, 12.0]
        bcs = [mod.BoundaryConditionEnum.REFLECTIVE_BOUNDARY, mod.BoundaryConditionEnum.REFLECTIVE_BOUNDARY,
               mod.BoundaryConditionEnum.INFINITE_BOUNDARY]
        substrate = Domain(bounds, bcs, 1.0)
        mod.MakeGeometry(substrate, mod.Plane([0.0, 0.0, 0.0], [0.0, 0.0, 1.0])).apply()
        ball = Domain(bounds, bcs, 1.0)
        mod.MakeGeometry(ball, mod.Sphere([0.0, 0.0, 1.0], 4.0)).apply()
        mod.BooleanOperation(substrate, ball, mod.BooleanOperationEnum.UNION).apply()
        adv = mod.Advect(substrate, mod.VelocityField.normal_power(2, 1, -0.1, -1.0))
        adv.setSpatialScheme(mod.SpatialSchemeEnum.WENO_5TH_ORDER)
        adv.setTemporalScheme(mod.TemporalSchemeEnum.RUNGE_KUTTA_3RD_ORDER)
        adv.setTimeStepRatio(0.4)
        for _ in range(3):
            adv.apply()
        return substrate

    a = script(vls, vls.Domain)
    b = script(core, lambda bo, bc, dx: core.Domain(bo, bc, dx))
    ia, va = a.get_runs()
    ib, vb = b.get_runs()
    assert cases.bit_equal(ia, va, ib, vb)
    c = vls.Domain(a)  # copy constructor
    ic, vc = c.get_runs()
    assert cases.bit_equal(ia, va, ic, vc) and c.getNumberOfSegments() == 1
