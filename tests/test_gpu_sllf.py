"""STENCIL_LOCAL_LAX_FRIEDRICHS_1ST_ORDER (enum 9) on the GPU: bit-exact
against the oracle (pinned in tests/test_oracle_sllf.py): the 3^D box
alphas, the time step from the last point's alphas, the rebuild width 3 and
the lower level sets left as they are."""
import numpy as np
import pytest

import cases
import pyoracle as po
from test_gpu_parity import assert_same, say, to_gpu

pytestmark = pytest.mark.gpu

SLLF = 9


def run_one(o, temporal, cfl, time, vel, tag):
    import viennals_amd as vls
    g = to_gpu(o)
    p = po.make_params(spatial=SLLF, temporal=temporal, cfl=cfl, time=time, vel=vel)
    st, r = po.advect(o, p)
    assert st == 0
    a = vls.Advect(g, vls.VelocityField(**vel))
    a.setSpatialScheme(SLLF)
    a.setTemporalScheme(temporal)
    a.setTimeStepRatio(cfl)
    a.setAdvectionTime(time)
    a.apply()
    assert a.getAdvectedTime() == r["advected_time"] and a.getNumberOfTimeSteps() == r["steps"], tag
    assert_same(o, g, tag)
    assert g.getLevelSetWidth() == o.width == 3
    return g


@pytest.mark.parametrize("temporal", [0, 2])
def test_sllf_sphere_directional(temporal):
    say(f"SLLF sphere r=6 dx=0.5, directional (egg) velocity, temporal {temporal}")
    run_one(cases.sphere(r=6.0, dx=0.5, origin=[0.3, 0.1, -0.2]), temporal, 0.4, 1.0, cases.VEL_EGG,
            f"SLLF sphere {temporal}")


def test_sllf_reference_sphere():
    # tests/StencilLocalLaxFriedrichs/StencilLocalLaxFriedrichs.cpp: r = 1,
    # dx 0.1, infinite boundaries, v = 1 for t = 0.5, forward Euler
    say("SLLF reference sphere r=1 dx=0.1 t=0.5")
    o = po.Domain(3, [-30] * 3, [30] * 3, [1] * 3, 0.1)
    o.make_sphere([0.0, 0.0, 0.0], 1.0)
    run_one(o, 0, 0.4999, 0.5, dict(b=1.0), "SLLF reference sphere")


@pytest.mark.parametrize("temporal", [1, 2])
def test_sllf_2d_and_periodic(temporal):
    say(f"SLLF 2D sphere + periodic wavy surface, temporal {temporal}")
    run_one(cases.sphere2d(), temporal, 0.4, 1.0, dict(b=1.0, vec=(0.3, 0.0, 0.0)), "SLLF 2D")
    run_one(cases.wavy_periodic(N=24), temporal, 0.4, 1.0, cases.VEL_NORMAL_SQ, "SLLF periodic")


@pytest.mark.parametrize("temporal", [0, 2])
def test_sllf_mask_keeps_lower_level_set(temporal):
    """Mask over a plane (2D), etch with the mask at rest: the scheme's
    material lookup per box and the lower level set is not intersected with
    the top (lsAdvect.hpp:802-813)."""
    from test_gpu_multimaterial import plane2, run_both
    say(f"SLLF mask over a plane (2D), temporal {temporal}")
    gmin, gmax, bc = [-16, -20], [16, 20], [0, 1]
    mask = po.Domain(2, gmin, gmax, bc, 1.0)
    i, v = cases.box_points(2, [2.5, -1.5], [12.5, 2.0], 1.0, gmin, gmax)
    mask.from_points(i, v, 2)
    top = plane2(0.0)
    top.boolean(mask, 1)
    before = mask.get()
    run_both([mask, top], SLLF, temporal, 0.4, 3.0, [dict(b=0.0), dict(b=-1.0)])
    after = mask.get()
    assert cases.bit_equal(before[0], before[1], after[0], after[1])


@pytest.mark.parametrize("temporal", [0, 2])
def test_sllf_host_velocity_matches_oracle(temporal):
    """SLLF with a host VelocityField subclass computing the egg velocity
    (same operations as the built-in functor): the scheme's call, the 3^D
    box calls and the calls at the box normals moved by -DN and +DN per axis
    all come from the host -- bit-exact against the oracle's built-in run,
    so the perturbed normals and DN are the reference's."""
    import viennals_amd as vls
    vel = dict(cases.VEL_EGG, vec=(0.2, 0.0, -0.1))
    say(f"SLLF host velocity (egg + vector), temporal {temporal}")

    class Egg(vls.HostVelocityField):
        def getScalarVelocity(self, coordinate, material, normal, pointId):
            c = normal[0]
            a = vel["a_pos"] if c > 0 else vel["a_neg"]
            ac = abs(c)
            return vel["b"] + a * (ac * ac)

        def getVectorVelocity(self, coordinate, material, normal, pointId):
            return vel["vec"]

    o = cases.sphere(r=4.0, dx=0.5, origin=[0.3, 0.1, -0.2])
    g = to_gpu(o)
    p = po.make_params(spatial=SLLF, temporal=temporal, cfl=0.4, time=0.6, vel=vel)
    st, r = po.advect(o, p)
    assert st == 0
    a = vls.Advect(g, Egg())
    a.setSpatialScheme(SLLF)
    a.setTemporalScheme(temporal)
    a.setTimeStepRatio(0.4)
    a.setAdvectionTime(0.6)
    a.apply()
    assert a.getAdvectedTime() == r["advected_time"] and a.getNumberOfTimeSteps() == r["steps"]
    assert_same(o, g, "SLLF host velocity")


def _oracle_prepare(levelSets, isDepo):
    """PrepareStencilLocalLaxFriedrichs (lsStencilLocalLaxFriedrichsScalar.hpp:
    423-479) composed from the oracle's band operations."""
    n = len(levelSets)
    isDepo = (list(isDepo) + [False] * n)[:n]
    g = levelSets[0]
    top = po.Domain(g.D, g.gmin, g.gmax, g.bc, g.delta)
    top.set(g.get()[0][:1], np.array([np.finfo(float).max]), 1)  # the grid's minimum point only
    layer, above = n - 1, False
    for k in range(n):
        mask = levelSets[n - 1 - k]
        if isDepo[int(isDepo[n - 1 - k])]:
            if not above:
                layer, above = n - 1 - k, True
        else:
            if above:
                la = levelSets[layer].copy()
                la.boolean(mask, 2)
                top.boolean(la, 1)
                top.prune(False)
            above = False
    if above:
        top.boolean(levelSets[layer], 1)
    top.boolean(top, 3)
    levelSets[-1].set(*top.get(), top.width)


def test_sllf_layer_wrapping_epitaxy_like():
    """examples/Epitaxy/Epitaxy.py's flow with built-in velocities: mask plane,
    fin box united with the mask, PrepareStencilLocalLaxFriedrichs, SLLF with
    two materials, FinalizeStencilLocalLaxFriedrichs -- bit-exact against
    the oracle at every stage."""
    import viennals_amd as vls
    say("SLLF layer wrapping (Epitaxy-like, 3D)")
    n = 20
    gmin, gmax, bc = [-n, -n, -n], [n, n, n], [0, 0, 1]
    dx = 0.05
    mask = po.Domain(3, gmin, gmax, bc, dx)
    i, v = cases.plane_points(3, gmin, gmax, bc, dx, 0.0)
    mask.from_points(i, v, 2)
    sub = po.Domain(3, gmin, gmax, bc, dx)
    i, v = cases.box_points(3, [-0.25, -0.25, 0.0], [0.25, 0.25, 0.2], dx, gmin, gmax)
    sub.from_points(i, v, 2)
    sub.boolean(mask, 1)
    gm, gs = to_gpu(mask), to_gpu(sub)
    _oracle_prepare([mask, sub], [False, True])
    vls.PrepareStencilLocalLaxFriedrichs([gm, gs], [False, True])
    assert_same(sub, gs, "prepared top")
    mats = [dict(b=0.0), dict(kind=1, dir=2, k=1, b=-0.2, a_pos=-0.3, a_neg=-0.3)]
    p = po.make_params(spatial=SLLF, temporal=0, cfl=0.4, time=0.1, vel=mats[-1], mat_vels=mats)
    st, r = po.advect_multi([mask, sub], p)
    assert st == 0
    a = vls.Advect()
    a.insertNextLevelSet(gm)
    a.insertNextLevelSet(gs)
    a.setVelocityField(vls.MaterialVelocityField([vls.VelocityField(**m) for m in mats]))
    a.setSpatialScheme(SLLF)
    a.setTimeStepRatio(0.4)
    a.setAdvectionTime(0.1)
    a.apply()
    assert a.getAdvectedTime() == r["advected_time"]
    assert_same(sub, gs, "advected top")
    assert_same(mask, gm, "mask")
    sub.boolean(sub, 3)
    sub.boolean(mask, 1)
    vls.FinalizeStencilLocalLaxFriedrichs([gm, gs])
    assert_same(sub, gs, "finalized")
