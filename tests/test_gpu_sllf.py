"""STENCIL_LOCAL_LAX_FRIEDRICHS_1ST_ORDER (enum 9) on the GPU: bit-exact
against the oracle (pinned in tests/test_oracle_sllf.py): the 3^D box
alphas, the time step from the last point's alphas, the rebuild width 3 and
the lower level sets left as they are."""
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


def test_sllf_rejects_host_velocity():
    import viennals_amd as vls
    say("SLLF with a host velocity field: error")
    d = to_gpu(cases.sphere(r=3.0, dx=0.5))
    a = vls.Advect(d, vls.HostVelocityField())
    a.setSpatialScheme(SLLF)
    with pytest.raises(vls.LsgError, match="host-evaluated"):
        a.apply()
