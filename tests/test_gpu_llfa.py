"""LOCAL_LAX_FRIEDRICHS_ANALYTICAL_1ST_ORDER (enum 4) on the GPU: bit-exact
against the oracle (pinned in tests/test_oracle_llfa.py), one level set and
the SquareEtch-style materials (zero velocity and alpha on the mask)."""
import pytest

import cases
import pyoracle as po
from test_gpu_parity import assert_same, say, to_gpu

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("temporal", [0, 2])
def test_llfa_sphere_matches_oracle(temporal):
    import viennals_amd as vls
    say(f"LLFA sphere r=6 dx=0.5, alpha (0.7,0.3,1.1), temporal {temporal}")
    o = cases.sphere(r=6.0, dx=0.5)
    g = to_gpu(o)
    vel = dict(kind=1, dir=2, k=1, b=0.4, a_pos=0.5, a_neg=0.1, alpha=(0.7, 0.3, 1.1))
    p = po.make_params(spatial=4, temporal=temporal, cfl=0.4, time=1.5, vel=vel)
    st, r = po.advect(o, p)
    assert st == 0
    a = vls.Advect(g, vls.VelocityField(**vel))
    a.setSpatialScheme(4)
    a.setTemporalScheme(temporal)
    a.setTimeStepRatio(0.4)
    a.setAdvectionTime(1.5)
    a.apply()
    assert a.getAdvectedTime() == r["advected_time"] and a.getNumberOfTimeSteps() == r["steps"]
    assert_same(o, g, "LLFA sphere")


def test_llfa_square_etch_materials():
    """SquareEtch.cpp's analytical field: material 1 etches with -|n_z|,
    alpha |v| in z; the mask (material 0) does not move."""
    import viennals_amd as vls
    from test_gpu_multimaterial import run_both
    say("LLFA with a mask (SquareEtch analytical field), 3D, FE")
    L = 16
    gmin, gmax, bc = [-L, -L, -L], [L, L, L], [0, 0, 1]
    mask = po.Domain(3, gmin, gmax, bc, 1.0)
    i, v = cases.box_points(3, [-L + 0.5, -L + 0.5, -0.5], [-3.5, L - 0.5, 1.5], 1.0, gmin, gmax)
    mask.from_points(i, v, 2)
    top = po.Domain(3, gmin, gmax, bc, 1.0)
    i, v = cases.plane_points(3, gmin, gmax, bc, 1.0, 0.0)
    top.from_points(i, v, 2)
    top.boolean(mask, 1)
    mats = [dict(b=0.0), dict(kind=1, dir=2, k=1, b=0.0, a_pos=-1.0, a_neg=-1.0, alpha=(0.0, 0.0, 1.0))]
    run_both([mask, top], 4, 0, 0.4, 2.0, mats)
