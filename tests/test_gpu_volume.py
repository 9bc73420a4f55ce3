"""CompareVolume on the GPU (lsg_compare_volume) against the oracle's
restatement (pinned in tests/test_oracle_volume.py): the same cell count on
planes, spheres, an advected trench and 2D circles, and narrow inputs
expanded to width 3 without being changed."""
import numpy as np
import pytest

import cases
import pyoracle as po
from test_gpu_parity import say, to_gpu
from test_oracle_volume import plane

pytestmark = pytest.mark.gpu


def gpu_volume(a, b):
    import viennals_amd as vls
    cv = vls.CompareVolume(to_gpu(a), to_gpu(b))
    cv.apply()
    return cv.getCellCount(), cv.getVolumeMismatch()


def pairs():
    c = cases.trench(L=24, width=10, depth=12)
    c2 = c.copy()
    p = po.make_params(spatial=11, temporal=2, cfl=0.4, time=2.0, vel=cases.VEL_ETCH_DIRECTIONAL)
    assert po.advect(c2, p)[0] == 0
    return {
        "planes": (plane(0.0), plane(2.0)),
        "plane_through_centres": (plane(0.5), plane(0.0)),
        "spheres": (cases.sphere(r=6.0, dx=0.5), cases.sphere(r=7.3, dx=0.5, origin=[0.3, 0.0, 0.1])),
        "trench_advected": (c, c2),
        "circles_2d": (cases.sphere2d(r=5.0), cases.sphere2d(r=6.5)),
    }


@pytest.mark.parametrize("case", ["planes", "plane_through_centres", "spheres", "trench_advected", "circles_2d"])
def test_volume_matches_oracle(case):
    say(f"CompareVolume: {case}")
    a, b = pairs()[case]
    n_o, v_o = po.compare_volume(a, b)
    n_g, v_g = gpu_volume(a, b)
    assert n_g == n_o, f"{case}: gpu {n_g} cells, oracle {n_o}"
    assert v_g == v_o
    assert n_o > 0


def test_narrow_inputs_keep_their_band():
    import viennals_amd as vls
    say("CompareVolume: width-1 inputs")
    a, b = cases.sphere(r=6.0, dx=0.5), cases.sphere(r=7.3, dx=0.5)
    a.reduce(1)
    b.reduce(1)
    ga, gb = to_gpu(a), to_gpu(b)
    ia, va = ga.get_runs()
    cv = vls.CompareVolume(ga, gb)
    cv.apply()
    assert cv.getCellCount() == po.compare_volume(a, b)[0]
    ia2, va2 = ga.get_runs()
    assert np.array_equal(ia, ia2) and np.array_equal(va.view(np.int64), va2.view(np.int64))
    assert ga.getLevelSetWidth() == 1
