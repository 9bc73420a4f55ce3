"""GPU surface points and CompareChamfer against the CPU judges
(tests/judges.py, pinned in tests/test_oracle_pins.py): the same surface
point set bit for bit, the Chamfer distances within 1e-12 relative (the sums
run in another order), and the reference's SLLF accuracy test judged on the
GPU."""
import numpy as np
import pytest

import cases
import judges
import pyoracle as po
from test_gpu_parity import say, to_gpu

pytestmark = pytest.mark.gpu


def sorted_rows(a):
    return a[np.lexsort(a.T[::-1])]


@pytest.mark.parametrize("geom", ["sphere", "trench", "sphere2d"])
def test_surface_points_match_the_judge(geom):
    import viennals_amd as vls
    say(f"surface points: {geom}")
    o = {"sphere": lambda: cases.sphere(r=6.3, dx=0.5, origin=[0.2, 0.1, -0.3]),
         "trench": lambda: cases.trench(L=16, width=8, depth=8),
         "sphere2d": lambda: cases.sphere2d()}[geom]()
    g = to_gpu(o)
    ref = judges.surface_points(*o.get(), o.delta)
    got = vls.surface_points(g)
    assert got.shape == ref.shape
    assert np.array_equal(sorted_rows(got), sorted_rows(ref))


def test_chamfer_matches_the_judge():
    import viennals_amd as vls
    say("CompareChamfer: spheres of radii 5 and 5.4, an advected sphere, planes")
    a = cases.sphere(r=5.0, dx=0.5, origin=[0.1, 0.0, 0.2])
    b = cases.sphere(r=5.4, dx=0.5, origin=[0.1, 0.0, 0.2])
    c = cases.sphere(r=5.0, dx=0.5, origin=[0.1, 0.0, 0.2])
    p = po.make_params(spatial=11, temporal=2, cfl=0.4, time=1.0, vel=cases.VEL_EGG)
    assert po.advect(c, p)[0] == 0
    for x, y in ((a, b), (a, c), (b, c), (a, a)):
        gx, gy = to_gpu(x), to_gpu(y)
        cc = vls.CompareChamfer(gx, gy)
        cc.apply()
        P = judges.surface_points(*x.get(), x.delta)
        Q = judges.surface_points(*y.get(), y.delta)
        assert cc.getNumTargetPoints() == len(P) and cc.getNumSamplePoints() == len(Q)
        ref = judges.chamfer(P, Q)
        assert cc.getChamferDistance() == pytest.approx(ref, rel=1e-12, abs=1e-15)
        from scipy.spatial import cKDTree
        f, _ = cKDTree(Q).query(P)
        bwd, _ = cKDTree(P).query(Q)
        assert cc.getForwardDistance() == pytest.approx(f.mean(), rel=1e-12, abs=1e-15)
        assert cc.getBackwardDistance() == pytest.approx(bwd.mean(), rel=1e-12, abs=1e-15)
        assert cc.getMaxDistance() == pytest.approx(max(f.max(), bwd.max()), rel=1e-12, abs=1e-15)
        rms = np.sqrt((np.sum(f ** 2) + np.sum(bwd ** 2)) / (len(f) + len(bwd)))
        assert cc.getRMSChamferDistance() == pytest.approx(rms, rel=1e-12, abs=1e-15)


def test_reference_sllf_test_judged_on_the_gpu():
    # tests/StencilLocalLaxFriedrichs/StencilLocalLaxFriedrichs.cpp:36-84 with
    # the advection and the Chamfer distance on the GPU
    import viennals_amd as vls
    say("SLLF sphere growth judged by the GPU CompareChamfer")
    s = vls.Domain([-3.0, 3.0] * 3, [vls.BoundaryConditionEnum.INFINITE_BOUNDARY] * 3, 0.1)
    vls.MakeGeometry(s, vls.Sphere([0.0, 0.0, 0.0], 1.0)).apply()
    a = vls.Advect(s, vls.VelocityField(b=1.0))
    a.setAdvectionTime(0.5)
    a.setSpatialScheme(vls.SpatialSchemeEnum.STENCIL_LOCAL_LAX_FRIEDRICHS_1ST_ORDER)
    a.apply()
    ref = vls.Domain([-3.0, 3.0] * 3, [vls.BoundaryConditionEnum.INFINITE_BOUNDARY] * 3, 0.1)
    vls.MakeGeometry(ref, vls.Sphere([0.0, 0.0, 0.0], 1.5)).apply()
    cc = vls.CompareChamfer(ref, s)
    cc.apply()
    assert cc.getChamferDistance() < 0.035


def test_chamfer_expands_narrow_level_sets_to_width_two():
    """lsCompareChamfer.hpp:118-134: a level set narrower than 2 is expanded
    (on a copy) before its surface is extracted, so a Reduce(1) input gives
    the distances of its width-2 expansion and stays unchanged itself."""
    import viennals_amd as vls
    say("CompareChamfer of Reduce(1) inputs")
    a = cases.sphere(r=5.0, dx=0.5, origin=[0.1, 0.0, 0.2])
    b = cases.sphere(r=5.4, dx=0.5, origin=[0.1, 0.0, 0.2])
    ga, gb = to_gpu(a), to_gpu(b)
    vls.Reduce(ga, 1).apply()
    vls.Reduce(gb, 1).apply()
    ia, va = ga.get_runs()
    wa, wb = to_gpu(a), to_gpu(b)  # width-2 copies of the same surfaces
    for g in (wa, wb):
        vls.Reduce(g, 1).apply()
        vls.Expand(g, 2).apply()
    narrow = vls.CompareChamfer(ga, gb)
    narrow.apply()
    wide = vls.CompareChamfer(wa, wb)
    wide.apply()
    assert narrow.getNumTargetPoints() == wide.getNumTargetPoints() > 0
    assert narrow.getNumSamplePoints() == wide.getNumSamplePoints() > 0
    for f in ("getForwardDistance", "getBackwardDistance", "getChamferDistance", "getRMSChamferDistance",
              "getMaxDistance"):
        assert getattr(narrow, f)() == getattr(wide, f)(), f
    ia2, va2 = ga.get_runs()
    assert np.array_equal(ia, ia2) and np.array_equal(va.view(np.int64), va2.view(np.int64))
    assert ga.getLevelSetWidth() == 1
