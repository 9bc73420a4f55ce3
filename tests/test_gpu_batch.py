"""Batched advection (BatchAdvect, lsg_batch_*): many independent level sets
stepped with one launch per kernel must give exactly what each level set
gives on its own through Advect (which is bit-exact against the oracle)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def make_sets(vls):
    BC = vls.BoundaryConditionEnum
    sets, vels = [], []
    rng = np.random.default_rng(42)
    for r in (3.0, 5.5, 8.0, 4.2):
        d = vls.Domain(gridDelta=1.0, dim=3)
        c = rng.uniform(-3, 3, 3)
        vls.MakeGeometry(d, vls.Sphere(list(c), r)).apply()
        sets.append(d)
        vels.append(vls.VelocityField.constant(float(rng.uniform(0.2, 1.0))))
    for L, w, depth in ((24, 8, 10), (40, 12, 20)):
        d = vls.Domain([-L, L, -L, L, -L, L], [BC.REFLECTIVE_BOUNDARY, BC.REFLECTIVE_BOUNDARY,
                                               BC.INFINITE_BOUNDARY], 1.0)
        vls.MakeGeometry(d, vls.Plane([0.0, 0.0, 0.0], [0.0, 0.0, 1.0])).apply()
        box = vls.Domain([-L, L, -L, L, -L, L], [BC.REFLECTIVE_BOUNDARY, BC.REFLECTIVE_BOUNDARY,
                                                 BC.INFINITE_BOUNDARY], 1.0)
        vls.MakeGeometry(box, vls.Box([-w / 2, -w / 2, -float(depth)], [w / 2, w / 2, 1.0])).apply()
        vls.BooleanOperation(d, box, vls.BooleanOperationEnum.RELATIVE_COMPLEMENT).apply()
        sets.append(d)
        rate = float(rng.uniform(0.2, 1.0))
        vels.append(vls.VelocityField.normal_power(dir=2, k=1, b=-0.1 * rate, a_pos=-rate, a_neg=0.0))
    return sets, vels


def copies(vls, sets):
    out = []
    for d in sets:
        c = vls.Domain.from_index_bounds(d.D, d.gmin, d.gmax, d.bc, d.gridDelta)
        c.deepCopy(d)
        out.append(c)
    return out


def individual(vls, sets, vels, steps, spatial, temporal):
    times = []
    for d, v in zip(sets, vels):
        a = vls.Advect(d, v)
        a.setSpatialScheme(spatial)
        a.setTemporalScheme(temporal)
        a.setTimeStepRatio(0.4)
        t = 0.0
        for _ in range(steps):
            a.apply()
            t += a.getAdvectedTime()
        times.append(t)
    return np.array(times)


@pytest.mark.parametrize("spatial,temporal,steps", [(11, 2, 3), (0, 0, 4), (10, 1, 2), (2, 0, 3), (3, 2, 2),
                                                    (5, 0, 3), (6, 2, 2), (7, 0, 2), (8, 2, 2)])
def test_batch_matches_individual_advection(spatial, temporal, steps):
    import viennals_amd as vls
    sets, vels = make_sets(vls)
    ref = copies(vls, sets)
    b = vls.BatchAdvect(sets, vels, spatial=spatial, temporal=temporal, cfl=0.4)
    for _ in range(steps):
        b.apply()
    b.writeBack()
    tb = b.advectedTimes()
    tr = individual(vls, ref, vels, steps, spatial, temporal)
    assert np.array_equal(tb, tr), (tb, tr)
    for k, (d, r) in enumerate(zip(sets, ref)):
        gi, gv = d.get_runs()
        ri, rv = r.get_runs()
        assert np.array_equal(gi, ri), k
        assert np.array_equal(gv, rv), k
    pu, st = b.pointUpdates()
    assert st == steps and pu > 0


def test_batch_autotune_and_reframe():
    """Autotuning restores the band; a long growth run moves the frames."""
    import viennals_amd as vls
    sets, vels = make_sets(vls)
    sets, vels = sets[:3], [vls.VelocityField.constant(1.0)] * 3
    ref = copies(vls, sets)
    b = vls.BatchAdvect(sets, vels, spatial=11, temporal=2, cfl=0.4)
    blocks, ms = b.autotune()
    assert blocks in (256, 512, 1024, 2048, 4096) and ms > 0
    assert b.pointUpdates() == (0.0, 0)
    steps = 36  # growth of about 14 cells: the frames (margin 24) must move
    for _ in range(steps):
        b.apply()
    b.writeBack()
    tr = individual(vls, ref, vels, steps, 11, 2)
    assert np.array_equal(b.advectedTimes(), tr)
    for d, r in zip(sets, ref):
        gi, gv = d.get_runs()
        ri, rv = r.get_runs()
        assert np.array_equal(gi, ri) and np.array_equal(gv, rv)
