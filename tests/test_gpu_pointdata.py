"""Point data on the GPU (Domain::getPointData carried through Expand, Reduce,
Prune, BooleanOperation, Advect::rebuildLS; updatePointData;
saveAdvectionVelocities): every array bit-exact against the oracle (pinned in
tests/test_oracle_pointdata.py), on the reference's Advection-test
configuration ("originalIDs", saved velocities), planes, booleans, the
stray-zero rewrite and several level sets."""
import numpy as np
import pytest

import cases
import pyoracle as po
from test_gpu_parity import assert_same, say, to_gpu

pytestmark = pytest.mark.gpu


def defined_count(o):
    idx, val = o.get()
    return int((np.abs(val) < 1e300).sum())


def seed(o, g, rng):
    """The same arrays on the oracle domain and its GPU copy."""
    n = defined_count(o)
    s = rng.standard_normal(n)
    v = rng.standard_normal((n, 3))
    ids = np.arange(n, dtype=float)
    for lab, arr, vec in (("originalIDs", ids, False), ("noise", s, False), ("vec", v, True)):
        o.insert_data(lab, arr, vector=vec)
        if vec:
            g.getPointData().insertNextVectorData(arr, lab)
        else:
            g.getPointData().insertNextScalarData(arr, lab)


def assert_same_data(o, g, tag):
    pd = g.getPointData()
    for vec in (False, True):
        labels = o.data_labels(vector=vec)
        n = pd.getVectorDataSize() if vec else pd.getScalarDataSize()
        got = [(pd.getVectorDataLabel(i) if vec else pd.getScalarDataLabel(i)) for i in range(n)]
        assert got == labels, (tag, got, labels)
        for lab in labels:
            a = o.data(lab, vector=vec)
            b = pd.getVectorData(lab) if vec else pd.getScalarData(lab)
            assert a.shape == b.shape, (tag, lab, a.shape, b.shape)
            assert np.array_equal(a.view(np.int64), b.view(np.int64)), f"{tag}: array {lab} differs"


def run_advect(o, g, spatial, temporal, time, vel, save=False, update=True):
    import viennals_amd as vls
    p = po.make_params(spatial=spatial, temporal=temporal, cfl=0.4, time=time, vel=vel,
                       save_velocities=save, update_point_data=update)
    st, r = po.advect(o, p)
    assert st == 0
    a = vls.Advect(g, vls.VelocityField(**vel))
    a.setSpatialScheme(spatial)
    a.setTemporalScheme(temporal)
    a.setTimeStepRatio(0.4)
    a.setAdvectionTime(time)
    a.setSaveAdvectionVelocities(save)
    a.setUpdatePointData(update)
    a.apply()
    assert a.getAdvectedTime() == r["advected_time"]


@pytest.mark.parametrize("spatial,temporal", [(0, 0), (11, 2), (8, 1), (9, 0)])
def test_reference_advection_config(spatial, temporal):
    """tests/Advection/Advection.cpp:66-100: sphere r = 7.3 at (5, 0, 0),
    "originalIDs" as point data, saveAdvectionVelocities on."""
    say(f"point data: sphere r=7.3 spatial {spatial} temporal {temporal}, saved velocities")
    o = cases.sphere(r=7.3, dx=0.5, origin=[5.0, 0.0, 0.0])
    g = to_gpu(o)
    seed(o, g, np.random.default_rng(spatial * 10 + temporal))
    run_advect(o, g, spatial, temporal, 1.0, cases.VEL_EGG, save=True)
    assert_same(o, g, "sphere")
    assert_same_data(o, g, "sphere")


@pytest.mark.parametrize("temporal", [0, 2])
def test_plane_and_update_off(temporal):
    say(f"point data: trench, temporal {temporal}, then updatePointData off")
    o = cases.trench(L=16, width=8, depth=8)
    g = to_gpu(o)
    seed(o, g, np.random.default_rng(5))
    run_advect(o, g, 11, temporal, 1.0, cases.VEL_ETCH_DIRECTIONAL, save=True)
    assert_same_data(o, g, "trench")
    run_advect(o, g, 11, temporal, 0.5, cases.VEL_ETCH_DIRECTIONAL, update=False)
    assert_same_data(o, g, "trench, update off")
    assert g.getPointData().empty()


def test_band_operations_and_stray_zeros():
    import viennals_amd as vls
    say("point data: expand / reduce / union / prune with stray zeros (2D)")
    g0 = ([-20, -20], [20, 20], [1, 1], 1.0)
    a = po.Domain(2, *g0)
    a.make_sphere([0.0, 0.0], 5.0)  # exact zeros on the axes: the stray-zero rewrite records no source
    b = po.Domain(2, *g0)
    b.make_sphere([9.3, 0.4], 5.1)
    ga, gb = to_gpu(a), to_gpu(b)
    rng = np.random.default_rng(3)
    seed(a, ga, rng)
    nb = defined_count(b)
    arr = rng.standard_normal(nb)
    b.insert_data("noise", arr)
    gb.getPointData().insertNextScalarData(arr, "noise")
    a.expand(5)
    vls.Expand(ga, 5).apply()
    assert_same_data(a, ga, "expand")
    a.reduce(3)
    vls.Reduce(ga, 3).apply()
    assert_same_data(a, ga, "reduce")
    a.boolean(b, 1)
    vls.BooleanOperation(ga, gb, vls.BooleanOperationEnum.UNION).apply()
    assert_same(a, ga, "union")
    assert_same_data(a, ga, "union")
    c = po.Domain(2, *g0)
    c.make_sphere([0.0, 0.0], 5.0)
    gc = to_gpu(c)
    seed(c, gc, rng)
    c.prune(True)
    p = vls.Prune(gc)
    p.setRemoveStrayZeros(True)
    p.apply()
    assert_same(c, gc, "prune")
    assert_same_data(c, gc, "prune with stray zeros")


@pytest.mark.parametrize("temporal", [0, 2])
def test_several_level_sets(temporal):
    import viennals_amd as vls
    from test_gpu_multimaterial import plane2
    say(f"point data: mask over a plane (2D), temporal {temporal}, saved velocities")
    gmin, gmax, bc = [-16, -20], [16, 20], [0, 1]
    mask = po.Domain(2, gmin, gmax, bc, 1.0)
    i, v = cases.box_points(2, [2.5, -1.5], [12.5, 2.0], 1.0, gmin, gmax)
    mask.from_points(i, v, 2)
    top = plane2(0.0)
    top.boolean(mask, 1)
    gm, gt = to_gpu(mask), to_gpu(top)
    rng = np.random.default_rng(11)
    seed(mask, gm, rng)
    seed(top, gt, rng)
    mats = [dict(b=0.0), dict(b=-1.0)]
    p = po.make_params(spatial=0, temporal=temporal, cfl=0.4, time=3.0, vel=mats[-1], mat_vels=mats,
                       save_velocities=True)
    st, r = po.advect_multi([mask, top], p)
    assert st == 0
    a = vls.Advect()
    a.insertNextLevelSet(gm)
    a.insertNextLevelSet(gt)
    a.setVelocityField(vls.MaterialVelocityField([vls.VelocityField(**m) for m in mats]))
    a.setTemporalScheme(temporal)
    a.setTimeStepRatio(0.4)
    a.setAdvectionTime(3.0)
    a.setSaveAdvectionVelocities(True)
    a.apply()
    assert a.getAdvectedTime() == r["advected_time"]
    assert_same(top, gt, "top")
    assert_same(mask, gm, "mask")
    assert_same_data(top, gt, "top")
    assert_same_data(mask, gm, "mask")
