"""The single-pass band operations (one kBandOp / kBandOpSeg launch per
Expand cycle, rebuild, Reduce, Prune, Boolean and update), switched on in
process with lsg_set_single_pass_band_ops: bit-identical to the oracle on
the band operations and advection cases of test_gpu_parity, and batches
equal to individual Advect runs of the multi-kernel pipeline."""
import numpy as np
import pytest

import cases
import pyoracle as po
from test_gpu_parity import _advect_gpu, _geom, assert_same, say, to_gpu

pytestmark = pytest.mark.gpu


def _set_single_pass(on):
    from viennals_amd._lib import check, lib
    check(lib().lsg_set_single_pass_band_ops(1 if on else 0))


@pytest.fixture
def single_pass():
    _set_single_pass(True)
    yield
    _set_single_pass(False)


@pytest.mark.parametrize("op", ["reduce", "rebuild", "prune_rsz", "prune"])
def test_band_ops_sphere(single_pass, op):
    import viennals_amd as vls
    say(f"single-pass band op {op} on sphere")
    o = cases.sphere(r=6.0, dx=0.25)
    g = to_gpu(o)
    for w in (3, 5, 7):
        o.expand(w)
        vls.Expand(g, w).apply()
        assert_same(o, g, f"expand {w}")
    if op == "reduce":
        o.reduce(1)
        vls.Reduce(g, 1).apply()
    elif op == "rebuild":
        o.reduce(1)
        vls.Reduce(g, 1).apply()
        o.rebuild()
        vls._lib.check(vls._lib.lib().lsg_rebuild(g._h))
    else:
        rsz = op == "prune_rsz"
        o.prune(rsz)
        p = vls.Prune(g)
        p.setRemoveStrayZeros(rsz)
        p.apply()
    assert_same(o, g, op)


def test_boolean_trench(single_pass):
    import viennals_amd as vls
    say("single-pass boolean: plane minus box")
    L = 32
    gmin, gmax, bc = [-L, -L, -L], [L, L, L], [0, 0, 1]
    a = po.Domain(3, gmin, gmax, bc, 1.0)
    ip, vp = cases.plane_points(3, gmin, gmax, bc, 1.0, 0.0)
    a.from_points(ip, vp, 2)
    b = po.Domain(3, gmin, gmax, bc, 1.0)
    ib, vb = cases.box_points(3, [-8, -8, -16], [8, 8, 1], 1.0, gmin, gmax)
    b.from_points(ib, vb, 2)
    ga, gb = to_gpu(a), to_gpu(b)
    a.boolean(b, 2)
    vls.BooleanOperation(ga, gb, vls.BooleanOperationEnum.RELATIVE_COMPLEMENT).apply()
    assert_same(a, ga, "relcomp")


CASES = [
    ("sphere", 11, 2, 0.4, 0.3, cases.VEL_TRANSLATE_X),
    ("sphere", 3, 2, 0.4, 0.3, cases.VEL_EGG),
    ("trench", 11, 2, 0.4, 3.0, cases.VEL_ETCH_DIRECTIONAL),
    ("periodic", 8, 2, 0.45, 2.0, cases.VEL_NORMAL_SQ),
    ("sphere2d", 11, 2, 0.4, 1.0, dict(kind=0, b=1.0, vec=(0.5, 0.0, 0.0))),
]


@pytest.mark.parametrize("geom,spatial,temporal,cfl,time,vel", CASES)
def test_advect_parity(single_pass, geom, spatial, temporal, cfl, time, vel):
    say(f"single-pass advect {geom} spatial={spatial} temporal={temporal}")
    o = _geom(geom)
    g = to_gpu(o)
    p = po.make_params(spatial=spatial, temporal=temporal, cfl=cfl, time=time, vel=vel)
    st, r = po.advect(o, p)
    assert st == 0
    a = _advect_gpu(g, p)
    assert a.getNumberOfTimeSteps() == r["steps"]
    assert a.getAdvectedTime() == r["advected_time"]
    assert_same(o, g, f"single-pass advect {geom} s{spatial} t{temporal}")


def test_batch_matches_multi_kernel_advection():
    import viennals_amd as vls
    from test_gpu_batch import copies, individual, make_sets
    say("single-pass batch of 6 level sets, WENO5 RK3, 3 steps")
    sets, vels = make_sets(vls)
    ref = copies(vls, sets)
    tr = individual(vls, ref, vels, 3, 11, 2)  # multi-kernel pipeline
    _set_single_pass(True)
    try:
        b = vls.BatchAdvect(sets, vels, spatial=11, temporal=2, cfl=0.4)
        for _ in range(3):
            b.apply()
        b.writeBack()
        tb = b.advectedTimes()
    finally:
        _set_single_pass(False)
    assert np.array_equal(tb, tr), (tb, tr)
    for k, (d, r) in enumerate(zip(sets, ref)):
        gi, gv = d.get_runs()
        ri, rv = r.get_runs()
        assert np.array_equal(gi, ri), k
        assert np.array_equal(gv.view(np.int64), rv.view(np.int64)), k
