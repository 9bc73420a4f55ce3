"""Point data through the advection path in the oracle
(PointData::translateFromMultiData via every newDataSourceIds of Expand,
Reduce, Prune, BooleanOperation and Advect::rebuildLS; updatePointData;
saveAdvectionVelocities).  Pinned by closed forms:

* a plane moving along z: updateLevelSet's Reduce(1) keeps the one active
  point of each column and rebuildLS gives every other point the data of its
  neighbour nearer the surface, so a whole column carries the data of its
  original active point;
* saveAdvectionVelocities on a plane: rate = dt * V / dx = cfl and
  dissipation 0 at every point (the rebuilt points copy their source);
* a union of two disjoint level sets keeps A's labels, B-sourced points
  take B's array of the same label, else 0;
* setUpdatePointData(false) drops the data at the rebuild;
* tests/Advection/Advection.cpp: "originalIDs" stay ids of original points
  near the point that carries them.
"""
import numpy as np
import pytest

import pyoracle as po
import cases

REF, INF = 0, 1


def plane3(z0=0.3):
    gmin, gmax, bc = [-5, -5, -12], [5, 5, 12], [REF, REF, INF]
    d = po.Domain(3, gmin, gmax, bc, 1.0)
    i, v = cases.plane_points(3, gmin, gmax, bc, 1.0, z0)
    d.from_points(i, v, 2)
    return d


def defined_points(d):
    idx, val = d.get()
    return idx[np.abs(val) < 1e300]


@pytest.mark.parametrize("spatial,v,time", [(0, 1.0, 0.0), (0, -1.0, 3.0), (11, 1.0, 2.0), (5, -0.7, 0.0)])
def test_plane_data_comes_from_the_active_point_of_the_column(spatial, v, time):
    # forward Euler: updateLevelSet's Reduce(1) keeps the active points (one
    # per column on this plane), and rebuildLS gives every other point the
    # data of its neighbour nearer the surface: the whole column carries the
    # data of its active point, step after step
    d = plane3()
    idx, val = d.get()
    m = np.abs(val) < 1e300
    pts0, v0 = idx[m], val[m]
    d.insert_data("z", pts0[:, 2].astype(float))
    d.insert_data("xyz", pts0.astype(float), vector=True)
    act = np.abs(v0) <= 0.5
    zact = {(a, b): c for a, b, c in pts0[act]}
    assert len(zact) == 121 and act.sum() == 121  # one active point per column
    p = po.make_params(spatial=spatial, temporal=0, cfl=0.4, time=time, vel=dict(b=v))
    st, r = po.advect(d, p)
    assert st == 0
    pts = defined_points(d)
    z = d.data("z")
    assert len(z) == len(pts)
    np.testing.assert_array_equal(z, [float(zact[(a, b)]) for a, b, _ in pts])
    xyz = d.data("xyz", vector=True)
    np.testing.assert_array_equal(xyz[:, :2], pts[:, :2].astype(float))  # the source lies in the same column
    np.testing.assert_array_equal(xyz[:, 2], z)


@pytest.mark.parametrize("temporal", [1, 2])
def test_plane_data_with_runge_kutta(temporal):
    # the combine takes the data of the current level set where both are
    # defined and of the step's start elsewhere: every value is still the z
    # of an original point of the same column
    d = plane3()
    pts0 = defined_points(d)
    d.insert_data("xyz", pts0.astype(float), vector=True)
    p = po.make_params(spatial=11, temporal=temporal, cfl=0.4, time=1.5, vel=dict(b=1.0))
    st, r = po.advect(d, p)
    assert st == 0
    pts = defined_points(d)
    xyz = d.data("xyz", vector=True)
    np.testing.assert_array_equal(xyz[:, :2], pts[:, :2].astype(float))
    assert set(xyz[:, 2]) <= set(pts0[:, 2].astype(float))


def test_saved_velocities_on_a_plane():
    d = plane3()
    p = po.make_params(spatial=0, temporal=0, cfl=0.4, time=0.0, vel=dict(b=1.0), save_velocities=True)
    st, r = po.advect(d, p)
    assert st == 0 and r["advected_time"] == pytest.approx(0.4, rel=1e-15)
    n = len(defined_points(d))
    assert d.data_labels() == ["AdvectionVelocities", "Dissipation"]
    np.testing.assert_array_equal(d.data("AdvectionVelocities"), np.full(n, 0.4))
    np.testing.assert_array_equal(d.data("Dissipation"), np.zeros(n))


def test_update_point_data_off_drops_the_data():
    d = plane3()
    d.insert_data("z", defined_points(d)[:, 2].astype(float))
    p = po.make_params(spatial=0, temporal=0, cfl=0.4, time=0.0, vel=dict(b=1.0),
                       update_point_data=False, save_velocities=True)
    st, r = po.advect(d, p)
    assert st == 0
    assert d.data_labels() == [] and d.data_labels(vector=True) == []


def test_union_takes_data_from_the_side_of_the_value():
    g = ([-20, -20], [20, 20], [INF, INF], 1.0)
    a = po.Domain(2, *g)
    a.make_sphere([-8.1, 0.2], 5.3)  # no grid point exactly on a surface (see the next test)
    b = po.Domain(2, *g)
    b.make_sphere([8.2, -0.1], 5.2)
    na, nb = len(defined_points(a)), len(defined_points(b))
    a.insert_data("m", np.full(na, 1.0))
    a.insert_data("onlyA", np.full(na, 7.0))
    b.insert_data("m", np.full(nb, 2.0))
    b.insert_data("onlyB", np.full(nb, 9.0))
    a.boolean(b, 1)  # UNION
    pts = defined_points(a)
    assert a.data_labels() == ["m", "onlyA"]
    m, oa = a.data("m"), a.data("onlyA")
    left = pts[:, 0] < 0
    np.testing.assert_array_equal(m, np.where(left, 1.0, 2.0))
    np.testing.assert_array_equal(oa, np.where(left, 7.0, 0.0))


def test_stray_zero_rewrite_records_no_source():
    # lsPrune.hpp:135-167: with removeStrayZeros a point of value exactly 0
    # whose neighbours are not monotone along some axis is rewritten without
    # recording a source id; every kept point records itself.  The data
    # array is then one value per kept point, shorter than the point count
    # (later operations read past its end in the reference; this
    # restatement gives 0 there).
    g = ([-20, -20], [20, 20], [INF, INF], 1.0)
    a = po.Domain(2, *g)
    a.make_sphere([0.0, 0.0], 5.0)
    idx, val = a.get()
    m = np.abs(val) < 1e300
    pts0, v0 = idx[m], val[m]
    a.insert_data("id", np.arange(len(pts0), dtype=float))
    zeros = {tuple(int(c) for c in p) for p, v in zip(pts0, v0) if v == 0.0}
    assert len(zeros) == 12  # (5, 0), (4, 3), (3, 4), ... lie exactly on the circle
    # at the four axis points both tangential neighbours lie outside (not
    # monotone); at (4, 3) and the like the values change sign along both axes
    rewritten = {(5, 0), (-5, 0), (0, 5), (0, -5)}
    a.prune(True)
    pts = defined_points(a)
    ids = a.data("id")
    pos0 = {tuple(int(c) for c in p): k for k, p in enumerate(pts0)}
    expect = [float(pos0[tuple(int(c) for c in p)]) for p in pts if tuple(int(c) for c in p) not in rewritten]
    assert len(ids) == len(pts) - 4
    np.testing.assert_array_equal(ids, expect)


def test_original_ids_stay_near():
    # tests/Advection/Advection.cpp:87-93: point data "originalIDs" = iota
    d = cases.sphere2d()
    pts0 = defined_points(d)
    d.insert_data("originalIDs", np.arange(len(pts0), dtype=float))
    p = po.make_params(spatial=11, temporal=2, cfl=0.4, time=2.0, vel=cases.VEL_EGG, save_velocities=True)
    st, r = po.advect(d, p)
    assert st == 0
    pts = defined_points(d)
    ids = d.data("originalIDs")
    assert len(ids) == len(pts)
    assert np.all(ids == np.round(ids)) and ids.min() >= 0 and ids.max() < len(pts0)
    # the surface moved by at most ~|v| t / dx cells; data travel with it
    dist = np.abs(pts - pts0[ids.astype(int)]).max(1)
    assert dist.max() <= 12
    assert "AdvectionVelocities" in d.data_labels() and "Dissipation" in d.data_labels()
