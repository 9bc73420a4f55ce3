"""Void detection (MarkVoidPoints with Advect's options,
lsMarkVoidPoints.hpp) and Advect::setIgnoreVoids in the oracle.  Pinned by:

* tests/MarkVoidPoints/MarkVoidPoints.cpp: a plane at y = 0.35 with two
  holes of radius 5.1 below it -- exactly the defined points below y = -2
  are void;
* connected components from scipy.ndimage.label on the sign field of the
  same grid (an independent labelling), on random blobs;
* with ignoreVoids a hole below a growing surface keeps its values while
  the surface moves (void points get velocity 0), without it the hole
  shrinks.
"""
import numpy as np
import pytest
from scipy import ndimage

import pyoracle as po
import cases

REF, INF = 0, 1


def defined(d):
    idx, val = d.get()
    m = np.abs(val) < 1e300
    return idx[m], val[m]


def plane_with_holes(holes, extent=15, z0=0.35):
    gmin, gmax, bc = [-extent, -extent], [extent, extent], [REF, INF]
    d = po.Domain(2, gmin, gmax, bc, 1.0)
    i, v = cases.plane_points(2, gmin, gmax, bc, 1.0, z0)
    d.from_points(i, v, 2)
    for o, r in holes:
        h = po.Domain(2, gmin, gmax, bc, 1.0)
        h.make_sphere(o, r)
        d.boolean(h, 2)  # RELATIVE_COMPLEMENT
    return d


def test_reference_mark_void_points():
    # tests/MarkVoidPoints/MarkVoidPoints.cpp:30-109
    d = plane_with_holes([([-8.0, -9.5], 5.1), ([8.0, -9.5], 5.1)])
    d.mark_void_points()
    idx, val = defined(d)
    m = d.data("VoidPointMarkers")
    assert len(m) == len(val)
    np.testing.assert_array_equal(m, (idx[:, 1] < -2).astype(float))


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_components_match_scipy_labels(seed):
    # random positive blobs (voids) inside a negative block below an open
    # top: label the sign field with scipy and apply the marker rule
    rng = np.random.default_rng(seed)
    gmin, gmax, bc = [-12, -12], [12, 12], [REF, INF]
    d = plane_with_holes([((float(x), float(y)), float(r)) for x, y, r in
                          zip(rng.uniform(-9, 9, 4), rng.uniform(-10, -3, 4), rng.uniform(1.2, 2.6, 4))],
                         extent=12, z0=0.3)
    d.mark_void_points()
    idx, val = defined(d)
    marks = d.data("VoidPointMarkers")
    # sign field on a box covering the grid in x and the band +-2 in y
    ylo, yhi = idx[:, 1].min() - 2, idx[:, 1].max() + 2
    xs, ys = np.arange(gmin[0], gmax[0] + 1), np.arange(ylo, yhi + 1)
    allIdx, allVal = d.get()
    sign = np.zeros((len(ys), len(xs)), bool)
    # value of every grid point: run-start list lookup (last entry at or before it)
    keys = allIdx[:, 1].astype(np.int64) * 10**6 + allIdx[:, 0]
    for a, y in enumerate(ys):
        k = np.searchsorted(keys, y * 10**6 + xs, side="right") - 1
        sign[a] = allVal[k] >= 0
    lab, n = ndimage.label(sign)
    labn, nn = ndimage.label(~sign)
    top = lab[-1, -1]
    assert top > 0  # the open side is up
    expect = []
    for (x, y), c in zip(idx, val):
        a, b = y - ylo, x - gmin[0]
        if c >= 0:
            expect.append(float(lab[a, b] != top))
        else:
            nbs = [(a, b + 1), (a + 1, b), (a, b - 1), (a - 1, b)]
            hit = any(0 <= p < len(ys) and 0 <= q < len(xs) and sign[p, q] and lab[p, q] == top for p, q in nbs)
            expect.append(float(not hit))
    np.testing.assert_array_equal(marks, expect)
    assert np.sum(marks) > 0


def test_ignore_voids_keeps_the_hole():
    # tests/VoidDetection/VoidDetection.cpp: plane y = 0 with a hole of radius
    # 3 at (0, -5), v = 1
    def run(ignore):
        d = plane_with_holes([([0.0, -5.0], 3.0)], extent=10, z0=0.0)
        idx0, val0 = defined(d)
        p = po.make_params(spatial=0, temporal=0, cfl=0.4, time=2.0, vel=dict(b=1.0), ignore_voids=ignore)
        st, r = po.advect(d, p)
        assert st == 0
        return (idx0, val0), defined(d), d
    (i0, v0), (i1, v1), d = run(True)
    hole0 = {tuple(p): v for p, v in zip(i0, v0) if p[1] < -1 and abs(v) <= 0.5}
    hole1 = {tuple(p): v for p, v in zip(i1, v1) if p[1] < -1 and abs(v) <= 0.5}
    assert hole0 and hole0 == hole1
    top = v1[i1[:, 1] > -1]
    assert len(top) > 0
    assert "VoidPointMarkers" in d.data_labels()
    (_, _), (j1, w1), _ = run(False)
    hole2 = {tuple(p): v for p, v in zip(j1, w1) if p[1] < -1 and abs(v) <= 0.5}
    assert hole2 != hole0
