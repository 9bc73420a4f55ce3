"""The oracle's MakeGeometry planes and boxes (FromSurfaceMesh,
lsFromSurfaceMesh.hpp:243-480) pinned by closed forms: along the grid lines
crossing an element the points next to the crossing get their distance along
the line times 1 - 1e-7 (distanceEps * 1e-3, :445), clipped to +-1, the
smallest sign distance wins where lines of several axes reach a point, and
elements on the grid boundary are dropped.  The reference's golden fixture
(tests/BooleanOperationExactZero) is rebuilt from these shapes."""
import json
import os

import numpy as np
import pytest

import cases
import pyoracle as po
from viennals_amd.hrle import to_hrle_runs

HERE = os.path.dirname(os.path.abspath(__file__))
REF, INF, PER = cases.REF, cases.INF, cases.PER
F = 1.0 - 1e-4 * 1e-3  # the distance factor of FromSurfaceMesh


def defined(d):
    idx, val = d.get()
    m = np.abs(val) < 1e300
    return idx[m], val[m]


@pytest.mark.parametrize("z0,sign", [(0.0, 1.0), (0.0, -1.0), (0.3, 1.0), (-1.75, 1.0)])
def test_axis_plane_values(z0, sign):
    """Plane z = z0 with normal +-e_z over a reflective 8 x 8 grid: every
    (x, y) column holds the points floor(z0 - 1e-4) .. ceil(z0 + 1e-4) with
    value sign * (z - z0) * (1 - 1e-7)."""
    L = 4
    d = po.Domain(3, [-L, -L, -L], [L, L, L], [REF, REF, INF], 1.0)
    assert d.make_plane([0.0, 0.0, z0], [0.0, 0.0, sign]) == 0
    idx, val = defined(d)
    ks = np.arange(int(np.floor(z0 - 1e-4)), int(np.ceil(z0 + 1e-4)) + 1)
    assert len(val) == (2 * L + 1) ** 2 * len(ks)
    for k in ks:
        m = idx[:, 2] == k
        assert m.sum() == (2 * L + 1) ** 2
        real = k - z0 if sign > 0 else -(k - z0)
        np.testing.assert_array_equal(val[m], np.full(m.sum(), real * F))


def test_tilted_line_values_2d():
    """2D line x + 2 y = 0 (normal (1, 2)), x reflective on [-4, 4], y
    infinite: a point (i, k) near the line gets its distance along y,
    (k + i / 2) * (1 - 1e-7) -- the crossing along x is twice as far, so its
    sign distance loses.  The grid lines x = 0 and y = 0 pass through the
    element's centre, where the direction to the centre is 0 / 0: their sign
    distances are NaN, and std::sort of (index, (sign, distance)) pairs then
    orders a point's crossings by distance alone (:445, :455): at (-1, 0) the
    crossing along x (distance -1) wins, at (1, 0) the one along y (0.5)."""
    d = po.Domain(2, [-4, -8], [4, 8], [REF, INF], 1.0)
    assert d.make_plane([0.0, 0.0], [1.0, 2.0]) == 0
    idx, val = defined(d)
    assert len(val) > 0
    pts = {(int(i), int(k)): v for (i, k), v in zip(idx, val)}
    for (i, k), v in pts.items():
        if i == 0 or k == 0:
            continue
        assert v == pytest.approx((k + i / 2.0) * F, abs=1e-15), (i, k, v)
    assert pts[(-1, 0)] == -F and pts[(1, 0)] == 0.5 * F
    # every column holds the points around its crossing y = -i / 2
    for i in range(-4, 5):
        ks = sorted(idx[idx[:, 0] == i][:, 1])
        y = -i / 2.0
        assert ks == list(range(int(np.floor(y - 1e-4)), int(np.ceil(y + 1e-4)) + 1))


def test_box_2d_values():
    """Box [-2.5, 2.5]^2 (2D, both axes infinite on [-8, 8]): the points
    next to an edge hold +-0.5 (1 - 1e-7) (inside negative); corners are not
    reached by any grid line crossing an edge, so they stay undefined."""
    d = po.Domain(2, [-8, -8], [8, 8], [INF, INF], 1.0)
    d.make_box([-2.5, -2.5], [2.5, 2.5])
    idx, val = defined(d)
    pts = {tuple(p): v for p, v in zip(idx, val)}
    for j in range(-2, 3):
        for x, s in ((-3, 1), (-2, -1), (2, -1), (3, 1)):
            assert pts[(x, j)] == s * 0.5 * F
            assert pts[(j, x)] == s * 0.5 * F
    for c in ((-3, -3), (3, 3), (-3, 3), (3, -3)):
        assert c not in pts
    assert len(pts) == 4 * 5 * 2 - 4  # the four inner corners are reached twice


def test_box_faces_on_the_grid_boundary():
    """A box face on the boundary of a reflective axis is dropped (all nodes
    on the boundary, :311-329): the box [-8, 2] x [-2, 2] x [-6, 1] in a grid
    reflective on [-8, 8] along x has no points around x = -8 (row y = 0,
    z = -3, away from the top and bottom faces).  Along an
    infinite axis the grid ends at -+715827882, so a face on the user's bound
    stays: the bottom face z = -8 of a box reaching that bound is there."""
    d = po.Domain(3, [-8, -8, -8], [8, 8, 8], [REF, REF, INF], 1.0)
    d.make_box([-8.0, -2.0, -6.0], [2.0, 2.0, 1.0])
    idx, _ = defined(d)
    row = (idx[:, 1] == 0) & (idx[:, 2] == -3)
    assert sorted(idx[row][:, 0]) == [1, 2, 3]  # only around the face x = 2
    d = po.Domain(3, [-8, -8, -8], [8, 8, 8], [REF, REF, INF], 1.0)
    d.make_box([-2.0, -2.0, -8.0], [2.0, 2.0, 1.0])
    idx, val = defined(d)
    col = (idx[:, 0] == 0) & (idx[:, 1] == 0)
    assert sorted(idx[col][:, 2]) == [-9, -8, -7, 0, 1, 2]


def test_plane_error_cases():
    d = po.Domain(3, [-4] * 3, [4] * 3, [REF, INF, INF], 1.0)
    assert d.make_plane([0, 0, 0], [0, 0, 1]) == -2
    d = po.Domain(3, [-4] * 3, [4] * 3, [REF, REF, REF], 1.0)
    assert d.make_plane([0, 0, 0], [0, 0, 1]) == -4
    d = po.Domain(3, [-4] * 3, [4] * 3, [REF, REF, INF], 1.0)
    assert d.make_plane([0, 0, 0], [1, 0, 0]) == -3


def test_boolean_exact_zero_fixture_from_meshes():
    """tests/BooleanOperationExactZero/BooleanOperationExactZero.cpp:60-111
    with MakeGeometry's box and plane restated through FromSurfaceMesh: the
    run starts and values of the union equal result.hpp."""
    g = json.load(open(os.path.join(HERE, "golden", "boolean_exact_zero.json")))
    I = np.array(g["indices"])
    V = np.array(g["values"])
    gmin, gmax, bc = [-10, -10], [10, 10], [REF, INF]
    mask = po.Domain(2, gmin, gmax, bc, 1.0)
    mask.make_box([-5.0, 1.0], [5.0, 10.0])
    sub = po.Domain(2, gmin, gmax, bc, 1.0)
    assert sub.make_plane([0.0, 1.0], [0.0, 1.0]) == 0
    assert sub.boolean(mask, 1) == 0
    io, vo = to_hrle_runs(*sub.get(), [-10, -715827882], [10, 715827882])
    assert len(io) == len(I)
    assert np.array_equal(io, I)
    dfn = np.abs(V) < 1e300
    np.testing.assert_allclose(vo[dfn], V[dfn], atol=1e-4)
    assert np.all((vo[~dfn] < 0) == (V[~dfn] < 0))


GEOMETRY_CASES = {
    "trench_plane": (0, 3, [-128] * 3, [128] * 3, [REF, REF, INF], 1.0, [0.0, 0.0, 0.0], [0.0, 0.0, 1.0], 0),
    "trench_box": (1, 3, [-128] * 3, [128] * 3, [REF, REF, INF], 1.0, [-32.0, -32.0, -128.0], [32.0, 32.0, 1.0], 0),
    "tilted_plane": (0, 3, [-20, -20, -20], [20, 20, 20], [PER, REF, INF], 0.5, [0.3, -0.2, 1.1], [0.2, -0.5, 1.0], 0),
    "tilted_line_2d": (0, 2, [-16, -16], [16, 16], [REF, INF], 0.7, [0.1, 0.4], [-0.6, 1.0], 0),
    "box_periodic": (1, 3, [-10, -10, -10], [10, 10, 10], [PER, PER, INF], 1.0, [-10.0, -3.0, -4.0], [5.0, 10.0, 2.5], 0),
    "box_periodic_ignore": (1, 3, [-10, -10, -10], [10, 10, 10], [PER, PER, INF], 1.0, [-10.0, -3.0, -4.0],
                            [5.0, 10.0, 2.5], 3),
    "box_2d": (1, 2, [-12, -12], [12, 12], [REF, INF], 0.5, [-2.2, -3.1], [4.3, 1.7], 0),
}


@pytest.mark.parametrize("case", sorted(GEOMETRY_CASES))
def test_library_geometry_matches_oracle(case):
    """The library's host code for planes and boxes (viennals_amd/csrc/
    lsg_geometry.cpp, through the host emulation build) gives the oracle's
    defined points bit for bit; the GPU domain built from them is compared
    in tests/test_gpu_geometry.py."""
    import pyemul as pe
    kind, D, gmin, gmax, bc, dx, a, b, mask = GEOMETRY_CASES[case]
    d = po.Domain(D, gmin, gmax, bc, dx)
    if kind == 0:
        assert d.make_plane(a, b) == 0
    else:
        d.make_box(a, b, mask)
    io, vo = defined(d)
    ie, ve = pe.geometry(kind, D, gmin, gmax, bc, dx, a, b, mask)
    assert len(ve) == len(vo) > 0
    assert np.array_equal(io, ie)
    assert np.array_equal(vo.view(np.int64), ve.view(np.int64))
