"""The oracle's CompareVolume (lsCompareVolume.hpp:176-289) pinned by closed
forms: planes a known number of cell layers apart, the inside rule
"corner average <= 0" (a plane through cell centres), symmetry, identity,
the width-3 expansion of narrow inputs and a sphere shell against its
analytic volume."""
import numpy as np
import pytest

import cases
import pyoracle as po

REF, INF = cases.REF, cases.INF


def plane(z0, L=8, dx=1.0, width=2):
    gmin, gmax, bc = [-L, -L, -L], [L, L, L], [REF, REF, INF]
    d = po.Domain(3, gmin, gmax, bc, dx)
    ip, vp = cases.plane_points(3, gmin, gmax, bc, dx, z0)
    d.from_points(ip, vp, 2)
    if width > 2:
        d.expand(width)
    return d


def test_planes_two_layers_apart():
    """phi = z - z0: a cell [k, k+1] has corner average k + 0.5 - z0, so
    planes at z = 0 and z = 2 differ in the layers k = 0 and 1: 2 x 16 x 16
    cells of the reflective 16 x 16 lateral grid."""
    n, vol = po.compare_volume(plane(0.0), plane(2.0))
    assert n == 2 * 16 * 16
    assert vol == n * 1.0


def test_inside_rule_is_less_or_equal():
    """A plane at z = 0.5 has corner average exactly 0 in the cells k = 0:
    they count as inside ("<= 0", lsCompareVolume.hpp:271-272), so they
    differ from the plane at z = 0 -- a strict "< 0" would give no
    difference at all."""
    n, _ = po.compare_volume(plane(0.5), plane(0.0))
    assert n == 16 * 16


def test_symmetry_identity_and_grid_delta():
    a, b = plane(0.0, dx=0.5), plane(1.5, dx=0.5)
    n_ab, v_ab = po.compare_volume(a, b)
    n_ba, v_ba = po.compare_volume(b, a)
    assert n_ab == n_ba == 3 * 16 * 16  # z0 = 1.5 is 3 cells of 0.5
    assert v_ab == pytest.approx(n_ab * 0.125)
    assert po.compare_volume(a, a)[0] == 0


def test_narrow_inputs_are_expanded_to_width_three():
    """Inputs narrower than 3 are expanded on copies first (:195-220): the
    result equals that of the width-3 expansions, and the inputs keep their
    width."""
    a, b = cases.sphere(r=6.0, dx=0.5), cases.sphere(r=7.3, dx=0.5, origin=[0.3, 0.0, 0.1])
    a.reduce(1)
    b.reduce(1)
    n1, _ = po.compare_volume(a, b)
    assert a.width == 1 and b.width == 1
    a3, b3 = a.copy(), b.copy()
    a3.expand(3)
    b3.expand(3)
    assert po.compare_volume(a3, b3)[0] == n1


def test_sphere_shell_volume():
    """Two concentric spheres: the mismatch approaches 4/3 pi (R^3 - r^3)."""
    a, b = cases.sphere(r=6.0, dx=0.25), cases.sphere(r=8.0, dx=0.25)
    _, vol = po.compare_volume(a, b)
    exact = 4.0 / 3.0 * np.pi * (8.0 ** 3 - 6.0 ** 3)
    assert vol == pytest.approx(exact, rel=0.01)
