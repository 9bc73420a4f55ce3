"""STENCIL_LOCAL_LAX_FRIEDRICHS_1ST_ORDER (enum 9,
lsStencilLocalLaxFriedrichsScalar.hpp) in the oracle.  Pinned by closed
forms and by the reference's own accuracy test:

* on a plane the central and one-sided differences agree, so the
  dissipation vanishes and the plane moves by v*dt exactly;
* for a constant velocity the alpha of direction k is max |v n_k| over the
  3^D box, and the time step uses the alphas of the LAST point in entry
  order (:385-397), not the maxima over all points;
* for v = a*max(n_x, 0) the Monti term lifts alpha_x to a (s^2 + c^2) = a and
  the Toifl term cancels the Osher term in y, so the alpha bound is dx / a;
* tests/StencilLocalLaxFriedrichs/StencilLocalLaxFriedrichs.cpp: a sphere of
  radius 1 grown for 0.5 lies within 0.035 of radius 1.5.
"""
import numpy as np
import pytest

import pyoracle as po
import cases

REF, INF = 0, 1
SLLF = 9


def plane3(z0, width):
    gmin, gmax, bc = [-6, -6, -12], [6, 6, 12], [REF, REF, INF]
    d = po.Domain(3, gmin, gmax, bc, 1.0)
    ks = np.arange(-12, 13)
    x, y, z = np.meshgrid(np.arange(-6, 7), np.arange(-6, 7), ks, indexing="ij")
    phi = (z.ravel() - z0).astype(float)
    m = np.abs(phi) <= width / 2
    idx = np.stack([x.ravel()[m], y.ravel()[m], z.ravel()[m]], 1).astype(np.int32)
    d.from_points(idx, phi[m], width)
    return d


@pytest.mark.parametrize("v", [1.0, -0.5])
def test_plane_moves_by_v_dt(v):
    # exact linear values over the whole stencil (band width 8 = the scheme's)
    d = plane3(0.25, 8)
    p = po.make_params(spatial=SLLF, temporal=0, cfl=0.4, time=0.0, vel=dict(b=v))
    st, r = po.advect(d, p)
    dt = 0.4 / abs(v)  # cfl / |H - D| with H = |v| |grad phi| = |v|, D = 0
    assert st == 0 and r["advected_time"] == pytest.approx(dt, rel=1e-15)
    idx, val = d.get()
    m = np.abs(val) <= 0.5
    z = idx[m, 2].astype(float)
    assert m.sum() > 0
    assert np.allclose(val[m], z - 0.25 - v * dt, atol=1e-12)
    assert d.width == 3  # rebuildLS final width for SLLF (lsAdvect.hpp:250-256)


def circle_band(r0, o, dx=1.0, n=40, width=8):
    """Exact signed distance of a circle (grid units) on |phi| <= width / 2."""
    x, y = np.meshgrid(np.arange(-n, n + 1), np.arange(-n, n + 1), indexing="ij")
    phi = (np.hypot(x * dx - o[0], y * dx - o[1]) - r0) / dx
    m = np.abs(phi) <= width / 2
    d = po.Domain(2, [-n, -n], [n, n], [INF, INF], dx)
    d.from_points(np.stack([x[m], y[m]], 1).astype(np.int32), phi[m], width)
    return d, lambda i, j: (np.hypot(i * dx - o[0], j * dx - o[1]) - r0) / dx


@pytest.mark.parametrize("r0,o", [(20.0, (0.1, 0.2)), (15.0, (0.3, -0.2)), (25.0, (0.0, 0.45))])
def test_alpha_of_last_point_bounds_dt(r0, o):
    # v = 1, cfl 2: the velocity bound (~2) lies above the alpha bound.  With
    # a constant velocity alpha_k = max over the box of |n_k|, where the
    # normal at box offset (ox, oy) takes d/dx over the row y + oy through
    # x - 1, x, x + 1 and d/dy over the column x + ox (:142-163); the time step
    # uses the alphas of the last active point in entry order (the top of the
    # circle, dt ~ 0.82).  Maxima over all points would give dt = 0.5.
    d, phi = circle_band(r0, o)
    idx, val = d.get()
    a = idx[np.abs(val) <= 0.5]
    X, Y = a[np.lexsort((a[:, 0], a[:, 1]))[-1]]
    ax = ay = 0.0
    for ox in (-1, 0, 1):
        for oy in (-1, 0, 1):
            gx = (phi(X + 1, Y + oy) - phi(X - 1, Y + oy)) / 2
            gy = (phi(X + ox, Y + 1) - phi(X + ox, Y - 1)) / 2
            nrm = np.hypot(gx, gy)
            ax, ay = max(ax, abs(gx / nrm)), max(ay, abs(gy / nrm))
    p = po.make_params(spatial=SLLF, temporal=0, cfl=2.0, time=0.0, vel=dict(b=1.0))
    st, r = po.advect(d, p)
    assert st == 0
    assert r["advected_time"] == pytest.approx(1.0 / (ax + ay), rel=1e-9)
    assert r["advected_time"] > 0.75


def test_monti_and_toifl_terms_fix_the_alphas():
    # v = a max(n_x, 0): alpha_x = |a s^2 + 0 + a c^2| = a, alpha_y =
    # |0 - a c s + a c s| = 0 at every box point with n_x > DN, so dt = dx / a
    # (up to the box points within DN of n_x = 0, relative 1e-4)
    a = 2.0
    for dx in (1.0, 0.5):
        d, _ = circle_band(10.0, (0.1, 0.2), dx, int(20 / dx))
        p = po.make_params(spatial=SLLF, temporal=0, cfl=2.0, time=0.0,
                           vel=dict(kind=1, dir=0, k=1, b=0.0, a_pos=a, a_neg=0.0))
        st, r = po.advect(d, p)
        assert st == 0
        assert r["advected_time"] == pytest.approx(dx / a, rel=1e-4)


def test_reference_sphere_accuracy():
    # tests/StencilLocalLaxFriedrichs/StencilLocalLaxFriedrichs.cpp:36-83:
    # sphere r = 1 at dx 0.1, infinite boundaries, v = 1 for t = 0.5
    d = po.Domain(3, [-30] * 3, [30] * 3, [INF] * 3, 0.1)
    d.make_sphere([0.0, 0.0, 0.0], 1.0)
    p = po.make_params(spatial=SLLF, temporal=0, time=0.5, vel=dict(b=1.0))
    st, r = po.advect(d, p)
    assert st == 0 and r["advected_time"] == pytest.approx(0.5, rel=1e-12)
    idx, val = d.get()
    m = np.abs(val) <= 0.5
    rad = np.sqrt((idx[m].astype(float) ** 2).sum(1)) * 0.1
    err = np.abs(rad - val[m] * 0.1 - 1.5)  # distance of the implied surface from r = 1.5
    assert err.mean() < 0.035 and err.max() < 0.1


def test_lower_level_sets_are_not_adjusted():
    # Advect::adjustLowerLayers is skipped for SLLF (lsAdvect.hpp:802-813)
    gmin, gmax, bc = [-8, -8, -12], [8, 8, 12], [REF, REF, INF]

    def make(z0):
        d = po.Domain(3, gmin, gmax, bc, 1.0)
        i, v = cases.plane_points(3, gmin, gmax, bc, 1.0, z0)
        d.from_points(i, v, 2)
        return d

    for spatial, kept in ((SLLF, True), (5, False)):
        lower, top = make(1.5), make(0.5)  # the lower level set reaches above the top
        before = lower.get()
        p = po.make_params(spatial=spatial, temporal=0, cfl=0.4, time=0.0, vel=dict(b=-1.0))
        st, r = po.advect_multi([lower, top], p)
        assert st == 0
        after = lower.get()
        same = cases.bit_equal(before[0], before[1], after[0], after[1])
        assert same == kept


def test_invert_negates_every_entry():
    # BooleanOperation INVERT (lsBooleanOperation.hpp:214-247): defined values
    # negated, the undefined signs swapped; twice is the identity
    d = cases.sphere2d()
    i0, v0 = d.get()
    d.boolean(d, 3)
    i1, v1 = d.get()
    assert np.array_equal(i0, i1) and np.array_equal(v1, -v0)
    d.boolean(d, 3)
    assert cases.bit_equal(i0, v0, *d.get())
