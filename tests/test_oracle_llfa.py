"""LOCAL_LAX_FRIEDRICHS_ANALYTICAL_1ST_ORDER (enum 4,
lsLocalLaxFriedrichsAnalytical.hpp) in the oracle, with the built-in
functor's analytical alphas (constant per direction).  Pinned by closed
forms: on a plane the one-sided differences agree, so the dissipation
vanishes and the plane moves by v*t exactly; the alphas bound the time step
by 1 / sum(alpha_d / dx) (:231-241)."""
import numpy as np
import pytest

import pyoracle as po
import cases

REF, INF = 0, 1


def plane3(z0):
    gmin, gmax, bc = [-6, -6, -12], [6, 6, 12], [REF, REF, INF]
    d = po.Domain(3, gmin, gmax, bc, 1.0)
    i, v = cases.plane_points(3, gmin, gmax, bc, 1.0, z0)
    d.from_points(i, v, 2)
    return d


@pytest.mark.parametrize("alpha_z,expected_dt", [(0.0, 0.4), (1.0, 0.4), (4.0, 0.25)])
def test_plane_moves_by_vt_and_alpha_bounds_dt(alpha_z, expected_dt):
    d = plane3(0.0)
    p = po.make_params(spatial=4, temporal=0, cfl=0.4, time=0.0, vel=dict(b=1.0, alpha=(0.0, 0.0, alpha_z)))
    st, r = po.advect(d, p)
    assert st == 0 and abs(r["advected_time"] - expected_dt) < 1e-15
    idx, val = d.get()
    m = np.abs(val) < 1e300
    z = idx[m, 2].astype(float)
    assert np.allclose(val[m], z - expected_dt, atol=1e-12)  # phi = z - v t: the plane moved up by v t


def test_alpha_dissipation_acts_on_a_kink():
    # a corner (kink) has different one-sided slopes: the dissipation term
    # alpha*(gradNeg - gradPos)/2 changes the result with alpha, and only in
    # the direction whose alpha is set
    def run(alpha):
        d = po.Domain(2, [-12, -12], [12, 12], [INF, INF], 1.0)
        i, v = cases.box_points(2, [-4.0, -4.0], [4.0, 4.0], 1.0)
        d.from_points(i, v, 2)
        p = po.make_params(spatial=4, temporal=0, cfl=0.4, time=0.0, vel=dict(b=1.0, alpha=alpha))
        st, r = po.advect(d, p)
        assert st == 0
        return d.get(), r
    (i0, v0), r0 = run((0.0, 0.0, 0.0))
    (i1, v1), r1 = run((0.5, 0.0, 0.0))
    (i2, v2), r2 = run((0.0, 0.5, 0.0))
    assert r1["advected_time"] == r0["advected_time"] == r2["advected_time"] == 0.4
    assert not np.array_equal(v0, v1)
    # mirror symmetry of the square: the x-alpha result, transposed, is the y-alpha result
    a = {tuple(k): x for k, x in zip(i1.tolist(), v1) if abs(x) < 1e300}
    b = {(k[1], k[0]): x for k, x in zip(i2.tolist(), v2) if abs(x) < 1e300}
    assert a and a == b
