"""Host-side pieces of the learned-WENO5 pipeline (viennals_amd/learned.py)
checked on the CPU against the oracle's weno5."""
import numpy as np

import pyoracle as po
from viennals_amd import learned as L


def test_weno5_parts_reproduce_weno5():
    rng = np.random.default_rng(3)
    st = np.cumsum(rng.uniform(-1.5, 1.5, (200, 7)), 1)
    for plus in (False, True):
        s = st[:, ::-1] if plus else st
        d, beta, p = L._weno5_parts(s)
        assert d.shape == (200, 5) and beta.shape == (200, 3) and p.shape == (200, 3)
        if plus:
            d, p = -d, -p
        al = L.IDEAL / (beta + 1e-6)
        mine = (al * p).sum(1) / al.sum(1)
        ref = np.array([po.weno5(x, 1.0, plus) for x in st])
        np.testing.assert_allclose(mine, ref, rtol=1e-12, atol=1e-12)


def test_star_shape_band_is_signed_distance_like():
    shape = L.StarShape(12.0, seed=4)
    idx, val = shape.band(1.0)
    assert np.all(np.abs(val) <= 1.0)
    # sorted in HRLE order (x fastest) and unique
    key = (idx[:, 2].astype(np.int64) << 40) + (idx[:, 1].astype(np.int64) << 20) + idx[:, 0]
    assert np.all(np.diff(key) > 0)
    # the zero level lies within the perturbation bounds of the radius
    r = np.linalg.norm(idx, axis=1)
    near = np.abs(val) < 0.2
    assert r[near].min() > 12.0 * 0.65 and r[near].max() < 12.0 * 1.35


def test_init_params_start_at_linear_weights():
    P = L.init_params(0)
    w = po.mlp_forward(P, np.array([0.2, 1.0, -0.3, 0.5, 0.1], np.float32))
    np.testing.assert_allclose(w, L.IDEAL, rtol=1e-6)


def test_star_perturbation_matches_scipy_harmonics():
    """The star radius perturbation (recurrence in l per order m) against
    the same sum over scipy's complex spherical harmonics."""
    from scipy.special import sph_harm_y
    shape = L.StarShape(10.0, seed=11)
    rng = np.random.default_rng(0)
    th = np.arccos(rng.uniform(-1, 1, 2000))
    ph = rng.uniform(0, 2 * np.pi, 2000)
    th[:4] = [0.0, np.pi, np.pi / 2, 1e-9]  # poles and the equator
    ref = np.zeros_like(th)
    for l, m, a in shape.terms:
        y = sph_harm_y(l, abs(m), th, ph)
        ref += a * (np.real(y) if m == 0 else np.sqrt(2.0) * (np.real(y) if m > 0 else np.imag(y)))
    np.testing.assert_allclose(shape._pert(th, ph), ref, rtol=0, atol=1e-12)
