"""Pins of the oracle's learned-WENO5 network (oracle/mlp_oracle.cpp).

The gradient is checked against central finite differences of the loss
computed in double precision; the forward pass against closed forms
(zero network, bias-only network) and the learned weno5 against the
textbook fifth-order upwind-biased derivative (exact for quartics) and
against the reference weno5 where the kink fallback applies.
"""
import numpy as np
import pytest

import pyoracle as po

OFF = po.MLP_OFFSETS


def rand_params(seed, scale=0.5):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(po.MLP_NPARAM) * scale).astype(np.float32)


def rand_samples(seed, n):
    rng = np.random.default_rng(seed)
    X = np.zeros((n, 8), np.float32)
    d = rng.standard_normal((n, 5))
    d /= np.abs(d).max(axis=1, keepdims=True)
    X[:, :5] = d
    X[:, 5:] = rng.uniform(-1, 1, (n, 3))
    T = rng.uniform(-1, 1, n).astype(np.float32)
    return X, T


def test_gradient_matches_finite_differences():
    P = rand_params(1)
    X, T = rand_samples(2, 64)
    loss, g = po.mlp_loss_grad(P, X, T, use_double=True)
    rng = np.random.default_rng(3)
    idx = np.concatenate([rng.choice(np.arange(a, a + int(np.prod(s))), 8, replace=False)
                          for a, s in OFF.values() if np.prod(s) >= 8] + [np.arange(1344, 1347)])
    for i in idx:
        h = np.float32(2e-3)
        Pp = P.copy(); Pp[i] += h
        Pm = P.copy(); Pm[i] -= h
        hp = float(Pp[i]) - float(Pm[i])
        lp, _ = po.mlp_loss_grad(Pp, X, T, use_double=True)
        lm, _ = po.mlp_loss_grad(Pm, X, T, use_double=True)
        fd = (lp - lm) / hp
        assert abs(fd - g[i]) <= 2e-5 + 2e-4 * abs(g[i]), (i, fd, g[i])
    # the float path agrees with the double path
    lf, gf = po.mlp_loss_grad(P, X, T)
    assert abs(lf - loss) <= 1e-5 * abs(loss)
    assert np.max(np.abs(gf - g)) <= 1e-5 * (1 + np.max(np.abs(g)))


def test_emulated_rounding_is_close():
    P = rand_params(4)
    X, T = rand_samples(5, 128)
    l0, g0 = po.mlp_loss_grad(P, X, T)
    l1, g1 = po.mlp_loss_grad(P, X, T, emulate=True)
    assert l1 != l0  # bf16 rounding does change the numbers ...
    assert abs(l1 - l0) <= 0.05 * abs(l0)  # ... by bf16-sized amounts
    assert np.max(np.abs(g1 - g0)) <= 0.05 * np.max(np.abs(g0))


def test_zero_network_gives_uniform_weights():
    P = np.zeros(po.MLP_NPARAM, np.float32)
    w = po.mlp_forward(P, np.array([0.3, -1, 0.2, 0.5, 1], np.float32))
    assert np.allclose(w, 1 / 3, atol=0, rtol=1e-7)


def ideal_params():
    P = np.zeros(po.MLP_NPARAM, np.float32)
    P[1344:1347] = np.log(np.array([0.1, 0.6, 0.3], np.float64)).astype(np.float32)
    return P


def test_bias_only_network_gives_ideal_weights():
    w = po.mlp_forward(ideal_params(), np.array([1, 0.5, -0.5, 0.2, 0.1], np.float32))
    assert np.allclose(w, [0.1, 0.6, 0.3], rtol=1e-6, atol=0)


@pytest.mark.parametrize("plus", [False, True])
def test_ideal_weights_are_fifth_order_exact(plus):
    # f(x) = x + 0.01 x^2 + 0.001 x^3 + 1e-4 x^4 sampled at x = -3..3: the
    # linear-weight WENO5 combination is the upwind-biased 5th-order stencil,
    # exact for polynomials of degree <= 5 (Jiang & Shu 1996)
    c = np.array([0.0, 1.0, 0.01, 0.001, 1e-4])
    xs = np.arange(-3, 4, dtype=np.float64)
    f = sum(c[k] * xs ** k for k in range(5))
    exact = c[1]
    got = po.weno5_learned(f, 1.0, plus, ideal_params())
    # weights are float32 (0.1, 0.6, 0.3 rounded), so exactness holds to ~1e-7
    assert abs(got - exact) < 2e-7
    # the plain (nonlinear) weno5 differs measurably on the same data
    assert abs(po.weno5(f, 1.0, plus) - exact) > abs(got - exact)


@pytest.mark.parametrize("plus", [False, True])
def test_kink_falls_back_to_plain_weno5(plus):
    # a kink seen by one sub-stencil only (beta ratio far above 10)
    xs = np.arange(-3, 4, dtype=np.float64)
    f = np.abs(xs - (2.4 if plus else -2.4)) * 1.3
    P = rand_params(7)
    assert po.weno5_learned(f, 1.0, plus, P) == po.weno5(f, 1.0, plus)
    z = np.zeros(7)
    assert po.weno5_learned(z, 1.0, plus, P) == po.weno5(z, 1.0, plus)


def test_learned_scheme_advects_plane_exactly():
    """A plane is linear: every candidate polynomial equals the slope, so
    any weights give the exact Hamiltonian and the plane moves by v*T."""
    import cases
    dx, T, z0 = 0.5, 1.3, 0.2
    gmin, gmax, bc = [-6, -6, -10], [6, 6, 10], [0, 0, 1]
    d = po.Domain(3, gmin, gmax, bc, dx)
    ip, vp = cases.plane_points(3, gmin, gmax, bc, dx, z0)
    d.from_points(ip, vp, 2)
    p = po.make_params(spatial=12, temporal=2, cfl=0.4, time=T, vel=dict(kind=0, b=1.0), mlp=rand_params(8))
    st, r = po.advect(d, p)
    assert st == 0 and r["advected_time"] == pytest.approx(T, abs=1e-12)
    idx, val = d.defined()
    zi = idx[:, 2] * dx - val * dx
    m = np.abs(val) <= 0.5
    np.testing.assert_allclose(zi[m], z0 + T, atol=1e-9)
