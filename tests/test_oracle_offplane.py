"""Pins of the oracle's Hamiltonians off the plane (no GPU).

On a plane D+ = D-, so every dissipation term, every box or global alpha and
every second-order correction vanishes there.  These fields make them
non-zero with values known by hand:

* a V profile phi = (k - k0) + a |i| (grid units, periodic in x and y,
  linear in z): at the valley one-sided x differences are -+a/dx, at the
  peak +-a/dx, and the central-difference normals of the slopes have
  n_x = a / sqrt(1 + a^2);
* a V whose slopes steepen from a to b beyond |i| = 3, so the global alpha
  (LaxFriedrichs, lsAdvect.hpp:90-195) differs from the local 3^D box alpha
  (LocalLaxFriedrichs, lsLocalLaxFriedrichs.hpp:184-232) and from the
  centre alpha (LocalLocalLaxFriedrichs, lsLocalLocalLaxFriedrichs.hpp:157-162);
* a parabola phi = (k - k0) + c i^2, on which the second-order ENO
  correction (lsEngquistOsher.hpp:70-108) turns the one-sided differences
  into the exact derivative 2 c i / dx, so the dissipation (D+ - D-) / 2
  vanishes for order 2 but not for order 1.

Every closed form below is derived from the reference's formulas for these
inputs, not by evaluating them in code.  test_pins_catch_planted_mistakes
builds the oracle with each planted mistake of ls_oracle.cpp (LSO_MUTATION)
and shows that these pins fail on it.
"""
import os
import subprocess
import sys
import tempfile

import numpy as np
import pytest

import pyoracle as po

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
EO1, EO2, LF1, LF2, LLLF1, LLLF2, LLF1, LLF2 = 0, 1, 2, 3, 5, 6, 7, 8  # SpatialSchemeEnum
DX = 0.5
K0 = -0.2  # phi(0, 0, 0) = 0.2 + g(0)
L = 16     # periodic extent in x and y (indices -8 .. 7)


def v_profile(a):
    return lambda i: a * abs(i)


def two_slopes(a, b):
    return lambda i: a * abs(i) if abs(i) <= 3 else 3 * a + b * (abs(i) - 3)


def parabola(c):
    return lambda i: c * i * i if abs(i) <= 3 else 9 * c + 6 * c * (abs(i) - 3)


def field(g, axis=0):
    """All points of the periodic box defined (width 64, so prepareLS adds
    nothing): phi = (k - K0) + g(i) with i the index along `axis`."""
    n = L // 2
    gmin, gmax, bc = [-n, -n, -30], [n, n, 30], [2, 2, 1]
    i, j, k = np.meshgrid(np.arange(-n, n), np.arange(-n, n), np.arange(-14, 15), indexing="ij")
    idx = np.stack([i.ravel(), j.ravel(), k.ravel()], 1).astype(np.int32)
    lateral = idx[:, axis]
    val = (idx[:, 2] - K0) + np.array([g(int(q)) for q in lateral])
    order = np.lexsort((idx[:, 0], idx[:, 1], idx[:, 2]))
    d = po.Domain(3, gmin, gmax, bc, DX)
    d.from_points(idx[order], val[order], 64)
    return d


def rates_at(d, spatial, V, points):
    """(dt, [(rate, dissipation) at each point]) of one rate pass with a
    constant scalar velocity V."""
    p = po.make_params(spatial=spatial, temporal=0, cfl=0.4, vel=dict(kind=0, b=V))
    dt, r, q = po.compute_rates_diss(d, p)
    idx, val = d.get()
    out = []
    for pt in points:
        e = np.where((idx == np.array(pt)).all(1))[0]
        assert len(e) == 1, pt
        assert abs(val[e[0]]) <= 0.5, f"{pt} is not active"
        out.append((r[e[0]], q[e[0]]))
    return dt, out


def active_k(g, i):
    """z index of the active point of column i (|phi| <= 0.5)."""
    return int(np.round(K0 - g(i)))


# ------------------------------------------------------ Engquist-Osher ----

@pytest.mark.parametrize("spatial", [EO1, EO2])
@pytest.mark.parametrize("V", [1.0, -1.0])
def test_eo_v_profile(spatial, V):
    """Valley and peak of a V (a = 0.3): the growing front (V > 0) sees only
    the z slope at the valley but both x slopes at the peak, the etching
    front (V < 0) the reverse: H = V/dx or V sqrt(1 + 2a^2)/dx
    (lsEngquistOsher.hpp:113-116,147-151; the order-2 correction is zero on
    straight slopes).  Catches swapped upwinding (LSO_MUTATION 4)."""
    a = 0.3
    g = v_profile(a)
    valley, peak = (0, 0, active_k(g, 0)), (-8, 0, active_k(g, -8))
    _, ((rv, dv), (rp, dp)) = rates_at(field(g), spatial, V, [valley, peak])
    flat, kink = abs(V) / DX, abs(V) * np.sqrt(1 + 2 * a * a) / DX
    exp_valley, exp_peak = (flat, kink) if V > 0 else (-kink, -flat)
    assert rv == pytest.approx(exp_valley, rel=1e-13)
    assert rp == pytest.approx(exp_peak, rel=1e-13)
    assert dv == 0.0 and dp == 0.0


@pytest.mark.parametrize("spatial", [EO1, EO2])
def test_eo_parabola_second_order(spatial):
    """Vertex of phi = (k - k0) + c i^2 (c = 0.25) etched with V = -1: first
    order sees x differences -+c/dx (H = -sqrt(1 + 2c^2)/dx), second order
    corrects both to 0 (H = -1/dx).  Catches a dropped order-2 term
    (LSO_MUTATION 2)."""
    c = 0.25
    g = parabola(c)
    _, ((r, _),) = rates_at(field(g), spatial, -1.0, [(0, 0, active_k(g, 0))])
    expected = -1.0 / DX if spatial == EO2 else -np.sqrt(1 + 2 * c * c) / DX
    assert r == pytest.approx(expected, rel=1e-13)


# ------------------------------------------------- Lax-Friedrichs family ----

A, B = 0.2, 0.5  # two-slope V


def _alpha(s):  # |V n_x| of a straight slope of steepness s, V = 1
    return s / np.sqrt(1 + s * s)


@pytest.mark.parametrize("spatial,alpha_x", [
    (LF1, _alpha(B)), (LF2, _alpha(B)),        # global: the steepest slope anywhere
    (LLF1, _alpha(A)), (LLF2, _alpha(A)),      # 3^D box: the slopes next to the valley
    (LLLF1, 0.0), (LLLF2, 0.0),                # centre: n_x = 0 at the valley
])
def test_lf_family_dissipation_at_valley(spatial, alpha_x):
    """Valley of the two-slope V, V = 1: H = V/dx (the centred gradient has
    no x part) and the dissipation is alpha_x (D+ - D-)/2 = alpha_x a/dx, the
    z term vanishing (lsLaxFriedrichs.hpp:139-140; lsLocalLaxFriedrichs.hpp:
    236-238; lsLocalLocalLaxFriedrichs.hpp:157-162).  Catches a flipped
    dissipation sign (LSO_MUTATION 1), centre instead of box alphas (3) and
    would catch global alphas used locally or the reverse."""
    g = two_slopes(A, B)
    _, ((r, d),) = rates_at(field(g), spatial, 1.0, [(0, 0, active_k(g, 0))])
    diss = alpha_x * A / DX
    assert d == pytest.approx(diss, rel=1e-13, abs=1e-15)
    assert r == pytest.approx(1.0 / DX - diss, rel=1e-13)


@pytest.mark.parametrize("spatial", [LF1, LF2, LLF1, LLF2, LLLF1, LLLF2])
def test_lf_family_dissipation_at_peak(spatial):
    """Peak of a V (a = 0.3), V = 1: D+ - D- = -2a/dx, so the dissipation is
    -alpha_x a/dx and the rate exceeds V/dx; every alpha is that of the
    straight slopes here except the centre alpha (0)."""
    a = 0.3
    g = v_profile(a)
    _, ((r, d),) = rates_at(field(g), spatial, 1.0, [(-8, 0, active_k(g, -8))])
    alpha_x = 0.0 if spatial in (LLLF1, LLLF2) else _alpha(a)
    assert d == pytest.approx(-alpha_x * a / DX, rel=1e-13, abs=1e-15)
    assert r == pytest.approx(1.0 / DX + alpha_x * a / DX, rel=1e-13)


@pytest.mark.parametrize("spatial", [LF1, LF2, LLF1, LLF2, LLLF1, LLLF2])
def test_lf_family_parabola_order(spatial):
    """phi = (k - k0) + c i^2 (c = 0.25) at i = 1, V = 1.  Both orders give
    the centred x derivative 2c/dx, so H = sqrt(1 + 4c^2)/dx.  Order 2
    corrects the one-sided differences to 2c/dx each, so no dissipation;
    order 1 keeps 3c/dx and c/dx, so the dissipation is alpha_x c/dx with
    alpha_x = |n_x| at the centre (LLLF: n = (2c, 0, 1)/norm), over the box
    (LLF: the largest centred slope in i = 0..2 is (g(3)-g(1))/2 = 4c), or
    over all active points (LF: the steepest centred slope, 6c beyond
    |i| = 3).  Catches a dropped order-2 term (LSO_MUTATION 2)."""
    c = 0.25
    g = parabola(c)
    _, ((r, d),) = rates_at(field(g), spatial, 1.0, [(1, 0, active_k(g, 1))])
    H = np.sqrt(1 + 4 * c * c) / DX
    if spatial in (LF2, LLF2, LLLF2):
        diss = 0.0
    else:
        slope = {LF1: 6 * c, LLF1: 4 * c, LLLF1: 2 * c}[spatial]
        diss = _alpha(slope) * c / DX
    assert d == pytest.approx(diss, rel=1e-12, abs=1e-14)
    assert r == pytest.approx(H - diss, rel=1e-12)


def test_lf_global_time_step():
    """LaxFriedrichs time step on the two-slope V: dt = min(cfl / max|v|,
    1 / sum(alpha_d / dx)) with alpha_x of the b slopes and alpha_z = 1 (the
    kinks' normals are +z) (lsLaxFriedrichs.hpp:160-172, lsAdvect.hpp:591-606).
    The largest rate is at the peak, (1 + alpha_x b)/dx (its dissipation is
    -alpha_x b/dx), above the b slopes' sqrt(1 + b^2)/dx."""
    g = two_slopes(A, B)
    dt, _ = rates_at(field(g), LF1, 1.0, [(0, 0, active_k(g, 0))])
    vmax = max((1 + _alpha(B) * B) / DX, np.sqrt(1 + B * B) / DX)
    cfl_dt = 0.4 / vmax
    alpha_dt = 1.0 / (_alpha(B) / DX + 1.0 / DX)
    assert dt == pytest.approx(min(cfl_dt, alpha_dt), rel=1e-13)


# ------------------------------------------------------------ symmetry ----

def _rate_map(d, spatial, V):
    p = po.make_params(spatial=spatial, temporal=0, cfl=0.4, vel=dict(kind=0, b=V))
    _, r, q = po.compute_rates_diss(d, p)
    idx, val = d.get()
    act = np.abs(val) <= 0.5
    return {tuple(k): (x, y) for k, x, y in zip(idx[act], r[act], q[act])}


@pytest.mark.parametrize("spatial", [EO1, EO2, LF1, LF2, LLF1, LLF2, LLLF1, LLLF2])
def test_axis_permutation_and_mirror(spatial):
    """An asymmetric profile (slope 0.3 for i >= 0, 0.1 below, periodic) along
    x and the same along y give the same rate at swapped points; the profile
    mirrored in x gives the same rate at mirrored points.  Catches mixed-up
    neighbour indices of one axis or of one side."""
    g = lambda i: 0.3 * i if i >= 0 else -0.1 * i  # noqa: E731
    gm = lambda i: g(-i if i != -8 else -8)        # noqa: E731  (mirror i -> -i mod 16)
    V = 1.0 if spatial != EO1 else -1.0
    rx = _rate_map(field(g, 0), spatial, V)
    ry = _rate_map(field(g, 1), spatial, V)
    rm = _rate_map(field(gm, 0), spatial, V)
    assert len(rx) == len(ry) == len(rm) > 0
    for (i, j, k), (x, dx_) in rx.items():
        y, dy = ry[(j, i, k)]
        assert y == pytest.approx(x, rel=1e-13, abs=1e-13) and dy == pytest.approx(dx_, rel=1e-13, abs=1e-13)
        mi = -i if i != -8 else -8
        m, dm = rm[(mi, j, k)]
        assert m == pytest.approx(x, rel=1e-13, abs=1e-13) and dm == pytest.approx(dx_, rel=1e-13, abs=1e-13)


# --------------------------------------------------------- convergence ----

@pytest.mark.parametrize("spatial", [LF2, LLF2, LLLF2])
def test_sphere_translation_converges(spatial):
    """The reference's TimeIntegration setup (sphere r = 1.5 moved by (1,0,0)
    for t = 0.5, RK3) with the Lax-Friedrichs schemes: the Chamfer distance to
    the sphere made at x = 0.5 shrinks when the grid is refined from
    dx 0.15 to 0.1 and stays below one cell."""
    import judges
    res = []
    for dx in (0.15, 0.1):
        ext = int(np.ceil(5 / dx))
        s = po.Domain(3, [-ext] * 3, [ext] * 3, [1] * 3, dx)
        s.make_sphere([0.0, 0.0, 0.0], 1.5)
        ref = po.Domain(3, [-ext] * 3, [ext] * 3, [1] * 3, dx)
        ref.make_sphere([0.5, 0.0, 0.0], 1.5)
        p = po.make_params(spatial=spatial, temporal=2, cfl=0.4999, time=0.5,
                           vel=dict(kind=0, b=0.0, vec=(1.0, 0.0, 0.0)))
        st, _ = po.advect(s, p)
        assert st == 0 and s.check()[0] == 0
        res.append(judges.chamfer(judges.surface_points(*ref.get(), dx), judges.surface_points(*s.get(), dx)))
    assert res[1] < res[0] < 0.15, res


# ------------------------------------------------- the pins catch mistakes ----

MUTATIONS = {
    1: "test_lf_family_dissipation_at_valley",
    2: "test_lf_family_parabola_order or test_eo_parabola_second_order",
    3: "test_lf_family_dissipation_at_valley",
    4: "test_eo_v_profile",
}


def test_pins_catch_planted_mistakes():
    """Builds the oracle with each planted mistake (ls_oracle.cpp,
    LSO_MUTATION 1-4) and runs the pins that name it against that build: each
    run must fail."""
    src = os.path.join(ROOT, "oracle")
    with tempfile.TemporaryDirectory() as tmp:
        builds = {}
        for m in MUTATIONS:
            out = os.path.join(tmp, f"liblsoracle_m{m}.so")
            builds[m] = (out, subprocess.Popen(
                ["g++", "-O1", "-std=c++17", "-fPIC", "-fopenmp", "-ffp-contract=off", "-shared",
                 f"-DLSO_MUTATION={m}", "-o", out, os.path.join(src, "ls_oracle.cpp"),
                 os.path.join(src, "mlp_oracle.cpp")]))
        for m, (out, pr) in builds.items():
            assert pr.wait() == 0, f"build of mutation {m} failed"
        runs = {}
        for m, (out, _) in builds.items():
            env = dict(os.environ, LSO_ORACLE_LIB=out, OMP_NUM_THREADS="2")
            runs[m] = subprocess.Popen(
                [sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", __file__, "-k", MUTATIONS[m]],
                cwd=ROOT, env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
        for m, pr in runs.items():
            log = pr.communicate()[0].decode()
            assert pr.returncode == 1, f"mutation {m}: the pins did not fail\n{log[-2000:]}"
            assert "failed" in log, log[-2000:]
