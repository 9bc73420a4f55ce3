"""Host-evaluated velocities (VelocityField subclasses on the host,
lsg_advect_set_velocity_host): a host callback that computes the same
function as a built-in functor gives bit-identical advection, and the
callback receives the reference's arguments (coordinate = index * gridDelta,
the scheme's normal, the defined-point id) for exactly the active points."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class NormalPower:
    """b + (n[dir] > 0 ? aPos : aNeg) * |n[dir]|^k, same operation order as the functor."""

    def __init__(self, dir, k, b, a_pos, a_neg, vec=(0.0, 0.0, 0.0)):
        self.dir, self.k, self.b, self.a_pos, self.a_neg, self.vec = dir, k, b, a_pos, a_neg, vec

    def scalar(self, n):
        c = n[self.dir]
        a = self.a_pos if c > 0 else self.a_neg
        ac = abs(c)
        m = 1.0
        if self.k >= 1:
            m = ac
            for _ in range(1, self.k):
                m = m * ac
        return self.b + a * m


def sphere(vls, r=6.0):
    d = vls.Domain(gridDelta=0.5, dim=3)
    vls.MakeGeometry(d, vls.Sphere([0.3, -0.2, 0.1], r)).apply()
    return d


@pytest.mark.parametrize("spatial,temporal", [(0, 0), (1, 2), (11, 2), (6, 1), (10, 0), (2, 0), (3, 2), (7, 0),
                                              (8, 1), (9, 0), (9, 2)])
def test_host_callback_matches_builtin(spatial, temporal):
    import viennals_amd as vls
    f = NormalPower(dir=0, k=2, b=1.0, a_pos=2.3, a_neg=0.5, vec=(0.2, 0.0, -0.1))

    class Host(vls.HostVelocityField):
        def getScalarVelocity(self, coordinate, material, normal, pointId):
            return f.scalar(normal)

        def getVectorVelocity(self, coordinate, material, normal, pointId):
            return f.vec

    results = []
    for vel in (vls.VelocityField.normal_power(dir=0, k=2, b=1.0, a_pos=2.3, a_neg=0.5, vector=f.vec), Host()):
        d = sphere(vls)
        a = vls.Advect(d, vel)
        a.setSpatialScheme(spatial)
        a.setTemporalScheme(temporal)
        a.setTimeStepRatio(0.4)
        for _ in range(2):
            a.apply()
        results.append((d.get_runs(), a.getAdvectedTime()))
    (i0, v0), t0 = results[0]
    (i1, v1), t1 = results[1]
    assert t0 == t1
    assert np.array_equal(i0, i1) and np.array_equal(v0, v1)


def test_host_callback_arguments():
    import viennals_amd as vls
    d = sphere(vls, 4.0)
    calls = []

    class Rec(vls.HostVelocityField):
        def getScalarVelocity(self, coordinate, material, normal, pointId):
            calls.append((coordinate, material, normal, pointId))
            return 1.0

    a = vls.Advect(d, Rec())
    a.setSpatialScheme(vls.SpatialSchemeEnum.ENGQUIST_OSHER_1ST_ORDER)
    a.setTemporalScheme(vls.TemporalSchemeEnum.FORWARD_EULER)
    a.prepareLS()  # expanded band as the first stage sees it
    idx, val = d.get_runs()
    defined = np.abs(val) < 1e300
    active = defined & (np.abs(val) <= 0.5)
    rank = np.cumsum(defined) - 1
    a.setSingleStep(True)
    a.apply()
    assert len(calls) == int(active.sum())
    coords = np.array([c[0] for c in calls])
    np.testing.assert_array_equal(coords, idx[active] * 0.5)
    assert all(c[1] == 0 for c in calls)
    np.testing.assert_array_equal(np.array([c[3] for c in calls]), rank[active])
    norms = np.linalg.norm(np.array([c[2] for c in calls]), axis=1)
    np.testing.assert_allclose(norms, 1.0, rtol=1e-12)


def test_host_callback_llf_box_calls():
    """LLF calls the velocity at the 3^D box around every active point with
    coordinate + offset * gridDelta, the box point's normal and the centre's
    point id (lsLocalLaxFriedrichs.hpp:193-222), after the scheme's call."""
    import viennals_amd as vls
    d = sphere(vls, 4.0)
    calls = []

    class Rec(vls.HostVelocityField):
        def getScalarVelocity(self, coordinate, material, normal, pointId):
            calls.append((tuple(coordinate), pointId))
            return 1.0

    a = vls.Advect(d, Rec())
    a.setSpatialScheme(vls.SpatialSchemeEnum.LOCAL_LAX_FRIEDRICHS_1ST_ORDER)
    a.setTemporalScheme(vls.TemporalSchemeEnum.FORWARD_EULER)
    a.prepareLS()
    idx, val = d.get_runs()
    active = (np.abs(val) < 1e300) & (np.abs(val) <= 0.5)
    a.setSingleStep(True)
    a.apply()
    n = int(active.sum())
    assert len(calls) == 28 * n
    centres = np.array([calls[28 * j][0] for j in range(n)])
    np.testing.assert_array_equal(centres, idx[active] * 0.5)
    for j in (0, n // 2, n - 1):
        c = np.array(calls[28 * j][0])
        for q in range(27):
            off = np.array([q % 3 - 1, (q // 3) % 3 - 1, q // 9 - 1]) * 0.5
            np.testing.assert_array_equal(np.array(calls[28 * j + 1 + q][0]), c + off)
            assert calls[28 * j + 1 + q][1] == calls[28 * j][1]


def test_host_callback_sllf_calls():
    """The stencil local Lax-Friedrichs scheme calls the velocity at the
    point (its centre normal), at the 3^D box points (coordinate + offset *
    gridDelta, the box point's normal) and then, per box point and axis k, at
    the box normal moved by -DN and +DN along k, DN = cbrt(eps) * gridDelta
    for a unit velocity (lsStencilLocalLaxFriedrichsScalar.hpp:62-74,
    252-338): 28 + 27 * 6 calls per active point, the second round after
    the first."""
    import viennals_amd as vls
    d = sphere(vls, 4.0)
    calls = []

    class Rec(vls.HostVelocityField):
        def getScalarVelocity(self, coordinate, material, normal, pointId):
            calls.append((tuple(coordinate), tuple(normal), pointId))
            return 1.0

    a = vls.Advect(d, Rec())
    a.setSpatialScheme(9)
    a.setTemporalScheme(vls.TemporalSchemeEnum.FORWARD_EULER)
    a.prepareLS()
    idx, val = d.get_runs()
    active = (np.abs(val) < 1e300) & (np.abs(val) <= 0.5)
    a.setSingleStep(True)
    a.apply()
    n = int(active.sum())
    assert len(calls) == (28 + 162) * n
    first, second = calls[:28 * n], calls[28 * n:]
    np.testing.assert_array_equal(np.array([first[28 * j][0] for j in range(n)]), idx[active] * 0.5)
    DN = np.cbrt(np.finfo(np.float64).eps) * 0.5
    for j in (0, n // 2, n - 1):
        for q in range(27):
            coord, ln, pid = first[28 * j + 1 + q]
            assert pid == first[28 * j][2]
            for k in range(3):
                cp, npn, pp = second[162 * j + 6 * q + 2 * k]
                cn, nnn, pn = second[162 * j + 6 * q + 2 * k + 1]
                assert cp == coord and cn == coord and pp == pid and pn == pid
                assert abs(npn[k] - (ln[k] - DN)) <= 1e-15 and abs(nnn[k] - (ln[k] + DN)) <= 1e-15
                for i in range(3):
                    if i != k:  # earlier axes moved and moved back, later ones untouched
                        assert abs(npn[i] - ln[i]) <= 1e-15 and abs(nnn[i] - ln[i]) <= 1e-15
                        if i > k:
                            assert npn[i] == ln[i] and nnn[i] == ln[i]


def test_host_callback_rejected_with_learned():
    import viennals_amd as vls
    d = sphere(vls, 3.0)
    a = vls.Advect(d, vls.HostVelocityField())
    a.setSpatialScheme(12)
    with pytest.raises(vls.LsgError, match="host-evaluated"):
        a.apply()


@pytest.mark.parametrize("temporal", [0, 2])
def test_host_llfa_alphas_match_builtin(temporal):
    """LOCAL_LAX_FRIEDRICHS_ANALYTICAL with getDissipationAlpha on the host
    (constant per direction, like the built-in functors' alphas): bit-identical
    to the built-in functor with the same alphas."""
    import viennals_amd as vls
    f = NormalPower(dir=0, k=2, b=1.0, a_pos=2.3, a_neg=0.5, vec=(0.2, 0.0, -0.1))
    al = (0.7, 0.4, 0.9)

    class Host(vls.HostVelocityField):
        def getScalarVelocity(self, coordinate, material, normal, pointId):
            return f.scalar(normal)

        def getVectorVelocity(self, coordinate, material, normal, pointId):
            return f.vec

        def getDissipationAlpha(self, direction, material, centralDifferences):
            return al[direction]

    builtin = vls.VelocityField(kind=1, dir=0, k=2, b=1.0, a_pos=2.3, a_neg=0.5, vec=f.vec, alpha=al)
    results = []
    for vel in (builtin, Host()):
        d = sphere(vls)
        a = vls.Advect(d, vel)
        a.setSpatialScheme(4)
        a.setTemporalScheme(temporal)
        a.setTimeStepRatio(0.4)
        for _ in range(2):
            a.apply()
        results.append((d.get_runs(), a.getAdvectedTime()))
    (i0, v0), t0 = results[0]
    (i1, v1), t1 = results[1]
    assert t0 == t1
    assert np.array_equal(i0, i1) and np.array_equal(v0, v1)


def test_host_llfa_alpha_box_calls():
    """getDissipationAlpha is called for every direction at the 27 box
    points of every active point, right after the point's velocity call; at
    the box centre its central differences, normalised, are the normal the
    velocity call received (two separate device computations of the same
    differences); alphas that grow with the differences advect the surface
    differently from constant ones."""
    import viennals_amd as vls
    events = []

    class Host(vls.HostVelocityField):
        def getScalarVelocity(self, coordinate, material, normal, pointId):
            events.append(("v", normal))
            return 1.0

        def getDissipationAlpha(self, direction, material, centralDifferences):
            events.append(("a", direction, centralDifferences))
            return 0.5 + min(abs(centralDifferences[direction]), 4.0)

    d = sphere(vls, 4.0)
    a = vls.Advect(d, Host())
    a.setSpatialScheme(4)
    a.setTemporalScheme(0)
    a.setTimeStepRatio(0.4)
    a.apply()
    per = 1 + 27 * 3
    assert len(events) > 0 and len(events) % per == 0
    for k in range(0, len(events), per):
        assert events[k][0] == "v"
        box = events[k + 1:k + per]
        assert [e[0] for e in box] == ["a"] * 81 and [e[1] for e in box] == [0, 1, 2] * 27
        cd = np.array(box[13 * 3][2])  # box point q = 13: offset (0, 0, 0)
        nm = 1.0 / np.sqrt(cd[0] * cd[0] + cd[1] * cd[1] + cd[2] * cd[2])
        assert tuple(cd * nm) == tuple(events[k][1])
    d2 = sphere(vls, 4.0)
    b = vls.Advect(d2, vls.VelocityField(kind=0, b=1.0, alpha=(0.5, 0.5, 0.5)))
    b.setSpatialScheme(4)
    b.setTemporalScheme(0)
    b.setTimeStepRatio(0.4)
    b.apply()
    _, v1 = d.get_runs()
    _, v2 = d2.get_runs()
    assert not (len(v1) == len(v2) and np.array_equal(v1, v2))


@pytest.mark.parametrize("spatial,temporal", [(3, 2), (8, 1), (2, 0), (4, 0), (4, 2), (9, 0), (9, 2)])
def test_host_callback_several_level_sets(spatial, temporal):
    """A host callback with a material-dependent velocity equals the
    built-in per-material functors with LF (global alphas over the points'
    materials), LLF (box velocities per material), the analytical LLF
    (box alphas per material) and the stencil local LF (box and perturbed
    normal velocities per material) on an etch stop."""
    import viennals_amd as vls
    from test_gpu_multimaterial import plane2
    from test_gpu_parity import to_gpu
    rates = [(-0.05, -0.2), (-0.1, -1.0)]
    fs = [NormalPower(dir=1, k=1, b=b, a_pos=ap, a_neg=0.0) for b, ap in rates]

    alphas = [(0.0, 0.3, 0.1), (0.2, 1.2, 0.0)]

    class Host(vls.HostVelocityField):
        def getScalarVelocity(self, coordinate, material, normal, pointId):
            return fs[material].scalar(normal)

        def getDissipationAlpha(self, direction, material, centralDifferences):
            return alphas[material][direction]

    builtin = vls.MaterialVelocityField([vls.VelocityField(kind=1, dir=1, k=1, b=b, a_pos=ap, a_neg=0.0, alpha=al)
                                         for (b, ap), al in zip(rates, alphas)])
    out = []
    for vel in (builtin, Host()):
        lv = [to_gpu(plane2(-2.3)), to_gpu(plane2(0.0))]
        a = vls.Advect()
        for g in lv:
            a.insertNextLevelSet(g)
        a.setVelocityField(vel)
        a.setSpatialScheme(spatial)
        a.setTemporalScheme(temporal)
        a.setTimeStepRatio(0.4)
        a.setAdvectionTime(3.0)
        a.apply()
        out.append(([g.get_runs() for g in lv], a.getAdvectedTime()))
    assert out[0][1] == out[1][1]
    for (i0, v0), (i1, v1) in zip(out[0][0], out[1][0]):
        assert np.array_equal(i0, i1) and np.array_equal(v0.view(np.int64), v1.view(np.int64))
