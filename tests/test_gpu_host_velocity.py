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


@pytest.mark.parametrize("spatial,temporal", [(0, 0), (1, 2), (11, 2), (6, 1), (10, 0)])
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


def test_host_callback_rejected_with_lax_friedrichs():
    import viennals_amd as vls
    d = sphere(vls, 3.0)
    a = vls.Advect(d, vls.HostVelocityField())
    a.setSpatialScheme(vls.SpatialSchemeEnum.LAX_FRIEDRICHS_1ST_ORDER)
    with pytest.raises(vls.LsgError, match="host-evaluated"):
        a.apply()
