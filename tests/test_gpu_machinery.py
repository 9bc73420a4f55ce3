"""GPU tests of the host machinery around the kernels: graph replay against
direct launches, the capacity-overflow retry, and frame moves of a
travelling band -- all bit-exact against the oracle."""
import os

import pytest

import cases
import pyoracle as po
from test_gpu_parity import _advect_gpu, assert_same, say, to_gpu

pytestmark = pytest.mark.gpu


def _run(o, p):
    g = to_gpu(o)
    _advect_gpu(g, p)
    return g


def test_graph_replay_equals_direct_launches(monkeypatch):
    say("graph replay vs direct launches: trench WENO5/RK3 10 steps")
    p = po.make_params(spatial=11, temporal=2, cfl=0.4, time=0.0, vel=cases.VEL_ETCH_DIRECTIONAL)
    import viennals_amd as vls
    base = cases.trench()
    o = base.copy()
    g_graph = to_gpu(base)
    g_direct = to_gpu(base)
    monkeypatch.delenv("LSG_NO_GRAPH", raising=False)
    vel = vls.VelocityField(**cases.VEL_ETCH_DIRECTIONAL)
    ag = vls.Advect(g_graph, vel)
    ag.setSpatialScheme(11)
    ag.setTemporalScheme(2)
    ag.setTimeStepRatio(0.4)
    for _ in range(10):
        ag.apply()
        po.advect(o, p)
    assert_same(o, g_graph, "graph replay")
    monkeypatch.setenv("LSG_NO_GRAPH", "1")
    ad = vls.Advect(g_direct, vel)
    ad.setSpatialScheme(11)
    ad.setTemporalScheme(2)
    ad.setTimeStepRatio(0.4)
    for _ in range(10):
        ad.apply()
    assert_same(o, g_direct, "direct launches")


def test_capacity_overflow_retry(monkeypatch):
    say("capacity overflow retry: tiny initial capacity, sphere WENO5/RK3")
    monkeypatch.setenv("LSG_TEST_CAPACITY", "2000")
    o = cases.sphere(r=4.0, dx=0.25)
    g = to_gpu(o)
    p = po.make_params(spatial=11, temporal=2, cfl=0.4, time=0.5, vel=cases.VEL_GROW)
    st, r = po.advect(o, p)
    a = _advect_gpu(g, p)
    assert a.getAdvectedTime() == r["advected_time"]
    assert_same(o, g, "after retries")


def test_frame_moves_with_band():
    say("travelling sphere (frame moves): EO1/FE, vector velocity, 32 cells")
    o = cases.sphere(r=4.0, dx=0.25)
    g = to_gpu(o)
    p = po.make_params(spatial=0, temporal=0, cfl=0.4, time=8.0, vel=dict(kind=0, b=0.0, vec=(1.0, 0.5, 0.0)))
    po.advect(o, p)
    _advect_gpu(g, p)
    assert_same(o, g, "travelling sphere")
