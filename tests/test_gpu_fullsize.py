"""GPU parity at the benchmark sizes.

The benchmark workloads are built exactly as bench.py builds them (on the
GPU, through the package's MakeGeometry / BooleanOperation / insertPoints),
their run lists are handed to the CPU oracle, and both sides advect:

* trench256 WENO5-RK3 (256x256 lateral cells, reflective x/y) up to
  advection time 40 in single steps of Advect::apply with the remaining time
  as the step bound (lsAdvect.hpp:1031-1040), split over several tests of
  about 25 steps so that every test ends within a minute;
* rough1024 LLF2-RK3 (1024x1024 periodic Gaussian height field) for 2 steps;
* the 512-level-set batch (seed 42) for 3 batched RK3 steps, of which the 4
  largest and the 4 smallest level sets are each advected alone by the
  oracle.

Run lists and advected times are compared bitwise after every step (every
batched step for the batch: its per-level-set times are checked at the end).
"""
import sys

import numpy as np
import pytest

import bench
import cases
import pyoracle as po

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def _vls():
    import viennals_amd as vls
    return vls


def oracle_of(g):
    o = po.Domain(g.D, list(g.gmin), list(g.gmax), list(g.bc), g.gridDelta)
    idx, val = g.get_runs()
    o.set(idx, val, g.getLevelSetWidth())
    return o


def assert_same(o, g, tag):
    io, vo = o.get()
    ig, vg = g.get_runs()
    if not cases.bit_equal(io, vo, ig, vg):
        i, a, b, c, d = cases.first_diff(io, vo, ig, vg)
        pytest.fail(f"{tag}: run lists differ (oracle {len(vo)} vs gpu {len(vg)}), first diff at {i}: "
                    f"oracle {a} {b} gpu {c} {d}")


# ------------------------------------------------------------ trench256 ----

T_END = 40.0
PARTS = 6  # about 120 RK3 steps at dt ~0.33: 6 parts of up to 25 steps
STEPS_PER_PART = 25


@pytest.fixture(scope="module")
def trench_state():
    vls = _vls()
    say("trench256 WENO5-RK3: building the workload")
    g, vel, sp, tp, cfl, _ = bench.build_workload(vls, "trench256")
    o = oracle_of(g)
    assert_same(o, g, "trench256 input")
    a = vls.Advect(g, vel)
    a.setSpatialScheme(sp)
    a.setTemporalScheme(tp)
    a.setTimeStepRatio(cfl)
    a.setSingleStep(True)
    st = dict(g=g, o=o, a=a, vel=vel.as_dict(), sp=int(sp), tp=int(tp), cfl=cfl, t=0.0, steps=0)
    yield st


@pytest.mark.parametrize("part", range(PARTS))
def test_trench256_weno5_rk3_to_t40(trench_state, part):
    st = trench_state
    g, o, a = st["g"], st["o"], st["a"]
    if st["t"] >= T_END:
        pytest.skip("advection time 40 already reached")
    say(f"trench256 part {part}: from t={st['t']:.4f} (step {st['steps']})")
    for _ in range(STEPS_PER_PART):
        if st["t"] >= T_END:
            break
        remaining = T_END - st["t"]
        a.setAdvectionTime(remaining)
        a.apply()
        p = po.make_params(spatial=st["sp"], temporal=st["tp"], cfl=st["cfl"], time=remaining, single_step=True,
                           vel=st["vel"])
        rc, r = po.advect(o, p)
        assert rc == 0
        tg = a.getAdvectedTime()
        assert tg == r["advected_time"], f"step {st['steps']}: advected time gpu {tg!r} oracle {r['advected_time']!r}"
        st["t"] += tg
        st["steps"] += 1
        assert_same(o, g, f"trench256 step {st['steps']} (t={st['t']:.4f})")
    if part == PARTS - 1:
        assert st["t"] >= T_END, f"advection time 40 not reached after {st['steps']} steps"


# ------------------------------------------------------------ rough1024 ----

def test_rough1024_llf2_rk3_two_steps():
    vls = _vls()
    say("rough1024 LLF2-RK3: building the workload")
    g, vel, sp, tp, cfl, _ = bench.build_workload(vls, "rough1024")
    o = oracle_of(g)
    a = vls.Advect(g, vel)
    a.setSpatialScheme(sp)
    a.setTemporalScheme(tp)
    a.setTimeStepRatio(cfl)
    p = po.make_params(spatial=int(sp), temporal=int(tp), cfl=cfl, time=0.0, vel=vel.as_dict())
    for step in range(2):
        say(f"rough1024 step {step}")
        a.apply()
        rc, r = po.advect(o, p)
        assert rc == 0
        assert a.getAdvectedTime() == r["advected_time"]
        assert_same(o, g, f"rough1024 step {step + 1}")


# ------------------------------------------------------------- batch512 ----

BATCH_STEPS = 3
N_PICK = 4  # largest and smallest


@pytest.fixture(scope="module")
def batch_state():
    vls = _vls()
    say("batch512: building 512 level sets (seed 42)")
    sets, vels = bench.build_batch(vls, 512)
    sizes = np.array([d.getNumberOfPoints() for d in sets])
    order = np.argsort(sizes, kind="stable")
    picks = [int(k) for k in list(order[:N_PICK]) + list(order[-N_PICK:])]
    inputs = {k: (sets[k].get_runs(), sets[k].getLevelSetWidth()) for k in picks}
    bat = vls.BatchAdvect(sets, vels, spatial=vls.SpatialSchemeEnum.WENO_5TH_ORDER,
                          temporal=vls.TemporalSchemeEnum.RUNGE_KUTTA_3RD_ORDER, cfl=0.4)
    for s in range(BATCH_STEPS):
        say(f"batch512 step {s}")
        bat.apply()
    times = bat.advectedTimes()
    bat.writeBack()
    outputs = {k: sets[k].get_runs() for k in picks}
    meta = {k: (sets[k].D, list(sets[k].gmin), list(sets[k].gmax), list(sets[k].bc), sets[k].gridDelta,
                vels[k].as_dict(), int(sizes[k])) for k in picks}
    del bat
    return dict(picks=picks, inputs=inputs, outputs=outputs, times=times, meta=meta)


@pytest.mark.parametrize("j", range(2 * N_PICK))
def test_batch512_level_set_against_oracle(batch_state, j):
    st = batch_state
    k = st["picks"][j]
    D, gmin, gmax, bc, dx, vel, size = st["meta"][k]
    say(f"batch512 level set {k} ({size} points)")
    (idx, val), width = st["inputs"][k]
    o = po.Domain(D, gmin, gmax, bc, dx)
    o.set(idx, val, width)
    p = po.make_params(spatial=11, temporal=2, cfl=0.4, time=0.0, vel=vel)
    t = 0.0
    for _ in range(BATCH_STEPS):
        rc, r = po.advect(o, p)
        assert rc == 0
        t += r["advected_time"]
    io, vo = o.get()
    ig, vg = st["outputs"][k]
    if not cases.bit_equal(io, vo, ig, vg):
        i, a, b, c, d = cases.first_diff(io, vo, ig, vg)
        pytest.fail(f"level set {k}: run lists differ (oracle {len(vo)} vs gpu {len(vg)}), first diff at {i}: "
                    f"oracle {a} {b} gpu {c} {d}")
    assert st["times"][k] == t, f"level set {k}: advected time gpu {st['times'][k]!r} oracle {t!r}"
