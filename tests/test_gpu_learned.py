"""Learned WENO5 weights on the GPU: inference and training kernels against
the oracle (oracle/mlp_oracle.cpp), and the learned scheme inside Advect."""
import numpy as np
import pytest

import pyoracle as po

pytestmark = pytest.mark.gpu


def rand_params(seed, scale=0.5):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(po.MLP_NPARAM) * scale).astype(np.float32)


def smooth_stencils(seed, n):
    rng = np.random.default_rng(seed)
    xs = np.arange(-3, 4, dtype=np.float64)
    c = rng.standard_normal((n, 4)) * np.array([1.0, 0.05, 0.01, 0.002])
    return c[:, :1] * xs + c[:, 1:2] * xs ** 2 + c[:, 2:3] * xs ** 3 + c[:, 3:4] * xs ** 4 + rng.standard_normal((n, 1))


def test_learned_weno5_inference_matches_oracle():
    import viennals_amd as vls
    P = rand_params(11)
    st = np.concatenate([smooth_stencils(1, 3000),
                         np.abs(np.arange(-3, 4) - np.random.default_rng(2).uniform(-3, 3, (500, 1))),
                         np.zeros((4, 7))])
    for plus in (False, True):
        gpu = vls.weno5_learned(st, 0.5, plus, P)
        cpu = np.array([po.weno5_learned(s, 0.5, plus, po.bf16_round(P)) for s in st])
        plain = np.array([po.weno5(s, 0.5, plus) for s in st])
        fallback = cpu == plain
        # kink / zero stencils take plain weno5: bit-exact
        assert np.array_equal(gpu[fallback], cpu[fallback])
        # network path: fp32 network, transcendental functions may differ by ulps
        err = np.abs(gpu - cpu) / np.maximum(1.0, np.abs(cpu))
        assert err.max() < 2e-6, err.max()
        assert (~fallback).sum() > 1000  # most smooth stencils use the network


def teacher_samples(seed, n, teacher):
    rng = np.random.default_rng(seed)
    X = np.zeros((n, 8), np.float32)
    d = rng.standard_normal((n, 5))
    d /= np.abs(d).max(axis=1, keepdims=True)
    X[:, :5] = d
    p = rng.uniform(-1, 1, (n, 3))
    X[:, 5:] = p
    w = np.array([po.mlp_forward(teacher, x) for x in X[:, :5]])
    T = (w * p).sum(1).astype(np.float32)
    return X, T


def test_training_gradient_matches_oracle():
    from viennals_amd.learned import Trainer
    P = rand_params(21, 0.3)
    X, T = teacher_samples(22, 1000, rand_params(23, 0.5))
    tr = Trainer(P)
    tr.set_samples(X, T)
    for off, cnt in ((0, 1000), (100, 37)):
        loss, g = tr.loss_grad(off, cnt)
        lref, gref = po.mlp_loss_grad(P, X[off:off + cnt], T[off:off + cnt], emulate=True)
        assert abs(loss - lref) <= 1e-4 * abs(lref), (loss, lref)
        scale = np.max(np.abs(gref))
        err = np.abs(g - gref)
        # bf16 operands: a last-bit difference of an fp32 activation can flip
        # one bf16 rounding; the fp32 accumulation order differs
        assert err.max() <= 3e-3 * scale, (err.max(), scale, int(np.argmax(err)))
        assert np.linalg.norm(g - gref) <= 1e-3 * np.linalg.norm(gref)


def test_training_reduces_loss_on_teacher_task():
    from viennals_amd.learned import Trainer, init_params
    teacher = rand_params(31, 0.4)
    X, T = teacher_samples(32, 20000, teacher)
    Xv, Tv = teacher_samples(33, 2000, teacher)
    tr = Trainer(init_params(0))
    tr.set_samples(Xv, Tv)
    l0, _ = tr.loss_grad()
    tr.set_samples(X, T)
    tr.train(600, 4096, 3e-3)
    tr.set_samples(Xv, Tv)
    l1, _ = tr.loss_grad()
    assert l1 < 0.2 * l0, (l0, l1)
    # the trained parameters evaluate to the same loss on the oracle
    lref, _ = po.mlp_loss_grad(tr.params(), Xv, Tv, emulate=True)
    assert abs(l1 - lref) <= 1e-3 * lref


def test_learned_scheme_step_matches_oracle():
    """One RK3 step of a star shape with the learned scheme: GPU vs oracle
    (values agree to 1e-5 grid units, same defined set: the network takes
    its inputs and activations as bf16, and a last-bit difference of an fp32
    value -- fused multiply-adds on the GPU, the hardware exp and reciprocal
    -- can move one bf16 rounding)."""
    import viennals_amd as vls
    from viennals_amd.learned import StarShape
    shape = StarShape(12.0, seed=5)
    idx, val = shape.band(1.0)
    P = rand_params(41, 0.3)
    d = vls.Domain(gridDelta=1.0, dim=3)
    d.insertPoints(idx, val, 2)
    a = vls.Advect(d, vls.VelocityField.constant(0.7))
    a.setSpatialScheme(vls.SpatialSchemeEnum.WENO_5TH_ORDER_LEARNED)
    a.setTemporalScheme(vls.TemporalSchemeEnum.RUNGE_KUTTA_3RD_ORDER)
    a.setTimeStepRatio(0.4)
    a.setLearnedWeights(P)
    a.apply()
    gi, gv = d.get_runs()
    o = po.Domain(3, d.gmin, d.gmax, d.bc, 1.0)
    o.from_points(idx, val, 2)
    prm = po.make_params(spatial=12, temporal=2, cfl=0.4, vel=dict(kind=0, b=0.7), mlp=po.bf16_round(P))
    st, r = po.advect(o, prm)
    assert st == 0
    oi, ov = o.get()
    assert abs(a.getAdvectedTime() - r["advected_time"]) <= 1e-6 * r["advected_time"]
    assert np.array_equal(gi, oi)
    both = (np.abs(gv) < 1e300) & (np.abs(ov) < 1e300)
    assert np.array_equal(np.abs(gv) < 1e300, np.abs(ov) < 1e300)
    assert np.max(np.abs(gv[both] - ov[both])) < 1e-5


def test_learned_scheme_requires_weights():
    import viennals_amd as vls
    d = vls.Domain(gridDelta=1.0, dim=3)
    vls.MakeGeometry(d, vls.Sphere([0.0, 0.0, 0.0], 5.0)).apply()
    a = vls.Advect(d, vls.VelocityField.constant(1.0))
    a.setSpatialScheme(vls.SpatialSchemeEnum.WENO_5TH_ORDER_LEARNED)
    with pytest.raises(vls.LsgError, match="network parameters"):
        a.apply()


def test_trained_network_gain_over_plain_coarse():
    """The committed network (trained on fine WENO5-RK3 runs by the GPU
    trainer) against plain WENO5 on the same coarse grid, both judged against
    the fine plain run of the same star shape sampled on the coarse grid,
    with CompareChamfer and CompareVolume: on this shape it reduces both the
    Chamfer distance and the volume mismatch (the wall-time ratio to the fine
    run is printed, not asserted: it is a timing)."""
    import os
    import viennals_amd as vls
    from viennals_amd import learned as L
    P = np.load(os.path.join(os.path.dirname(vls.__file__), "data", "weno5_learned.npy"))
    rng = np.random.default_rng(1234)
    shape = L.StarShape(32.0, seed=int(rng.integers(1 << 30)))
    vel = L.random_velocity(rng)
    r = L.compare(shape, vel, 6.0, P)
    plain, learned = r["coarse_plain"], r["coarse_learned"]
    print("chamfer plain %.5f learned %.5f; volume error plain %.3e learned %.3e; speedup vs fine %.2f" %
          (plain["chamfer"], learned["chamfer"], plain["volume_error"], learned["volume_error"],
           learned["speedup_vs_fine"]))
    assert learned["chamfer"] < plain["chamfer"]
    assert learned["volume_mismatch"] <= plain["volume_mismatch"]
