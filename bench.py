"""Benchmark of the level-set advection step on one MI355X.

A "step" is one advection time step (Advect::apply with advectionTime 0: one
CFL-limited step; three stages for RK3) over the whole band.  Default
workload: the 3D square-trench etch (256x256 lateral cells, gridDelta 1,
reflective x/y, infinite z, plane z=0 minus a centred 64x64x128 trench,
directional etch velocity -(max(0,n_z) + 0.1)), WENO5 + RK3, CFL 0.4, FP64.

value = point-updates/s: active points (|phi| <= 0.5) summed over the RK
stages of the K timed steps, divided by the wall time of those steps with
the band resident in HBM.  Two kernels of the step are then timed with HIP
events on the library's stream over K more steps: the rule evaluation of
the star band operations (Expand cycles and rebuilds; kEvalKeep<EvalStar>,
algorithmic bytes 8 B per input entry and 8 B per candidate; with
--single-pass the whole single-pass operation kBandOp<AtStar>, 12 B per
input entry read and per output entry written, value and position, plus
24 B per bitmap word for both bitmaps with their prefix counts) and the rate
(Hamiltonian) kernel (8 B per band entry read plus 8 B per active point
written).  The one with more time is reported as `roofline`.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, 8 TB/s


def build_workload(vls, name):
    BC = vls.BoundaryConditionEnum
    if name == "trench256":
        L = 128
        d = vls.Domain([-L, L, -L, L, -L, L], [BC.REFLECTIVE_BOUNDARY, BC.REFLECTIVE_BOUNDARY,
                                               BC.INFINITE_BOUNDARY], 1.0)
        vls.MakeGeometry(d, vls.Plane([0.0, 0.0, 0.0], [0.0, 0.0, 1.0])).apply()
        box = vls.Domain([-L, L, -L, L, -L, L], [BC.REFLECTIVE_BOUNDARY, BC.REFLECTIVE_BOUNDARY,
                                                 BC.INFINITE_BOUNDARY], 1.0)
        vls.MakeGeometry(box, vls.Box([-32.0, -32.0, -128.0], [32.0, 32.0, 1.0])).apply()
        vls.BooleanOperation(d, box, vls.BooleanOperationEnum.RELATIVE_COMPLEMENT).apply()
        vel = vls.VelocityField.normal_power(dir=2, k=1, b=-0.1, a_pos=-1.0, a_neg=0.0)
        cfg = dict(workload="trench256_weno5_rk3",
                   domain="256x256 lateral cells, reflective x/y, infinite z, gridDelta 1.0",
                   geometry="plane z=0 minus centred 64x64x128 box trench",
                   velocity="-(1.0*max(0,n_z) + 0.1)", spatial="WENO_5TH_ORDER",
                   temporal="RUNGE_KUTTA_3RD_ORDER", cfl=0.4)
        return d, vel, vls.SpatialSchemeEnum.WENO_5TH_ORDER, vls.TemporalSchemeEnum.RUNGE_KUTTA_3RD_ORDER, 0.4, cfg
    if name == "rough1024":
        import numpy as np
        N = 1024
        rng = np.random.Generator(np.random.PCG64(1234))
        white = rng.standard_normal((N, N))
        k = np.fft.fftfreq(N)
        kx, ky = np.meshgrid(k, k, indexing="ij")
        ell = 16.0
        kern = np.exp(-0.5 * (2 * np.pi) ** 2 * (kx ** 2 + ky ** 2) * ell ** 2)
        h = np.real(np.fft.ifft2(np.fft.fft2(white) * kern))
        h = 4.0 * (h - h.mean()) / h.std()
        hx = 0.5 * (np.roll(h, -1, 0) - np.roll(h, 1, 0))
        hy = 0.5 * (np.roll(h, -1, 1) - np.roll(h, 1, 1))
        gn = np.sqrt(1.0 + hx ** 2 + hy ** 2)
        zlo = int(np.floor((h - gn).min())) - 1
        zhi = int(np.ceil((h + gn).max())) + 1
        xs, ys = np.meshgrid(np.arange(N), np.arange(N), indexing="ij")
        pts, vals = [], []
        for z in range(zlo, zhi + 1):
            phi = (z - h) / gn
            m = np.abs(phi) <= 1.0
            pts.append(np.stack([xs[m], ys[m], np.full(int(m.sum()), z)], 1))
            vals.append(phi[m])
        idx = np.concatenate(pts).astype(np.int32)
        val = np.concatenate(vals)
        d = vls.Domain([0, N, 0, N, -64, 64], [BC.PERIODIC_BOUNDARY, BC.PERIODIC_BOUNDARY,
                                               BC.INFINITE_BOUNDARY], 1.0)
        d.insertPoints(idx, val, 2)
        vel = vls.VelocityField.normal_power(dir=2, k=2, b=0.05, a_pos=1.0, a_neg=0.0)
        cfg = dict(workload="rough1024_llf2_rk3",
                   domain="1024x1024 periodic x/y, infinite z, gridDelta 1.0",
                   geometry="Gaussian random height field (PCG64 seed 1234, correlation 16, sigma 4)",
                   velocity="max(0,n_z)^2 + 0.05", spatial="LOCAL_LAX_FRIEDRICHS_2ND_ORDER",
                   temporal="RUNGE_KUTTA_3RD_ORDER", cfl=0.45)
        return d, vel, vls.SpatialSchemeEnum.LOCAL_LAX_FRIEDRICHS_2ND_ORDER, \
            vls.TemporalSchemeEnum.RUNGE_KUTTA_3RD_ORDER, 0.45, cfg
    if name == "sphere_eo1":
        d = vls.Domain([-20, 20] * 3, [BC.INFINITE_BOUNDARY] * 3, 0.25)
        vls.MakeGeometry(d, vls.Sphere([0.0, 0.0, 0.0], 10.0)).apply()
        vel = vls.VelocityField.constant(1.0)
        cfg = dict(workload="sphere_r10_eo1_fe", domain="infinite, gridDelta 0.25", geometry="sphere r=10",
                   velocity="+1", spatial="ENGQUIST_OSHER_1ST_ORDER", temporal="FORWARD_EULER", cfl=0.4)
        return d, vel, 0, 0, 0.4, cfg
    raise SystemExit(f"unknown workload {name}")


def build_batch(vls, n=512, seed=42):
    """512 independent level sets of uneven size: half spheres (radius
    U[8,64] cells, constant speed), half square trenches (width U[16,96],
    depth U[32,192], lateral domain U[128,512] cells, reflective x/y,
    directional etch); speeds U[0.2,1.0]; gridDelta 1."""
    BC = vls.BoundaryConditionEnum
    rng = np.random.default_rng(seed)
    sets, vels = [], []
    for k in range(n):
        rate = float(rng.uniform(0.2, 1.0))
        if k % 2 == 0:
            r = float(rng.uniform(8, 64))
            d = vls.Domain(gridDelta=1.0, dim=3)
            vls.MakeGeometry(d, vls.Sphere([0.0, 0.0, 0.0], r)).apply()
            vels.append(vls.VelocityField.constant(rate))
        else:
            w = float(rng.uniform(16, 96))
            depth = float(rng.uniform(32, 192))
            half = int(rng.uniform(128, 512)) // 2
            w = min(w, half)
            bounds = [-half, half, -half, half, -half, half]
            bcs = [BC.REFLECTIVE_BOUNDARY, BC.REFLECTIVE_BOUNDARY, BC.INFINITE_BOUNDARY]
            d = vls.Domain(bounds, bcs, 1.0)
            vls.MakeGeometry(d, vls.Plane([0.0, 0.0, 0.0], [0.0, 0.0, 1.0])).apply()
            box = vls.Domain(bounds, bcs, 1.0)
            vls.MakeGeometry(box, vls.Box([-w / 2, -w / 2, -depth], [w / 2, w / 2, 1.0])).apply()
            vls.BooleanOperation(d, box, vls.BooleanOperationEnum.RELATIVE_COMPLEMENT).apply()
            vels.append(vls.VelocityField.normal_power(dir=2, k=1, b=-0.1 * rate, a_pos=-rate, a_neg=0.0))
        sets.append(d)
    return sets, vels


def under_profiler():
    """True when a rocprofv3 tool library is preloaded into this process."""
    return "rocprofiler" in os.environ.get("LD_PRELOAD", "")


def run_batch(args):
    if args.seq_steps > 0 and under_profiler():
        # one profiled run of this leg (64 graph-captured Advects after a
        # batch) aborted inside a HIP call of the step launch; its cause is
        # not known, so the leg does not run under the profiler
        print("bench.py: the sequential per-level-set leg does not run under rocprofv3; pass --seq-steps 0",
              file=sys.stderr)
        return 2
    import viennals_amd as vls
    t0 = time.perf_counter()
    sets, vels = build_batch(vls, args.batch_size)
    setup = time.perf_counter() - t0
    W = vls.SpatialSchemeEnum.WENO_5TH_ORDER
    R3 = vls.TemporalSchemeEnum.RUNGE_KUTTA_3RD_ORDER
    # sequential per-case launches (same code path, one Advect per level set)
    seq_sets = []
    for d in (sets if args.seq_steps > 0 else []):
        c = vls.Domain.from_index_bounds(d.D, d.gmin, d.gmax, d.bc, d.gridDelta)
        c.deepCopy(d)
        seq_sets.append(c)
    bat = vls.BatchAdvect(sets, vels, spatial=W, temporal=R3, cfl=0.4)
    blocks, tune_ms = bat.autotune()
    for _ in range(args.warmup):
        bat.apply()
    vls.synchronize()
    pu0, _ = bat.pointUpdates()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        bat.apply()
    vls.synchronize()
    wall = time.perf_counter() - t0
    pu = bat.pointUpdates()[0] - pu0
    del bat
    seq = None
    if args.seq_steps > 0:
        seq = sequential(vls, seq_sets, vels, W, R3, args.seq_steps)
    rec = {
        "metric": "aggregate point-updates/s (batched 512 level sets, WENO5-RK3 3D)",
        "value": pu / wall,
        "unit": "point-updates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"batch{args.batch_size}_weno5_rk3",
                   "level_sets": "half spheres r~U[8,64], half trenches w~U[16,96] d~U[32,192] L~U[128,512]",
                   "seed": 42, "active_point_updates_per_step": pu / args.steps, "setup_seconds": setup,
                   "autotuned_scan_blocks": blocks, "autotune_best_step_ms": tune_ms},
        "sequential": seq,
        "speedup_vs_sequential": (pu / wall) / seq["value"] if seq else None,
    }
    print(json.dumps(rec))
    return 0


def sequential(vls, seq_sets, vels, W, R3, steps):
    """One Advect per level set, launched one after another (graph-captured steps)."""
    advs = []
    for d, v in zip(seq_sets, vels):
        a = vls.Advect(d, v)
        a.setSpatialScheme(W)
        a.setTemporalScheme(R3)
        a.setTimeStepRatio(0.4)
        advs.append(a)
    for a in advs:  # warm-up step (captures each level set's graphs)
        a.apply()
    vls.synchronize()
    seq_pu = 0.0
    t0 = time.perf_counter()
    for _ in range(steps):
        for a in advs:
            a.apply()
            seq_pu += a.getPointUpdates()
    vls.synchronize()
    seq_wall = time.perf_counter() - t0
    return {"value": seq_pu / seq_wall, "unit": "point-updates/s", "steps": steps,
            "ms_per_step": seq_wall * 1e3 / steps,
            "what": "one Advect (graph-captured steps) per level set, launched one after another"}


def measured_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` (file tag: kRates, kBandOpStar) from
    the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary of this workload
    (scaled by the copy-kernel calibration), or None."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}_traffic.md")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":")[1])
    return None


def cpu_baseline(idx, val, width, grid, params, steps, threads):
    """The oracle (CPU restatement, OpenMP) on the same input band: `steps`
    steps from the initial band with `threads` threads; returns
    (point-updates/s, seconds)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    po.set_threads(threads)
    D, gmin, gmax, bc, dx = grid
    o = po.Domain(D, gmin, gmax, bc, dx)
    o.set(idx, val, width)
    pu = 0.0
    t0 = time.perf_counter()
    for _ in range(steps):
        st, r = po.advect(o, params)
        pu += r["point_updates"]
    dt = time.perf_counter() - t0
    return pu / dt, dt


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def host_threads():
    """Threads for the CPU baseline: the CPUs this process may run on."""
    try:
        return len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        return os.cpu_count() or 1


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="trench256")
    ap.add_argument("--cpu-steps", type=int, default=4)
    ap.add_argument("--cpu-steps-1t", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batch-size", type=int, default=512)
    ap.add_argument("--seq-steps", type=int, default=2)
    ap.add_argument("--single-pass", action="store_true",
                    help="band operations as single-pass kernels (LSG_FUSED_BANDOPS=1 does the same)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the path is single-GPU; "
              "multi-GPU sharding is described in DESIGN.md)", file=sys.stderr)
        return 2

    single_pass = args.single_pass or os.environ.get("LSG_FUSED_BANDOPS", "") == "1"
    if single_pass:
        from viennals_amd._lib import check, lib
        check(lib().lsg_set_single_pass_band_ops(1))
    if args.workload == "batch512":
        return run_batch(args)

    import viennals_amd as vls

    d, vel, spatial, temporal, cfl, cfg = build_workload(vls, args.workload)
    cfg["band_operations"] = "single-pass" if single_pass else "multi-kernel"
    idx0, val0 = d.get_runs()
    width0 = d.getLevelSetWidth()

    adv = vls.Advect(d, vel)
    adv.setSpatialScheme(spatial)
    adv.setTemporalScheme(temporal)
    adv.setTimeStepRatio(cfl)
    adv.setAdvectionTime(0.0)
    for _ in range(args.warmup):
        adv.apply()
    vls.synchronize()

    pu = 0.0
    t_adv = 0.0
    vls.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        adv.apply()
        pu += adv.getPointUpdates()
        t_adv += adv.getAdvectedTime()
    vls.synchronize()
    t1 = time.perf_counter()
    wall = t1 - t0
    # Kernel durations: HIP events around every launch of the two measured
    # kernels, recorded on the library stream over another K steps (the
    # event records add launch gaps, so they stay out of the timed steps).
    adv.setProfiling(True)
    for _ in range(args.steps):
        adv.apply()
    vls.synchronize()
    value = pu / wall

    def roofline(tag, name, stats):
        ms, launches, nbytes = stats
        achieved = (nbytes / (ms * 1e-3)) / 1e9 if ms > 0 else 0.0
        return {
            "bound": "hbm",
            "kernel": name,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(tag, cfg["workload"]),
            "launches": launches,
            "avg_launch_us": (ms * 1e3 / launches) if launches else None,
            "algorithmic_bytes_per_launch": (nbytes / launches) if launches else None,
            "total_ms": ms,
        }

    # the dominant kernel of the step (most time over the timed steps) gets
    # "roofline"; the other measured kernel is reported beside it
    if single_pass:
        star = roofline("kBandOpStar", "kBandOp<AtStar> (single-pass Expand / rebuild band operation)",
                        adv.getStarKernelStats())
    else:
        star = roofline("kEvalKeepStar", "kEvalKeep<EvalStar> (Expand / rebuild rule evaluation)",
                        adv.getStarKernelStats())
    rate = roofline("kRates", "kRates (Hamiltonian + CFL reduction)", adv.getRateKernelStats())
    measured = [star, rate]
    if not single_pass:
        measured.append(roofline("kEvalKeepStar2", "kEvalKeep<EvalStar2> (two Expand cycles at once)",
                                 adv.getExpand2KernelStats()))
    measured = [m for m in measured if m["launches"]]
    measured.sort(key=lambda m: -m["total_ms"])
    dominant, others = measured[0], measured[1:]

    rec = {
        "metric": "point-updates/s + %HBM peak (WENO5-RK3 3D)",
        "value": value,
        "unit": "point-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": dict(cfg, active_point_updates_per_step=pu / args.steps,
                       advected_time_per_step=t_adv / args.steps, initial_runs=int(len(val0))),
        "roofline": dominant,
        "roofline_other_kernels": others,
    }
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as po
        threads = host_threads()
        params = po.make_params(spatial=spatial, temporal=temporal, cfl=cfl, time=0.0, vel=vel.as_dict())
        grid = (d.D, d.gmin, d.gmax, d.bc, d.gridDelta)
        runs = [cpu_baseline(idx0, val0, width0, grid, params, args.cpu_steps, threads) for _ in range(3)]
        values = sorted(r[0] for r in runs)
        one_value, one_wall = cpu_baseline(idx0, val0, width0, grid, params, args.cpu_steps_1t, 1)
        rec["cpu_baseline"] = {
            "value": values[1],
            "unit": "point-updates/s",
            "cores": threads,
            "kind": "port",
            "sample": f"median of 3 runs of the first {args.cpu_steps} steps of the same workload from the same "
                      f"initial band (OpenMP, {threads} threads; "
                      f"{', '.join(f'{r[1]:.1f}' for r in runs)} s)",
            "runs": values,
            "single_thread": {"value": one_value, "steps": args.cpu_steps_1t, "seconds": one_wall},
            "cpu_model": cpu_model(),
        }
    print(json.dumps(rec))
    return 0


if __name__ == "__main__":
    sys.exit(main())
