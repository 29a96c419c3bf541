#!/usr/bin/env python3
"""Benchmark: KP07_dimsplit shallow water, 4096x4096 float32, on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--schemes kp07,hll2,waf,ee2d,lxf,hll,force,kp07u]
                    [--no-cpu-baseline] [--ensemble 512] [--serve 2000 --serve-rate 300]

A "step" is one full time step of the plain Simulator API (two fused substep
launches: x-y then y-x sweeps), on the seeded dam break with 16 Gaussian bumps,
reflective walls, g = 9.81, theta = 1.3, at a dt pinned to 0.45 x the initial
CFL.  The state stays resident in HBM; the timed region holds exactly K steps
between device synchronisations, timed with HIP events on the simulator's
stream.  One JSON line is printed:

* value / unit      cell-updates per second (cells x full steps / s)
* roofline          algorithmic HBM bytes of one substep launch (24 B per cell:
                    3 fields read + written) / average launch time, against
                    8 TB/s; `traffic` = measured HBM bytes per launch from the
                    rocprofv3 counter summary in profiles/ when present
* cpu_baseline      the float32 NumPy oracle (one core) on a bounded sample of
                    the same workload (512x512 crop of the generator, a few steps)
* schemes           the same measurement for HLL2 and WAF at 4096^2,
                    EE2D_KP07_dimsplit at 8192^2 (64 B per cell-step) and the
                    reference's other schemes LxF, HLL, FORCE (24 B per
                    cell-step) and the unsplit KP07 with RK2 (60 B) at 4096^2
* ensemble          512 dam breaks of uneven size in the batched engine against
                    the plain simulate() loop over the same instances
* serving           the continuous-batching server on 2000 Poisson-arriving
                    requests of mixed scheme and size: p50/p99 latency and
                    aggregate cell-updates/s against sequential plain solves
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0


def _parse():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[1])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=4096, help="grid side of the SWE workload")
    ap.add_argument("--schemes", default="kp07,hll2,waf,ee2d,lxf,hll,force,kp07u,waf_fast,ee2d_fast")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--tile-rows", type=int, default=0)
    ap.add_argument("--waves-per-block", type=int, default=0)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 256 MiB device copy kernel 5 times (HBM counter calibration)")
    ap.add_argument("--ensemble", type=int, default=512,
                    help="instances of the batched-ensemble leg (0 = skip the leg)")
    ap.add_argument("--serve", type=int, default=2000, help="requests of the serving leg (0 = skip the leg)")
    ap.add_argument("--serve-rate", type=float, default=300.0,
                    help="Poisson arrival rate, requests/s (default just above the measured capacity, ~267/s)")
    ap.add_argument("--sequential", type=int, default=-1,
                    help="requests of the sequential serving baseline (-1 = all; the first K by arrival)")
    return ap.parse_args()


def _copy_calibration(ctx):
    """A streaming copy of a known size: 256 MiB read + 256 MiB written per launch."""
    import ctypes
    from finitevolumegpu_amd import _lib
    from finitevolumegpu_amd.HipContext import Event
    n = 64 * 1024 * 1024
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    _lib.call("fvmi_malloc", ctypes.byref(a), 4 * n)
    _lib.call("fvmi_malloc", ctypes.byref(b), 4 * n)
    _lib.call("fvmi_fill_f32", a, 4 * 4096, 4096, n // 4096, 1.0, None)
    s, e = Event(), Event()
    s.record()
    for _ in range(5):
        _lib.call("fvmi_copy_f32", b, a, n, None)
    e.record()
    e.synchronize()
    ms = e.time_since(s) / 5
    print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "avg_launch_ms": ms,
                      "gbs": 8 * n / (ms * 1e-3) / 1e9}), flush=True)
    _lib.call("fvmi_free", a)
    _lib.call("fvmi_free", b)


def _time_steps(sim, dt, steps, warmup, ctx):
    from finitevolumegpu_amd.HipContext import Event
    for _ in range(warmup):
        sim.step(dt)
    ctx.synchronize()
    start, end = Event(), Event()
    wall0 = time.perf_counter()
    start.record(sim.stream)
    for _ in range(steps):
        sim.step(dt)
    end.record(sim.stream)
    end.synchronize()
    ctx.synchronize()
    wall = time.perf_counter() - wall0
    return end.time_since(start), wall


def _steady_ms(sim, dt, steps):
    """Median time of one step over ``steps`` further steps, each between two
    events: the rate once the shader clock has settled (reported beside the
    timed region, never as ``value``)."""
    from finitevolumegpu_amd.HipContext import Event
    evs = [Event() for _ in range(steps + 1)]
    evs[0].record(sim.stream)
    for i in range(steps):
        sim.step(dt)
        evs[i + 1].record(sim.stream)
    evs[-1].synchronize()
    return float(np.median([evs[i + 1].time_since(evs[i]) for i in range(steps)]))


def _swe_sim(ctx, scheme, n, args, math=None):
    from finitevolumegpu_amd.initial_conditions import dam_break_bumps
    h, hu, hv = dam_break_bumps(n, seed=0)
    kw = {"math": math}
    if args.tile_rows:
        kw["tile_rows"] = args.tile_rows
    if args.waves_per_block:
        kw["waves_per_block"] = args.waves_per_block
    if scheme == "kp07":
        from finitevolumegpu_amd.KP07_dimsplit import KP07_dimsplit as C
    elif scheme == "hll2":
        from finitevolumegpu_amd.HLL2 import HLL2 as C
    elif scheme == "waf":
        from finitevolumegpu_amd.WAF import WAF as C
    elif scheme == "lxf":
        from finitevolumegpu_amd.LxF import LxF as C
    elif scheme == "hll":
        from finitevolumegpu_amd.HLL import HLL as C
    elif scheme == "force":
        from finitevolumegpu_amd.FORCE import FORCE as C
    elif scheme == "kp07u":
        from finitevolumegpu_amd.KP07 import KP07 as C
    else:
        raise ValueError("unknown scheme %r" % scheme)
    sim = C(ctx, h, hu, hv, n, n, 1.0, 1.0, 9.81, **kw)
    dt = np.float32(sim.computeDt() * 0.9)
    return sim, dt


# HBM bytes per cell of one full step as the algorithm needs them (fields read
# once and written once per launch; the second RK2 stage of the unsplit KP07
# also reads the output buffer), and launches per step
STEP_BYTES = {"kp07": (48, 2), "hll2": (48, 2), "waf": (48, 2), "lxf": (24, 1), "hll": (24, 1), "force": (24, 1),
              "kp07u": (60, 2), "ee2d": (64, 2)}


def _euler_sim(ctx, n, math=None):
    """One rank's 8192^2 subdomain: Rayleigh-Taylor, reflective north/south,
    periodic east/west through the one-rank MPISimulator (Dirichlet edges and
    an on-device halo exchange after every substep)."""
    from finitevolumegpu_amd.EE2D_KP07_dimsplit import EE2D_KP07_dimsplit
    from finitevolumegpu_amd.MPISimulator import MPISimulator
    from finitevolumegpu_amd.Simulator import BoundaryCondition
    from finitevolumegpu_amd.initial_conditions import rayleigh_taylor
    rho, ru, rv, E, dx, dy = rayleigh_taylor(n, n, seed=0)
    T = BoundaryCondition.Type
    bc = BoundaryCondition({"north": T.Reflective, "south": T.Reflective, "east": T.Periodic, "west": T.Periodic})
    sim = MPISimulator(EE2D_KP07_dimsplit(ctx, rho, ru, rv, E, n, n, dx, dy, 0.1, 1.4, boundary_conditions=bc,
                                          math=math))
    c = np.sqrt(1.4 * 2.6 / 1.0)
    dt = np.float32(0.45 * dx / c)
    return sim, dt


def _measure(ctx, kind, args):
    """One scheme's timing; a kind ending in "_fast" runs that scheme in the
    fast arithmetic policy instead of its default (Simulator.DEFAULT_MATH)."""
    scheme = kind[:-5] if kind.endswith("_fast") else kind
    math = "fast" if kind.endswith("_fast") else None
    if scheme == "ee2d":
        n = 8192
        sim, dt = _euler_sim(ctx, n, math)
        steps = max(1, args.steps // 3)
    else:
        n = args.n
        sim, dt = _swe_sim(ctx, scheme, n, args, math)
        steps = args.steps
    step_bytes, per_step = STEP_BYTES[scheme]
    ms, wall = _time_steps(sim, dt, steps, args.warmup, ctx)
    steady = _steady_ms(sim, dt, 30 if scheme == "ee2d" else 100)
    out = sim.download()
    finite = all(bool(np.isfinite(a).all()) for a in out)
    launches = per_step * steps
    avg_launch_ms = ms / launches
    cells = n * n
    res = {
        "n": n, "steps": steps, "ms_per_step": ms / steps, "cell_updates_per_s": cells * steps / (ms * 1e-3),
        "avg_launch_ms": avg_launch_ms,
        "algorithmic_gbs": step_bytes * cells * steps / (ms * 1e-3) / 1e9,
        "wall_s": wall, "finite": finite, "math": getattr(sim, "sim", sim).math,
        "steady_ms_per_step": steady,
        "launch": [getattr(sim, "sim", sim).launch.tile_rows, getattr(sim, "sim", sim).launch.waves_per_block],
    }
    del sim
    return res


def _cpu_baseline(budget_s=12.0):
    """Float32 NumPy oracle, one core, 512x512 crop of the same generator."""
    from finitevolumegpu_amd.initial_conditions import dam_break_bumps
    from oracle import fv_oracle as O
    n = 512
    h, hu, hv = dam_break_bumps(n, seed=0)
    sim = O.OracleSimulator("kp07_dimsplit", [h, hu, hv], n, n, 1.0, 1.0, 9.81, dtype=np.float32)
    dt = np.float32(0.45 * float(sim.cfl))
    steps = 0
    t0 = time.perf_counter()
    while True:
        sim.step(dt)
        steps += 1
        el = time.perf_counter() - t0
        if el > budget_s or steps >= 200:
            break
    return {"value": n * n * steps / el, "unit": "cell-updates/s", "cores": 1, "kind": "port",
            "sample": "NumPy float32 oracle, KP07_dimsplit %dx%d dam break + bumps, %d full steps in %.1f s "
                      "(host os.cpu_count()=%s)" % (n, n, steps, el, os.cpu_count())}


def _ensemble_leg(ctx, count):
    """512 KP07_dimsplit dam breaks (side log-uniform in [64, 1024], per-instance
    CFL dt to t_end U(5, 50)) in one batched engine, against the plain
    Simulator.simulate() loop over the same instances.  Inputs are resident
    before either clock starts; step counts must agree for every instance and
    the fields bitwise for every 32nd instance."""
    from finitevolumegpu_amd.Autotuner import Autotuner
    from finitevolumegpu_amd.Ensemble import Ensemble, _bc_code
    from finitevolumegpu_amd.KP07_dimsplit import KP07_dimsplit
    from finitevolumegpu_amd.Simulator import BoundaryCondition
    from finitevolumegpu_amd.initial_conditions import ensemble_instances
    insts = ensemble_instances(count, seed=0)
    specs = [{"fields": [d["h"], d["hu"], d["hv"]], "nx": d["nx"], "ny": d["ny"], "dx": 1.0, "dy": 1.0,
              "g": 9.81, "t_end": d["t_end"], "bc": d["bc"]} for d in insts]
    # schedule from the autotuner, timed on every 4th instance (untimed here);
    # fixed segment lengths only: the balanced plan's length depends on the
    # instance count, so a sample does not represent it
    sched = Autotuner.ensemble_schedule(ctx, "kp07_dimsplit", specs[::4], seg_rows=(16, 24, 32, 48), blocks=4)
    print("# ensemble schedule: %s" % json.dumps(sched), file=sys.stderr, flush=True)
    ens = Ensemble(ctx, "kp07_dimsplit", specs, steps_per_graph=8, seg_rows=sched["seg_rows"],
                   waves_per_block=sched["waves_per_block"])
    ctx.synchronize()
    t0 = time.perf_counter()
    blocks = ens.run()
    ctx.synchronize()
    wall_ens = time.perf_counter() - t0
    cells_ens = ens.cell_steps()
    nt_ens = ens.tab["nt"][:count].copy()
    print("# ensemble: %d instances, %d blocks, %.3f s" % (count, blocks, wall_ens), file=sys.stderr, flush=True)
    wall_plain, cells_plain, checked, equal, steps_equal = 0.0, 0, 0, 0, 0
    for i, sp in enumerate(specs):
        bc = BoundaryCondition(BoundaryCondition.getTypes(_bc_code(sp["bc"])))
        sim = KP07_dimsplit(ctx, *sp["fields"], sp["nx"], sp["ny"], 1.0, 1.0, 9.81, boundary_conditions=bc)
        sim.synchronize()
        t0 = time.perf_counter()
        sim.simulate(sp["t_end"])
        sim.synchronize()
        wall_plain += time.perf_counter() - t0
        cells_plain += sp["nx"] * sp["ny"] * sim.simSteps()
        steps_equal += int(sim.simSteps() == nt_ens[i])
        if (i + 1) % 64 == 0:
            print("# plain loop: %d/%d instances, %.2f s" % (i + 1, count, wall_plain), file=sys.stderr, flush=True)
        if i % 32 == 0:
            checked += 1
            equal += int(all(np.array_equal(a, b) for a, b in zip(sim.download(), ens.download(i))))
        del sim
    ens.close()
    print("# plain loop: %.3f s" % wall_plain, file=sys.stderr, flush=True)
    return {"workload": "%d KP07_dimsplit dam breaks, side log-U[64,1024], t_end U(5,50), per-instance CFL dt, "
                        "reflective/outflow 50/50, fast arithmetic" % count,
            "instances": count, "schedule": {"seg_rows": sched["seg_rows"],
                                             "waves_per_block": sched["waves_per_block"]},
            "cell_steps": cells_ens, "wall_s": wall_ens,
            "cell_updates_per_s": cells_ens / wall_ens, "plain_loop_wall_s": wall_plain,
            "plain_loop_cell_updates_per_s": cells_plain / wall_plain, "speedup": wall_plain / wall_ens,
            "step_counts_equal": "%d/%d" % (steps_equal, count), "bitwise_equal_sampled": "%d/%d" % (equal, checked)}


def _serving_leg(ctx, count, rate, sequential):
    """Continuous-batching server on Poisson arrivals (60 % KP07_dimsplit, 20 %
    HLL2/WAF, 20 % EE2D, side log-uniform in [128, 2048], t_end U(2, 40),
    200 GB budget) against sequential plain solves of the same requests."""
    from finitevolumegpu_amd.Serving import Server, run_sequential
    from finitevolumegpu_amd.initial_conditions import serving_requests
    reqs = serving_requests(count, seed=0, rate=rate)

    def progress(tag):
        t0 = time.perf_counter()
        mark = [0]

        def report(done, total):
            if done * 10 // total > mark[0] or done == total:
                mark[0] = done * 10 // total
                print("# %s: %d/%d requests done, %.1f s" % (tag, done, total, time.perf_counter() - t0),
                      file=sys.stderr, flush=True)
        return report

    server = Server(ctx, budget_bytes=200e9, steps_per_block=8, keep_results=False)
    served = server.serve([dict(r) for r in reqs], realtime=True, progress=progress("served"))
    server.close()
    print("# served: %s" % json.dumps(served), file=sys.stderr, flush=True)
    k = count if sequential < 0 else min(count, sequential)
    _, seq = run_sequential(ctx, [dict(r) for r in reqs[:k]], realtime=True, keep_results=False,
                            progress=progress("sequential"))
    print("# sequential: %s" % json.dumps(seq), file=sys.stderr, flush=True)
    return {"workload": "%d requests, Poisson %.0f/s, 60%% KP07 / 20%% HLL2|WAF / 20%% EE2D, side log-U[128,2048], "
                        "t_end U(2,40), budget 200 GB, each scheme's default arithmetic (fast KP07_dimsplit and HLL2, "
                        "exact WAF and EE2D)" % (count, rate),
            "served": served, "sequential": dict(seq, requests_of_total="%d/%d" % (k, count)),
            "throughput_ratio": served["cell_updates_per_s"] / seq["cell_updates_per_s"]}


def _traffic():
    path = os.path.join(ROOT, "profiles", "kp07_dimsplit_4096_hbm_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("bytes_per_launch")


def main():
    args = _parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (--gpus %d requested)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    from finitevolumegpu_amd.HipContext import HipContext
    ctx = HipContext(device=0, autotuning=False)
    if args.copy_calibration:
        _copy_calibration(ctx)
        return
    kinds = [k.strip() for k in args.schemes.split(",") if k.strip()]
    if "kp07" not in kinds:
        kinds.insert(0, "kp07")
    results = {}
    for k in kinds:
        results[k] = _measure(ctx, k, args)
        print("# %s: %s" % (k, json.dumps(results[k])), file=sys.stderr, flush=True)
    main_res = results["kp07"]
    achieved = main_res["algorithmic_gbs"]
    cpu = None if args.no_cpu_baseline else _cpu_baseline()
    line = {
        "metric": "cell-updates/s + HBM GB/s (KP07 4096^2); ensemble p50/p99 vs plain loop",
        "value": main_res["cell_updates_per_s"],
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": main_res["steps"],
        "warmup": args.warmup,
        "ms_per_step": main_res["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded dam break + 16 Gaussian bumps)",
        "config": {"workload": "KP07_dimsplit shallow water %dx%d float32, reflective walls, RK2 dimensional "
                               "splitting, dt = 0.45 x initial CFL" % (main_res["n"], main_res["n"]),
                   "tile_rows": main_res["launch"][0], "waves_per_block": main_res["launch"][1]},
        "hbm_gbs": achieved,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": _traffic()},
        "cpu_baseline": cpu,
        "schemes": {k: {kk: v for kk, v in r.items() if kk != "launch"} for k, r in results.items()},
    }
    if args.ensemble > 0:
        line["ensemble"] = _ensemble_leg(ctx, args.ensemble)
    if args.serve > 0:
        line["serving"] = _serving_leg(ctx, args.serve, args.serve_rate, args.sequential)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
