#!/usr/bin/env python3
"""Benchmark: KP07_dimsplit shallow water, 4096x4096 float32, on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--schemes kp07,hll2,waf,ee2d]
                    [--no-cpu-baseline]

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
* schemes           the same measurement for HLL2 and WAF at 4096^2 and
                    EE2D_KP07_dimsplit at 8192^2 (64 B per cell-step)
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

SWE_BYTES_PER_CELL_SUBSTEP = 24  # 3 float32 fields, read once + written once
EULER_BYTES_PER_CELL_SUBSTEP = 32
HBM_PEAK_GBS = 8000.0


def _parse():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[1])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=4096, help="grid side of the SWE workload")
    ap.add_argument("--schemes", default="kp07,hll2,waf,ee2d")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--tile-rows", type=int, default=0)
    ap.add_argument("--waves-per-block", type=int, default=0)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 256 MiB device copy kernel 5 times (HBM counter calibration)")
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


def _swe_sim(ctx, scheme, n, args):
    from finitevolumegpu_amd.initial_conditions import dam_break_bumps
    h, hu, hv = dam_break_bumps(n, seed=0)
    kw = {}
    if args.tile_rows:
        kw["tile_rows"] = args.tile_rows
    if args.waves_per_block:
        kw["waves_per_block"] = args.waves_per_block
    if scheme == "kp07":
        from finitevolumegpu_amd.KP07_dimsplit import KP07_dimsplit as C
    elif scheme == "hll2":
        from finitevolumegpu_amd.HLL2 import HLL2 as C
    else:
        from finitevolumegpu_amd.WAF import WAF as C
    sim = C(ctx, h, hu, hv, n, n, 1.0, 1.0, 9.81, **kw)
    dt = np.float32(sim.computeDt() * 0.9)
    return sim, dt


def _euler_sim(ctx, n):
    from finitevolumegpu_amd.EE2D_KP07_dimsplit import EE2D_KP07_dimsplit
    from finitevolumegpu_amd.Simulator import BoundaryCondition
    from finitevolumegpu_amd.initial_conditions import rayleigh_taylor
    rho, ru, rv, E, dx, dy = rayleigh_taylor(n, n, seed=0)
    T = BoundaryCondition.Type
    bc = BoundaryCondition({"north": T.Reflective, "south": T.Reflective, "east": T.Periodic, "west": T.Periodic})
    sim = EE2D_KP07_dimsplit(ctx, rho, ru, rv, E, n, n, dx, dy, 0.1, 1.4, boundary_conditions=bc)
    c = np.sqrt(1.4 * 2.6 / 1.0)
    dt = np.float32(0.45 * dx / c)
    return sim, dt


def _measure(ctx, kind, args):
    if kind == "ee2d":
        n = 8192
        sim, dt = _euler_sim(ctx, n)
        bpc = EULER_BYTES_PER_CELL_SUBSTEP
        steps = max(1, args.steps // 3)
    else:
        n = args.n
        sim, dt = _swe_sim(ctx, kind, n, args)
        bpc = SWE_BYTES_PER_CELL_SUBSTEP
        steps = args.steps
    ms, wall = _time_steps(sim, dt, steps, args.warmup, ctx)
    out = sim.download()
    finite = all(bool(np.isfinite(a).all()) for a in out)
    launches = 2 * steps
    avg_launch_ms = ms / launches
    cells = n * n
    res = {
        "n": n, "steps": steps, "ms_per_step": ms / steps, "cell_updates_per_s": cells * steps / (ms * 1e-3),
        "avg_launch_ms": avg_launch_ms, "algorithmic_gbs": bpc * cells / (avg_launch_ms * 1e-3) / 1e9,
        "wall_s": wall, "finite": finite, "launch": [sim.launch.tile_rows, sim.launch.waves_per_block],
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
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
