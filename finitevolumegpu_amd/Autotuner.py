"""Launch-shape autotuner (GPUSimulators/Autotuner.py:34-278, rewritten for gfx950).

The reference sweeps the CUDA block size (8..31 x 8..31) at 2048x2048 and picks
the fastest configuration; it calls ``sim.stepEuler``, which no simulator
defines, so every run records NaN.  This tuner keeps the interface
(``benchmark``, ``get_peak_performance`` returning the fastest configuration
and its Mcells/s, a per-host ``.npz`` cache, ``run_benchmark`` with warm-up and
event timing, ``gen_test_data`` and the ``sanity_check`` bounds of
Autotuner.py:207-213) and sweeps the two knobs of the gfx950 kernels: output
rows per 64-lane strip (``tile_rows``) and strips per workgroup
(``waves_per_block``).  Timing uses ``step()``.

``ensemble_schedule`` extends it to the batched engine: it times candidate
segment heights and workgroup sizes on the actual instance mix and returns
the fastest.  ``tile_rows`` 0 is the library's own plan (the round-weighted
split where it fits, else equal segments balanced over the device); the tuned
shape is applied at every grid size, so a tuned 0 keeps that choice per size.
``seg_rows`` None is the ensemble's balanced plan.
"""

import logging
import os
from socket import gethostname

import numpy as np

from . import Common
from .HipContext import Event, HipContext


class Autotuner:
    def __init__(self, nx=2048, ny=2048, tile_rows=(0, 16, 32, 48, 64, 96, 128), waves_per_block=(1, 2, 4),
                 filename=None, timesteps=10, warmup_timesteps=2):
        self.filename = filename or os.path.join(os.getcwd(), "autotuning_data_" + gethostname() + ".npz")
        self.nx = nx
        self.ny = ny
        self.tile_rows = tuple(int(v) for v in tile_rows)
        self.waves_per_block = tuple(int(v) for v in waves_per_block)
        self.timesteps = timesteps
        self.warmup_timesteps = warmup_timesteps
        self.performance = {}

    # -- cache -------------------------------------------------------------
    def _load(self):
        if not os.path.isfile(self.filename):
            return {"simulators": np.array([], dtype="<U64")}
        with np.load(self.filename, allow_pickle=False) as data:
            return {k: data[k] for k in data.files}

    def benchmark(self, simulator, force=False, context=None):
        """Times every (tile_rows, waves_per_block) pair for ``simulator`` and stores
        the Mcells/s table in the cache file."""
        logger = logging.getLogger(__name__)
        key = str(simulator.__name__)
        data = self._load()
        if not force and key in list(data["simulators"]):
            logger.info("%s already benchmarked - skipping", key)
            return
        if context is None:
            context = HipContext(autotuning=False)
        arguments = Autotuner.gen_test_data(simulator, self.nx, self.ny)
        arguments["context"] = context
        table = Autotuner.benchmark_single_simulator(simulator, arguments, self.tile_rows, self.waves_per_block,
                                                     self.timesteps, self.warmup_timesteps)
        data[key + "_megacells"] = table
        data[key + "_tile_rows"] = np.array(self.tile_rows)
        data[key + "_waves_per_block"] = np.array(self.waves_per_block)
        sims = [str(s) for s in data["simulators"]]
        if key not in sims:
            sims.append(key)
        data["simulators"] = np.array(sims, dtype="<U64")
        np.savez_compressed(self.filename, **data)

    def get_peak_performance(self, simulator):
        """Fastest configuration: {"tile_rows", "waves_per_block", "megacells",
        "block_width", "block_height"} (the last two echo the shape for callers
        written against the reference)."""
        key = simulator.__name__
        if key in self.performance:
            return self.performance[key]
        data = self._load()
        if key not in list(data["simulators"]):
            self.benchmark(simulator)
            data = self._load()
        megacells = data[key + "_megacells"]
        if not np.any(np.isfinite(megacells)):
            raise RuntimeError("autotuning of %s produced no valid configuration" % key)
        j, i = np.unravel_index(np.nanargmax(megacells), megacells.shape)
        tr = int(data[key + "_tile_rows"][j])
        wpb = int(data[key + "_waves_per_block"][i])
        self.performance[key] = {"tile_rows": tr, "waves_per_block": wpb, "megacells": float(megacells[j, i]),
                                 "block_width": 64 * wpb, "block_height": tr}
        return self.performance[key]

    @staticmethod
    def benchmark_single_simulator(simulator, arguments, tile_rows, waves_per_block, timesteps=10,
                                   warmup_timesteps=2):
        megacells = np.full((len(tile_rows), len(waves_per_block)), np.nan)
        for j, tr in enumerate(tile_rows):
            for i, wpb in enumerate(waves_per_block):
                args = dict(arguments)
                args.update({"tile_rows": tr, "waves_per_block": wpb})
                megacells[j, i] = Autotuner.run_benchmark(simulator, args, timesteps, warmup_timesteps)
        return megacells

    @staticmethod
    def run_benchmark(simulator, arguments, timesteps=10, warmup_timesteps=2):
        """Mcells/s of ``timesteps`` full steps after ``warmup_timesteps``; NaN if
        construction fails or the result leaves the sanity bounds."""
        logger = logging.getLogger(__name__)
        args = dict(arguments)
        dt = args.pop("dt")
        bounds = args.pop("bounds")
        try:
            sim = simulator(**args)
        except Exception as e:  # noqa: BLE001 - a failing configuration scores NaN
            logger.debug("Failed creating %s: %s", simulator.__name__, e)
            return np.nan
        start, end = Event(), Event()
        for _ in range(warmup_timesteps):
            sim.step(dt)
        start.record(sim.stream)
        for _ in range(timesteps):
            sim.step(dt)
        end.record(sim.stream)
        end.synchronize()
        gpu_elapsed = end.time_since(start) * 1.0e-3
        megacells = (sim.nx * sim.ny * timesteps / 1.0e6) / gpu_elapsed
        out = sim.download()
        sane = all(Autotuner.sanity_check(v, lo, hi) for v, (lo, hi) in zip(out, bounds))
        return float(megacells) if sane else np.nan

    @staticmethod
    def gen_test_data(simulator, nx, ny, g=9.81):
        """Smooth test state for ``simulator`` (SWE: the Gaussian hump of
        Autotuner.py:228-268; Euler: a smooth density bump at rest)."""
        width, height = 100.0, 100.0
        dx, dy = width / nx, height / ny
        x = (dx * (np.arange(nx) + 0.5) - width / 2) / (width / 5.0)
        y = (dy * (np.arange(ny) + 0.5) - height / 2) / (height / 5.0)
        r = np.minimum(1.0, np.hypot(x[None, :], y[:, None]))
        bump = 0.5 * (1.0 + np.cos(np.pi * r))
        if getattr(simulator, "nvars", 3) == 4:
            gamma = 1.4
            rho = (1.0 + 0.1 * bump).astype(np.float32)
            z = np.zeros((ny, nx), dtype=np.float32)
            E = np.full((ny, nx), 2.5 / (gamma - 1.0), dtype=np.float32)
            dt = 0.3 * min(dx, dy) / np.sqrt(gamma * 2.5 / 1.0)
            return {"rho": rho, "rho_u": z, "rho_v": z.copy(), "E": E, "nx": nx, "ny": ny, "dx": dx, "dy": dy,
                    "g": 0.0, "gamma": gamma, "dt": np.float32(dt),
                    "bounds": [(0.5, 2.0), (-1.0, 1.0), (-1.0, 1.0), (0.0, 100.0)]}
        h = (0.5 + 0.1 * bump).astype(np.float32)
        hu = (0.1 * bump).astype(np.float32)
        hv = hu.copy()
        dt = 0.7 * min(dx, dy) / (0.1 * np.sqrt(2.0) + np.sqrt(g * 0.6))
        return {"h0": h, "hu0": hu, "hv0": hv, "nx": nx, "ny": ny, "dx": dx, "dy": dy, "g": g,
                "dt": np.float32(dt), "bounds": [(0.3, 0.7), (-0.2, 0.2), (-0.2, 0.2)]}

    @staticmethod
    def ensemble_schedule(context, scheme, instances, seg_rows=(None, 24, 48, 96), waves_per_block=(2, 4),
                          steps_per_graph=8, blocks=3, math=None):
        """Schedule of the batched engine for ``instances`` (Ensemble dicts): for
        every (segment rows, waves per workgroup) pair, builds the ensemble,
        replays one graph block to warm up and times ``blocks`` more with events.
        ``None`` segment rows is the balanced plan (all strip rows split over 4
        waves per SIMD).  Returns {"seg_rows", "waves_per_block",
        "cell_steps_per_s", "table"} for the fastest pair; the instances passed
        in are not modified (each candidate runs on its own copy)."""
        from .Ensemble import Ensemble
        table = []
        best = None
        for sr in seg_rows:
            for wpb in waves_per_block:
                ens = Ensemble(context, scheme, instances, math=math, steps_per_graph=steps_per_graph,
                               waves_per_block=wpb, seg_rows=sr)
                try:
                    ens.advance()
                    c0 = ens.cell_steps()
                    start, end = Event(), Event()
                    start.record(ens.stream)
                    for _ in range(blocks):
                        ens.advance(wait=False)
                    end.record(ens.stream)
                    end.synchronize()
                    seconds = end.time_since(start) * 1.0e-3
                    rate = (ens.cell_steps() - c0) / seconds if seconds > 0 else 0.0
                finally:
                    ens.close()
                table.append({"seg_rows": sr, "waves_per_block": wpb, "cell_steps_per_s": rate})
                if best is None or rate > best["cell_steps_per_s"]:
                    best = dict(table[-1])
        best["table"] = table
        return best

    @staticmethod
    def sanity_check(variable, bound_min, bound_max):
        maxval = np.amax(variable)
        minval = np.amin(variable)
        return not (np.isnan(maxval) or np.isnan(minval) or maxval > bound_max or minval < bound_min)


def _main(argv=None):
    """Command line sweep: prints one JSON line per configuration and the best."""
    import argparse
    import json
    ap = argparse.ArgumentParser(description="Sweep the launch shape of one scheme")
    ap.add_argument("--scheme", default="kp07", choices=["kp07", "hll2", "waf", "ee2d"])
    ap.add_argument("--n", type=int, default=4096)
    ap.add_argument("--tile-rows", default="24,32,40,48,56,60,64,68,70,72,80,96,128")
    ap.add_argument("--waves-per-block", default="1,2,4")
    ap.add_argument("--timesteps", type=int, default=20)
    args = ap.parse_args(argv)
    from . import KP07_dimsplit, HLL2, WAF, EE2D_KP07_dimsplit
    cls = {"kp07": KP07_dimsplit.KP07_dimsplit, "hll2": HLL2.HLL2, "waf": WAF.WAF,
           "ee2d": EE2D_KP07_dimsplit.EE2D_KP07_dimsplit}[args.scheme]
    ctx = HipContext(autotuning=False)
    data = Autotuner.gen_test_data(cls, args.n, args.n)
    data["context"] = ctx
    trs = [int(v) for v in args.tile_rows.split(",")]
    wps = [int(v) for v in args.waves_per_block.split(",")]
    best = None
    for tr in trs:
        for wp in wps:
            d = dict(data)
            d.update({"tile_rows": tr, "waves_per_block": wp})
            mc = Autotuner.run_benchmark(cls, d, args.timesteps, 2)
            print(json.dumps({"scheme": args.scheme, "n": args.n, "tile_rows": tr, "waves_per_block": wp,
                              "megacells": mc}), flush=True)
            if np.isfinite(mc) and (best is None or mc > best[0]):
                best = (mc, tr, wp)
    if best is None:
        print(json.dumps({"best": None}), flush=True)
        return 1
    print(json.dumps({"best": {"megacells": best[0], "tile_rows": best[1], "waves_per_block": best[2]}}), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(_main())
