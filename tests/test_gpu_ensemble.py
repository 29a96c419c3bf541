"""Batched ensemble on the GPU: every instance equals the plain
Simulator.simulate(t_end) run of the same grid alone (bitwise, both arithmetic
policies), graph replay equals direct launches, and the exact-arithmetic
ensemble reproduces the float32 oracle's simulate() for a grid that crosses the
100-step dt refresh."""

import numpy as np
import pytest

from oracle import fv_oracle as O

pytestmark = pytest.mark.gpu


def _instances(seed=0, count=9):
    rng = np.random.default_rng(seed)
    sides = [(40, 37), (64, 64), (130, 70), (61, 200), (33, 33), (150, 121), (97, 64), (250, 48), (70, 70)]
    bcs = ["reflective", "outflow", "periodic"]
    out = []
    for i in range(count):
        nx, ny = sides[i % len(sides)]
        y, x = np.mgrid[0:ny, 0:nx] + 0.5
        r = np.hypot(x - rng.uniform(0.3, 0.7) * nx, y - rng.uniform(0.3, 0.7) * ny)
        h = np.where(r < rng.uniform(0.1, 0.3) * min(nx, ny), rng.uniform(1.1, 3.0), 1.0).astype(np.float32)
        z = np.zeros_like(h)
        out.append({"fields": [h, z, z.copy()], "nx": nx, "ny": ny, "dx": 1.0, "dy": 1.0, "g": 9.81,
                    "t_end": float(rng.uniform(2.0, 9.0)), "bc": bcs[i % 3]})
    return out


def _plain(ctx, spec, math, scheme="kp07_dimsplit"):
    from finitevolumegpu_amd.Ensemble import _bc_code
    from finitevolumegpu_amd.Simulator import BoundaryCondition
    from finitevolumegpu_amd.KP07_dimsplit import KP07_dimsplit
    from finitevolumegpu_amd.HLL2 import HLL2
    from finitevolumegpu_amd.WAF import WAF
    cls = {"kp07_dimsplit": KP07_dimsplit, "hll2": HLL2, "waf": WAF}[scheme]
    bc = BoundaryCondition(BoundaryCondition.getTypes(_bc_code(spec["bc"])))
    sim = cls(ctx, *spec["fields"], spec["nx"], spec["ny"], spec["dx"], spec["dy"], spec["g"],
              boundary_conditions=bc, math=math)
    sim.simulate(spec["t_end"])
    return sim.download(), sim.simTime(), sim.simSteps()


@pytest.mark.parametrize("math", ["exact", "fast"])
@pytest.mark.parametrize("scheme", ["kp07_dimsplit", "waf"])
def test_ensemble_equals_plain_simulate(ctx, math, scheme):
    from finitevolumegpu_amd.Ensemble import Ensemble
    specs = _instances(1)
    print("case ensemble", scheme, math, flush=True)
    ens = Ensemble(ctx, scheme, specs, math=math, steps_per_graph=7)
    ens.run()
    for i, spec in enumerate(specs):
        ref, t_ref, nt_ref = _plain(ctx, spec, math, scheme)
        got = ens.download(i)
        t, nt = ens.state(i)
        assert nt == nt_ref and t == t_ref, (i, nt, nt_ref, t, t_ref)
        for k in range(3):
            assert np.array_equal(got[k], ref[k]), (i, k, float(np.abs(got[k] - ref[k]).max()))
    ens.close()


def test_ensemble_graph_equals_direct_launches(ctx):
    from finitevolumegpu_amd.Ensemble import Ensemble
    specs = _instances(2, count=6)
    a = Ensemble(ctx, "kp07_dimsplit", specs, steps_per_graph=5, use_graph=True)
    b = Ensemble(ctx, "kp07_dimsplit", specs, steps_per_graph=5, use_graph=False)
    a.run()
    b.run()
    for i in range(len(specs)):
        for x, y in zip(a.download(i), b.download(i)):
            assert np.array_equal(x, y)
        assert a.state(i) == b.state(i)


def test_ensemble_matches_oracle_simulate(ctx):
    """Exact arithmetic: one grid that runs past the 100-step refresh equals the
    float32 oracle's simulate() bitwise, inside a mix of other grids."""
    from finitevolumegpu_amd.Ensemble import Ensemble
    specs = _instances(3, count=5)
    n = 48
    y, x = np.mgrid[0:n, 0:n] + 0.5
    h = (1.0 + 0.5 * (np.hypot(x - 20, y - 26) < 9)).astype(np.float32)
    z = np.zeros_like(h)
    target = {"fields": [h, z, z.copy()], "nx": n, "ny": n, "dx": 1.0, "dy": 1.0, "g": 9.81, "t_end": 12.0,
              "bc": "reflective"}
    specs.insert(2, target)
    ens = Ensemble(ctx, "kp07_dimsplit", specs, math="exact", steps_per_graph=16)
    ens.run()
    ora = O.OracleSimulator("kp07_dimsplit", target["fields"], n, n, 1.0, 1.0, 9.81)
    ora.simulate(12.0)
    t, nt = ens.state(2)
    assert nt == ora.nt and nt > 100
    assert t == ora.t
    for k, (a, b) in enumerate(zip(ens.download(2), ora.download())):
        assert np.array_equal(a, b), (k, float(np.abs(a - b).max()))


def test_autotuned_schedule_keeps_results(ctx):
    """The autotuner's ensemble schedule (segment rows x waves per workgroup)
    times every candidate, leaves its input untouched, and an ensemble run
    with the chosen schedule still equals the plain simulate() bitwise."""
    from finitevolumegpu_amd.Autotuner import Autotuner
    from finitevolumegpu_amd.Ensemble import Ensemble
    specs = _instances(4, count=6)
    before = [f.copy() for f in specs[0]["fields"]]
    print("case ensemble schedule", flush=True)
    sched = Autotuner.ensemble_schedule(ctx, "kp07_dimsplit", specs, seg_rows=(None, 16), waves_per_block=(2, 4),
                                        steps_per_graph=4, blocks=2)
    assert len(sched["table"]) == 4 and all(row["cell_steps_per_s"] > 0 for row in sched["table"])
    assert sched["cell_steps_per_s"] == max(row["cell_steps_per_s"] for row in sched["table"])
    assert all(np.array_equal(a, b) for a, b in zip(before, specs[0]["fields"]))
    ens = Ensemble(ctx, "kp07_dimsplit", specs, seg_rows=sched["seg_rows"], waves_per_block=sched["waves_per_block"])
    ens.run()
    for i, spec in enumerate(specs):
        ref, _, nt_ref = _plain(ctx, spec, "fast")
        assert ens.state(i)[1] == nt_ref
        for k in range(3):
            assert np.array_equal(ens.download(i)[k], ref[k]), (i, k)
    ens.close()


def test_generated_ensemble_workload(ctx):
    """The batched-ensemble workload's own generator (64 dam breaks, side
    log-uniform in [64, 1024], reflective or outflow walls, per-instance CFL dt),
    with t_end shortened to U(1, 4): in the default (fast) arithmetic every
    instance takes the plain simulate() loop's step count and every 8th equals
    its plain solve bitwise; one instance in exact arithmetic equals the float32
    oracle's simulate() bitwise."""
    from finitevolumegpu_amd.Autotuner import Autotuner
    from finitevolumegpu_amd.Ensemble import Ensemble, _bc_code
    from finitevolumegpu_amd.initial_conditions import ensemble_instances
    insts = ensemble_instances(64, seed=3, t_min=1.0, t_max=4.0)
    sides = [d["nx"] for d in insts]
    assert min(sides) < 100 and max(sides) > 600
    specs = [{"fields": [d["h"], d["hu"], d["hv"]], "nx": d["nx"], "ny": d["ny"], "dx": 1.0, "dy": 1.0,
              "g": 9.81, "t_end": d["t_end"], "bc": d["bc"]} for d in insts]
    print("case generated ensemble", flush=True)
    sched = Autotuner.ensemble_schedule(ctx, "kp07_dimsplit", specs[::8], seg_rows=(16, 32), waves_per_block=(2, 4),
                                        blocks=1)
    ens = Ensemble(ctx, "kp07_dimsplit", specs, steps_per_graph=8, seg_rows=sched["seg_rows"],
                   waves_per_block=sched["waves_per_block"])
    ens.run()
    ens.cell_steps()  # downloads the table
    nt = ens.tab["nt"][:len(specs)].copy()
    for i, spec in enumerate(specs):
        if i % 8 == 0:
            ref, t_ref, nt_ref = _plain(ctx, spec, None)
            assert nt[i] == nt_ref
            for k, (a, b) in enumerate(zip(ens.download(i), ref)):
                assert np.array_equal(a, b), (i, k)
        else:
            from finitevolumegpu_amd.KP07_dimsplit import KP07_dimsplit
            from finitevolumegpu_amd.Simulator import BoundaryCondition
            bc = BoundaryCondition(BoundaryCondition.getTypes(_bc_code(spec["bc"])))
            sim = KP07_dimsplit(ctx, *spec["fields"], spec["nx"], spec["ny"], 1.0, 1.0, 9.81, boundary_conditions=bc)
            sim.simulate(spec["t_end"])
            assert nt[i] == sim.simSteps(), i
    ens.close()
    # one instance (the smallest) in exact arithmetic against the oracle
    j = int(np.argmin(sides))
    spec = specs[j]
    ens = Ensemble(ctx, "kp07_dimsplit", [spec], math="exact", steps_per_graph=8)
    ens.run()
    ens.cell_steps()  # downloads the table
    ora = O.OracleSimulator("kp07_dimsplit", spec["fields"], spec["nx"], spec["ny"], 1.0, 1.0, 9.81,
                            bc=_bc_code(spec["bc"]))
    ora.simulate(spec["t_end"])
    assert int(ens.tab["nt"][0]) == ora.nt
    for k, (a, b) in enumerate(zip(ens.download(0), ora.download())):
        assert np.array_equal(a, b), k
    ens.close()
