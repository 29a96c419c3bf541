"""Continuous-batching server on the GPU: requests of mixed scheme and size that
arrive while others run are admitted against a memory budget, and every result
equals the same request solved alone through the plain Simulator API
(bitwise, same arithmetic policy)."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _requests(count=10, seed=5):
    from finitevolumegpu_amd.initial_conditions import rayleigh_taylor
    from finitevolumegpu_amd.Simulator import BoundaryCondition
    rng = np.random.default_rng(seed)
    schemes = ["kp07_dimsplit", "hll2", "waf", "ee2d_kp07_dimsplit"]
    out = []
    for i in range(count):
        scheme = schemes[i % 4]
        n = int(rng.integers(40, 140))
        if scheme == "ee2d_kp07_dimsplit":
            rho, ru, rv, E, dx, dy = rayleigh_taylor(n, n, width=1.0, seed=i)
            T = BoundaryCondition.Type
            bc = BoundaryCondition({"north": T.Reflective, "south": T.Reflective, "east": T.Periodic,
                                    "west": T.Periodic})
            req = {"fields": [rho, ru, rv, E], "dx": dx, "dy": dy, "g": 0.1, "gamma": 1.4,
                   "t_end": float(rng.uniform(20, 60)) * dx, "bc": bc}
        else:
            y, x = np.mgrid[0:n, 0:n] + 0.5
            h = np.where(np.hypot(x - 0.5 * n, y - 0.45 * n) < 0.2 * n, rng.uniform(1.2, 2.5), 1.0)
            h = h.astype(np.float32)
            z = np.zeros_like(h)
            req = {"fields": [h, z, z.copy()], "dx": 1.0, "dy": 1.0, "g": 9.81,
                   "t_end": float(rng.uniform(1.0, 4.0)), "bc": ["reflective", "outflow"][i % 2]}
        req.update({"id": i, "scheme": scheme, "nx": n, "ny": n, "arrival": 0.004 * i})
        out.append(req)
    return out


@pytest.mark.parametrize("math", ["fast", "exact"])
def test_served_results_equal_plain_solves(ctx, math):
    from finitevolumegpu_amd.Serving import Server, make_simulator
    reqs = _requests()
    # a budget that holds only a few requests at a time forces queueing
    budget = 2.5e6
    server = Server(ctx, budget_bytes=budget, steps_per_block=6, math=math, capacity=64)
    peak = {"bytes": 0, "live": 0}
    orig = server.admit

    def admit(req, now, spec=None):
        slot = orig(req, now, spec)
        peak["bytes"] = max(peak["bytes"], server.used_bytes)
        peak["live"] = max(peak["live"], len(server.owner))
        return slot
    server.admit = admit
    print("case serving", math, flush=True)
    summary = server.serve(reqs, realtime=True)
    assert summary["requests"] == len(reqs) and summary["rejected"] == 0
    assert peak["bytes"] <= budget
    assert peak["live"] > 1  # requests overlapped on the device
    by_id = {c.id: c for c in server.completed}
    for req in reqs:
        sim = make_simulator(ctx, req, math)
        sim.simulate(req["t_end"])
        ref = sim.download()
        got = by_id[req["id"]]
        assert got.steps == sim.simSteps()
        for k, (a, b) in enumerate(zip(got.fields, ref)):
            assert np.array_equal(a, b), (req["id"], req["scheme"], k, float(np.abs(a - b).max()))
    server.close()


def test_budget_admission_and_rejection(ctx):
    from finitevolumegpu_amd.Serving import Server
    reqs = _requests(count=6, seed=7)
    huge = dict(reqs[0])
    huge.update({"id": 99, "nx": 600, "ny": 600, "arrival": 0.0,
                 "fields": [np.ones((600, 600), np.float32), np.zeros((600, 600), np.float32),
                            np.zeros((600, 600), np.float32)]})
    budget = 1.2e6  # every small request fits alone; the 600x600 one never does
    server = Server(ctx, budget_bytes=budget, steps_per_block=4, capacity=16)
    peak = [0]
    orig = server.admit

    def admit(req, now, spec=None):
        slot = orig(req, now, spec)
        peak[0] = max(peak[0], server.used_bytes)
        return slot
    server.admit = admit
    summary = server.serve(reqs + [huge], realtime=False)
    assert summary["rejected"] == 1 and 99 in server.rejected
    assert summary["requests"] == len(reqs)
    assert peak[0] <= budget
    assert server.used_bytes == 0
    server.close()


def test_generated_requests_served_equal_sequential(ctx):
    """The generated serving workload (lazily materialised requests of all four
    schemes, Poisson arrivals) served concurrently equals the sequential plain
    solves of the same requests, bitwise."""
    from finitevolumegpu_amd.Serving import Server, run_sequential
    from finitevolumegpu_amd.initial_conditions import serving_requests
    reqs = serving_requests(16, seed=3, rate=400.0, side_min=64, side_max=200, t_min=1.0, t_max=4.0)
    assert {r["scheme"] for r in reqs} >= {"kp07_dimsplit", "ee2d_kp07_dimsplit"}
    print("case generated serving", flush=True)
    server = Server(ctx, steps_per_block=8, capacity=32)
    summary = server.serve([dict(r) for r in reqs], realtime=True)
    assert summary["requests"] == len(reqs) and summary["p99_latency_s"] >= summary["p50_latency_s"]
    seq, seq_summary = run_sequential(ctx, [dict(r) for r in reqs], realtime=False)
    by_id = {c.id: c for c in server.completed}
    for c in seq:
        got = by_id[c.id]
        assert got.steps == c.steps and got.cells == c.cells
        for k, (a, b) in enumerate(zip(got.fields, c.fields)):
            assert np.array_equal(a, b), (c.id, c.scheme, k)
    assert summary["cell_steps"] == seq_summary["cell_steps"]
    server.close()
