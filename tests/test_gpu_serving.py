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


def test_served_exact_requests_equal_the_oracle(ctx):
    """One exact-arithmetic request of each scheme of the generated serving
    workload, served together, equals the float32 oracle's simulate() of the
    same request (the reference's dt policy), bitwise."""
    from finitevolumegpu_amd.Ensemble import _bc_code
    from finitevolumegpu_amd.Serving import Server, request_spec
    from finitevolumegpu_amd.initial_conditions import serving_requests
    from oracle import fv_oracle as O
    reqs = serving_requests(40, seed=11, rate=1000.0, side_min=48, side_max=96, t_min=0.5, t_max=2.0)
    picked, seen = [], set()
    for r in reqs:
        if r["scheme"] not in seen:
            seen.add(r["scheme"])
            picked.append(dict(r, arrival=0.0))
    assert seen == {"kp07_dimsplit", "hll2", "waf", "ee2d_kp07_dimsplit"}
    print("case served exact vs oracle", flush=True)
    server = Server(ctx, steps_per_block=6, math="exact", capacity=8)
    summary = server.serve([dict(r) for r in picked], realtime=False)
    assert summary["requests"] == len(picked) and summary["failed"] == 0
    by_id = {c.id: c for c in server.completed}
    for r in picked:
        spec = request_spec(r)
        gamma = spec.get("gamma") if spec.get("gamma") is not None else None
        ora = O.OracleSimulator(r["scheme"], spec["fields"], r["nx"], r["ny"], spec["dx"], spec["dy"], spec["g"],
                                gamma=gamma, bc=_bc_code(spec["bc"]))
        ora.simulate(spec["t_end"])
        got = by_id[r["id"]]
        assert got.steps == ora.nt, (r["scheme"], got.steps, ora.nt)
        for k, (a, b) in enumerate(zip(got.fields, ora.download())):
            assert np.array_equal(a, b), (r["scheme"], k, float(np.abs(a - b).max()))
    server.close()


def test_nan_request_is_reported_failed_and_isolated(ctx):
    """A request whose initial data hold a NaN is reported as failed; the
    requests served beside it equal their solo solves bitwise."""
    from finitevolumegpu_amd.Serving import Server, make_simulator
    reqs = _requests(count=8, seed=9)
    bad_ids = []
    for r in reqs:
        if r["id"] in (1, 4):  # an HLL2 and a KP07_dimsplit request
            f0 = np.array(r["fields"][0], dtype=np.float32)
            f0[r["ny"] // 3, r["nx"] // 2] = np.nan if r["id"] == 1 else np.inf
            r["fields"] = [f0] + list(r["fields"][1:])
            bad_ids.append(r["id"])
    print("case serving nan isolation", flush=True)
    server = Server(ctx, steps_per_block=4, capacity=16)
    summary = server.serve(reqs, realtime=False)
    assert summary["failed"] == len(bad_ids)
    assert summary["requests"] == len(reqs) - len(bad_ids)
    by_id = {c.id: c for c in server.completed}
    for r in reqs:
        got = by_id[r["id"]]
        if r["id"] in bad_ids:
            assert got.error is not None and got.error.startswith("non-finite"), got.error
            continue
        assert got.error is None
        sim = make_simulator(ctx, r)
        sim.simulate(r["t_end"])
        assert got.steps == sim.simSteps()
        for k, (a, b) in enumerate(zip(got.fields, sim.download())):
            assert np.array_equal(a, b), (r["id"], k)
    assert server.used_bytes == 0
    server.close()


def test_cancelled_requests_leave_the_others_unchanged(ctx):
    """A waiting request cancelled before admission is never run; a running
    one cancelled mid-flight frees its slot and memory; both are reported as
    cancelled, and the other requests equal their solo solves bitwise."""
    from finitevolumegpu_amd.Serving import Server, make_simulator
    reqs = _requests(count=6, seed=13)
    long_req = dict(reqs[0])
    long_req.update({"id": 50, "t_end": 1.0e4, "arrival": 0.0, "cancel_at": 0.05})  # would run ~10^5 steps
    early = dict(reqs[1])
    early.update({"id": 51, "arrival": 0.01, "cancel_at": 0.0})  # cancelled as soon as it arrives
    print("case serving cancellation", flush=True)
    server = Server(ctx, steps_per_block=4, capacity=16)
    summary = server.serve(reqs + [long_req, early], realtime=True)
    assert summary["cancelled"] == 2
    by_id = {c.id: c for c in server.completed}
    assert by_id[50].error == "cancelled" and by_id[50].admitted is not None and by_id[50].steps > 0
    assert by_id[51].error == "cancelled" and by_id[51].admitted is None and by_id[51].steps == 0
    for r in reqs:
        got = by_id[r["id"]]
        assert got.error is None
        sim = make_simulator(ctx, r)
        sim.simulate(r["t_end"])
        for k, (a, b) in enumerate(zip(got.fields, sim.download())):
            assert np.array_equal(a, b), (r["id"], k)
    assert server.used_bytes == 0 and not server.owner
    server.close()


def test_cancel_from_another_thread(ctx):
    """Server.cancel() from another thread stops a long request."""
    import threading
    from finitevolumegpu_amd.Serving import Server
    req = _requests(count=1, seed=17)[0]
    req.update({"t_end": 1.0e5})
    server = Server(ctx, steps_per_block=4, capacity=4)
    timer = threading.Timer(0.1, server.cancel, args=(req["id"],))
    timer.start()
    summary = server.serve([req], realtime=True, max_wall=30.0)
    timer.join()
    assert summary["cancelled"] == 1 and server.completed[0].error == "cancelled"
    assert summary["wall_s"] < 10.0
    server.close()
