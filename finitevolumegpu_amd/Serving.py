"""Continuous-batching simulation server (no counterpart in the reference).

Requests (a scheme, initial fields, grid spacing, boundary conditions and a
t_end) arrive over time.  The server keeps one ``Ensemble`` slot pool per
scheme on its own stream, all state resident in device memory, and loops:

1. admit every request that has arrived, while the device-memory budget
   allows (two padded buffers per field); the rest wait in arrival order;
2. enqueue one block of steps for every pool with work (pools run
   concurrently on their streams, requests that joined at different times
   share each launch);
3. retire finished instances: download the result, record the completion
   time, free the slot and its memory.

Because an instance's arithmetic and dt policy do not depend on what else is
in its pool, each result equals the plain Simulator.simulate() result of the
same request alone (bitwise, same arithmetic policy).

``run_sequential`` is the baseline: the same requests served one at a time
through the plain Simulator API, in arrival order.
"""

import collections
import concurrent.futures
import logging
import time

import numpy as np

from .Ensemble import DeviceArena, Ensemble, PinnedPool, _bc_code, instance_initial_cfl
from .HipContext import Stream
from .Simulator import BoundaryCondition


# error: None for a request that ran to t_end with finite fields; otherwise
# what went wrong ("non-finite ...": its state holds NaN or inf values when it
# retires; "cancelled")
Completed = collections.namedtuple("Completed", "id scheme arrival admitted finished steps cells fields error",
                                   defaults=(None,))


def request_spec(req, pinned=None):
    """The Ensemble instance dict of a request.  A request without ``fields``
    (``initial_conditions.serving_requests``) is materialised here, so the host
    never holds the initial data of requests still queued.  The initial CFL is
    computed here too, and with a ``PinnedPool`` the fields are staged in
    page-locked memory (``_staging`` lists the buffers to return once the
    upload has completed), so that admitting the request costs the admitting
    thread only queued copies."""
    if "fields" in req and pinned is None:
        return req
    spec = dict(req)
    if "fields" not in req:
        from .initial_conditions import materialise
        spec.update(materialise(req))
        spec.setdefault("nx", req["n"])
        spec.setdefault("ny", req["n"])
    nvars = len(spec["fields"])
    gamma = spec.get("gamma")
    gamma = (1.4 if nvars == 4 else 0.0) if gamma is None else gamma
    spec["cfl0"] = instance_initial_cfl(nvars, spec["fields"], spec["dx"], spec["dy"], spec["g"], np.float32(gamma))
    if pinned is not None:
        staged, bufs = [], []
        for f in spec["fields"]:
            f = np.asarray(f, dtype=np.float32)
            buf = pinned.get(f.nbytes)
            dst = pinned.view(buf, f.shape)
            np.copyto(dst, f)
            staged.append(dst)
            bufs.append(buf)
        spec["fields"] = staged
        spec["_staging"] = bufs
    return spec


class _Prefetcher:
    """Materialises requests on worker threads (NumPy releases the GIL), at
    most ``depth`` ahead of the one being admitted."""

    def __init__(self, workers=8, depth=32, pinned=None):
        self.pool = concurrent.futures.ThreadPoolExecutor(max_workers=workers) if workers else None
        self.depth = depth
        self.pinned = pinned
        self.futures = {}

    def ahead(self, reqs):
        """Starts materialising the first ``depth`` of ``reqs`` (arrived, in order)."""
        if self.pool is None:
            return
        for k, req in enumerate(reqs):
            if k >= self.depth:
                break
            if req["id"] not in self.futures:
                self.futures[req["id"]] = self.pool.submit(request_spec, req, self.pinned)

    def ready(self, req):
        f = self.futures.get(req["id"])
        return f is None or f.done()

    def take(self, req):
        f = self.futures.pop(req["id"], None)
        return request_spec(req, self.pinned) if f is None else f.result()

    def close(self):
        if self.pool is not None:
            self.pool.shutdown(wait=True, cancel_futures=True)
        if self.pinned is not None:  # staged requests never admitted
            for f in self.futures.values():
                # a cancelled future has no result (f.exception() would raise)
                if f.done() and not f.cancelled() and f.exception() is None:
                    for buf in f.result().get("_staging", ()):
                        self.pinned.put(buf)
        self.futures = {}


class Server:
    """Continuous-batching server: one Ensemble slot pool per scheme, all pools
    carving instance memory out of one DeviceArena, admission against
    ``budget_bytes`` of instance memory."""

    def __init__(self, context, budget_bytes=200e9, steps_per_block=8, math=None, capacity=4096,
                 keep_results=True, use_graph=True, prefetch_workers=8, chunk_bytes=1 << 30):
        self.logger = logging.getLogger(__name__)
        self.context = context
        self.budget = float(budget_bytes)
        self.steps_per_block = int(steps_per_block)
        self.math = math
        self.capacity = int(capacity)
        self.keep_results = keep_results
        self.use_graph = use_graph
        self.prefetch_workers = int(prefetch_workers)
        self.arena = DeviceArena(chunk_bytes=int(min(chunk_bytes, max(self.budget, 1 << 20))))
        self.pinned = PinnedPool()
        self.staged = collections.defaultdict(list)  # scheme -> page-locked buffers of queued uploads
        self.pools = {}
        self.used_bytes = 0
        self.owner = {}  # (scheme, slot) -> request record
        self.completed = []
        self.rejected = []
        self.cancel_requested = set()  # ids; see cancel()
        self.timing = collections.Counter()

    def _pool(self, scheme):
        if scheme not in self.pools:
            self.pools[scheme] = Ensemble(self.context, scheme, capacity=self.capacity, math=self.math,
                                          steps_per_graph=self.steps_per_block, use_graph=self.use_graph,
                                          stream=Stream(), arena=self.arena)
        return self.pools[scheme]

    def fits(self, req):
        pool = self._pool(req["scheme"])
        need = pool.bytes_of(req)
        if need > self.budget:
            return None  # can never fit
        return self.used_bytes + need <= self.budget and pool.free_slot() is not None

    def admit(self, req, now, spec=None):
        pool = self._pool(req["scheme"])
        spec = spec if spec is not None else request_spec(req)
        slot = pool.add(spec)
        self.staged[req["scheme"]].extend(spec.get("_staging", ()))
        self.used_bytes += pool.bytes_of(req)
        self.owner[(req["scheme"], slot)] = {"req": req, "admitted": now}
        return slot

    def cancel(self, request_id):
        """Asks the server to drop a request: a waiting one is never admitted, a
        running one leaves its pool at the next iteration (its slot and memory
        are freed; the other instances of the pool are not affected).  Either
        way it is reported with error "cancelled".  May be called from another
        thread while serve() runs; a request may also carry ``cancel_at``
        (seconds from the start of serve())."""
        self.cancel_requested.add(request_id)

    def _cancel(self, waiting, now_fn):
        """Drops the requests asked for by cancel() or whose ``cancel_at`` has passed."""
        now = now_fn()
        ids = set(self.cancel_requested)
        for rec in self.owner.values():
            if rec["req"].get("cancel_at") is not None and rec["req"]["cancel_at"] <= now:
                ids.add(rec["req"]["id"])
        for req in waiting:
            if req.get("cancel_at") is not None and req["cancel_at"] <= now:
                ids.add(req["id"])
        if not ids:
            return 0
        dropped = 0
        for req in [r for r in waiting if r["id"] in ids]:
            waiting.remove(req)
            self.completed.append(Completed(req["id"], req["scheme"], req["arrival_time"], None, now, 0, 0, None,
                                            "cancelled"))
            self.cancel_requested.discard(req["id"])
            dropped += 1
        for (scheme, slot), rec in list(self.owner.items()):
            req = rec["req"]
            if req["id"] not in ids:
                continue
            pool = self.pools[scheme]
            del self.owner[(scheme, slot)]
            nt = int(pool.state(slot)[1])
            self.used_bytes -= pool.bytes_of(req)
            pool.remove(slot)  # waits for the pool's stream before the memory is reused
            self.completed.append(Completed(req["id"], scheme, req["arrival_time"], rec["admitted"], now, nt,
                                            int(req["nx"]) * int(req["ny"]) * nt, None, "cancelled"))
            self.cancel_requested.discard(req["id"])
            dropped += 1
        return dropped

    def _retire(self, now_fn):
        done = 0
        for scheme, pool in self.pools.items():
            if not pool.live.any():
                continue
            slots = pool.finished()  # downloads the pool's table once
            # failure isolation: an instance whose state went non-finite (NaN
            # initial data, a blown-up solution) is reported as failed; the
            # others of the pool never read its memory
            bad = pool.nonfinite_counts(slots)
            for slot, nbad in zip(slots, bad):
                rec = self.owner.pop((scheme, slot))
                req = rec["req"]
                fields = pool.download(slot) if self.keep_results else None
                nt = int(pool.tab["nt"][slot])
                now = now_fn()
                error = None if nbad == 0 else "non-finite values in %d cells of the final state" % int(nbad)
                self.completed.append(Completed(req["id"], scheme, req["arrival_time"], rec["admitted"], now, nt,
                                                int(req["nx"]) * int(req["ny"]) * nt, fields, error))
                self.used_bytes -= pool.bytes_of(req)
                pool.remove(slot)
                done += 1
        return done

    def _admit_waiting(self, waiting, prefetch, now):
        """Admits waiting requests in arrival order while the budget and the
        pools' slots allow; stops at a request whose data are not ready yet if
        the device has other work."""
        while waiting:
            ok = self.fits(waiting[0])
            if ok is None:
                self.rejected.append(waiting.popleft()["id"])
                continue
            if not ok or (self.owner and not prefetch.ready(waiting[0])):
                break
            req = waiting.popleft()
            c = time.perf_counter()
            spec = prefetch.take(req)
            c1 = time.perf_counter()
            self.admit(req, now(), spec)
            self.timing["admit_data_wait_s"] += c1 - c
            self.timing["admit_queue_s"] += time.perf_counter() - c1
            prefetch.ahead(waiting)

    def serve(self, requests, realtime=True, max_wall=None, progress=None):
        """Serves ``requests`` (dicts as Ensemble instances plus ``id``, ``scheme``,
        ``arrival`` seconds from the start; without ``fields`` they are
        materialised at admission).  With ``realtime`` the arrivals are
        replayed against the wall clock; otherwise all requests are available
        at once.  Each iteration enqueues one block of steps on every pool,
        admits arrivals while the device runs it (their copies queue behind
        the block), waits, and retires finished requests.  ``progress(done,
        total)`` is called after each retirement pass.  Returns a summary dict."""
        pending = collections.deque(sorted(requests, key=lambda r: r["arrival"]))
        waiting = collections.deque()
        prefetch = _Prefetcher(self.prefetch_workers, pinned=self.pinned)
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0  # noqa: E731
        clock = time.perf_counter
        for r in requests:
            r["arrival_time"] = r["arrival"] if realtime else 0.0
        try:
            while pending or waiting or self.owner:
                while pending and (not realtime or pending[0]["arrival"] <= now()):
                    waiting.append(pending.popleft())
                if self.cancel_requested or any("cancel_at" in r for r in requests):
                    self._cancel(waiting, now)
                    if not (pending or waiting or self.owner):
                        break
                prefetch.ahead(waiting)
                c = clock()
                busy = [pool for pool in self.pools.values() if pool.live.any()]
                for pool in busy:  # one block of steps per pool, concurrently on their streams
                    pool.advance(wait=False)
                self.timing["enqueue_s"] += clock() - c
                c = clock()
                self._admit_waiting(waiting, prefetch, now)  # overlaps the blocks just enqueued
                self.timing["admit_s"] += clock() - c
                if not busy:
                    if not self.owner and pending and realtime:
                        time.sleep(max(0.0, min(0.002, pending[0]["arrival"] - now())))
                    continue
                c = clock()
                for scheme, pool in self.pools.items():
                    if pool in busy or self.staged[scheme]:
                        pool.stream.synchronize()
                    for buf in self.staged.pop(scheme, ()):  # their uploads are done
                        self.pinned.put(buf)
                self.timing["wait_s"] += clock() - c
                c = clock()
                if self._retire(now) and progress is not None:
                    progress(len(self.completed), len(requests))
                self.timing["retire_s"] += clock() - c
                self.timing["iterations"] += 1
                if max_wall is not None and now() > max_wall:
                    break
        finally:
            prefetch.close()
        wall = now()
        return self.summary(wall)

    def summary(self, wall):
        ok = [c for c in self.completed if c.error is None]
        lat = np.array([c.finished - c.arrival for c in ok])
        cells = sum(c.cells for c in self.completed)
        out = {"requests": len(ok), "rejected": len(self.rejected),
               "failed": sum(1 for c in self.completed if c.error is not None and c.error != "cancelled"),
               "cancelled": sum(1 for c in self.completed if c.error == "cancelled"), "wall_s": wall,
               "cell_updates_per_s": cells / wall if wall > 0 else 0.0,
               "p50_latency_s": float(np.percentile(lat, 50)) if lat.size else None,
               "p99_latency_s": float(np.percentile(lat, 99)) if lat.size else None,
               "cell_steps": cells}
        out["host_time"] = {k: (round(v, 4) if isinstance(v, float) else v) for k, v in self.timing.items()}
        out["host_time"]["graph_captures"] = sum(p.captures for p in self.pools.values())
        out["arena_reserved_bytes"] = self.arena.reserved()
        return out

    def close(self):
        for pool in self.pools.values():
            pool.close()
        self.pools = {}
        self.arena.close()
        self.pinned.close()


def make_simulator(context, req, math=None):
    """Plain Simulator for one request (the baseline path)."""
    from .EE2D_KP07_dimsplit import EE2D_KP07_dimsplit
    from .HLL2 import HLL2
    from .KP07_dimsplit import KP07_dimsplit
    from .WAF import WAF
    req = request_spec(req)
    bc = BoundaryCondition(BoundaryCondition.getTypes(_bc_code(req.get("bc", BoundaryCondition()))))
    f = req["fields"]
    scheme = req["scheme"]
    if scheme == "ee2d_kp07_dimsplit":
        return EE2D_KP07_dimsplit(context, *f, req["nx"], req["ny"], req["dx"], req["dy"], req["g"],
                                  req.get("gamma", 1.4), boundary_conditions=bc, math=math)
    cls = {"kp07_dimsplit": KP07_dimsplit, "hll2": HLL2, "waf": WAF}[scheme]
    return cls(context, *f, req["nx"], req["ny"], req["dx"], req["dy"], req["g"], boundary_conditions=bc, math=math)


def run_sequential(context, requests, realtime=True, math=None, keep_results=True, prefetch_workers=8,
                   progress=None):
    """Baseline: one plain simulate() per request, in arrival order (initial
    data materialised ahead on worker threads, as the server does)."""
    t0 = time.perf_counter()
    now = lambda: time.perf_counter() - t0  # noqa: E731
    out = []
    order = sorted(requests, key=lambda r: r["arrival"])
    prefetch = _Prefetcher(prefetch_workers)
    for k, req in enumerate(order):
        if realtime:
            while now() < req["arrival"]:
                time.sleep(min(0.002, req["arrival"] - now()))
        prefetch.ahead([r for r in order[k:k + prefetch.depth] if not realtime or r["arrival"] <= now()])
        spec = prefetch.take(req)
        sim = make_simulator(context, spec, math)
        sim.simulate(spec["t_end"])  # the materialised t_end (scaled for Euler requests)
        fields = sim.download() if keep_results else None
        sim.synchronize()
        arrival = req["arrival"] if realtime else 0.0
        out.append(Completed(req["id"], req["scheme"], arrival, arrival, now(), sim.simSteps(),
                             int(req["nx"]) * int(req["ny"]) * sim.simSteps(), fields))
        del sim
        if progress is not None:
            progress(len(out), len(order))
    prefetch.close()
    wall = now()
    lat = np.array([c.finished - c.arrival for c in out])
    cells = sum(c.cells for c in out)
    return out, {"requests": len(out), "wall_s": wall, "cell_updates_per_s": cells / wall,
                 "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
                 "cell_steps": cells}
