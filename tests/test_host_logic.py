"""Host-side logic of the Python layer (CPU only)."""

import numpy as np
import pytest

from finitevolumegpu_amd.Simulator import BaseSimulator, BoundaryCondition


def test_boundary_condition_packing():
    T = BoundaryCondition.Type
    bc = BoundaryCondition({"north": T.Reflective, "south": T.Periodic, "east": T.Dirichlet, "west": T.Outflow})
    code = int(bc.asCodedInt())
    assert code == (3 << 24) | (2 << 16) | (0 << 8) | 4
    back = BoundaryCondition(BoundaryCondition.getTypes(code))
    assert (back.north, back.south, back.east, back.west) == (T.Reflective, T.Periodic, T.Dirichlet, T.Outflow)
    assert int(BoundaryCondition().asCodedInt()) == 0x03030303


def test_neumann_rejected():
    T = BoundaryCondition.Type
    with pytest.raises(NotImplementedError):
        BoundaryCondition({"north": T.Neumann, "south": T.Reflective, "east": T.Reflective, "west": T.Reflective})


class _FakeSim(BaseSimulator):
    """BaseSimulator without a device: substeps only record their dt."""

    def __init__(self, cfl):
        self.t = 0.0
        self.nt = 0
        self.dt = None
        self.cfl_scale = 0.9
        self.num_substeps = 2
        self.cfl = cfl
        self.substeps = []
        self.dt_calls = []
        import logging
        self.logger = logging.getLogger("fake")

    def substep(self, dt, step_number):
        self.substeps.append((float(dt), step_number))

    def computeDt(self):
        self.dt_calls.append(self.nt)
        return float(np.float32(self.cfl)) * 0.5

    def check(self):
        pass


def test_simulate_dt_policy():
    sim = _FakeSim(0.1)
    t_end = 0.045 * 150.25
    sim.simulate(t_end)
    assert sim.dt_calls[:2] == [0, 100]
    assert sim.nt in (151, 152)
    assert abs(sim.t - t_end) < 1e-6
    # every step passes a float32 dt to both substeps, numbered 0 and 1
    assert [s for _, s in sim.substeps[:4]] == [0, 1, 0, 1]
    assert all(np.float32(d) == d for d, _ in sim.substeps)
    assert isinstance(sim.t, float)


def test_simulate_fixed_dt():
    sim = _FakeSim(0.1)
    sim.simulate(1.0, dt=0.3)
    assert sim.dt_calls == []
    assert sim.nt == 4
    assert abs(sim.substeps[-1][0] - np.float32(0.1)) < 1e-7


def test_ensemble_plan_segments_cover_each_row_once():
    """The vectorised work list equals the plain nested loop (instance, strip,
    row runs of `seg`, then a stable longest-first sort) and covers every row of
    every strip (60 or 124 columns) exactly once."""
    from finitevolumegpu_amd.Ensemble import plan_segments
    rng = np.random.default_rng(4)
    for seg, width in ((1, 60), (7, 124), (24, 60), (100, 124)):
        ids = np.sort(rng.choice(50, size=13, replace=False))
        nx = rng.integers(2, 300, size=ids.size)
        ny = rng.integers(2, 300, size=ids.size)
        got = plan_segments(ids, nx, ny, seg, width)
        want = []
        for i, a, b in zip(ids, nx, ny):
            for s in range(-(-int(a) // width)):
                for r0 in range(0, int(b), seg):
                    want.append((int(i), s, r0, min(seg, int(b) - r0)))
        want.sort(key=lambda t: -t[3])
        assert got.dtype == np.int32 and got.shape == (len(want), 4)
        assert [tuple(r) for r in got.tolist()] == want
        for i, a, b in zip(ids, nx, ny):
            for s in range(-(-int(a) // width)):
                rows = np.zeros(int(b), dtype=int)
                for _, _, r0, n in got[(got[:, 0] == i) & (got[:, 1] == s)]:
                    rows[r0:r0 + n] += 1
                assert (rows == 1).all()
    assert plan_segments([], [], [], 24).shape == (0, 4)


def test_instance_dtype_matches_ctypes_layout():
    from finitevolumegpu_amd import _lib
    t = (_lib.Instance * 3)()
    view = np.frombuffer(t, dtype=_lib.INSTANCE_DTYPE)
    t[1].nx, t[2].t, t[0].cfl[1], t[2].cfl_scale = 77, 5.5, 2.25, 0.9
    t[1].field[1][3] = 0x1234
    assert view["nx"].tolist() == [0, 77, 0] and view["t"][2] == 5.5 and view["cfl"][0, 1] == 2.25
    assert view["cfl_scale"][2] == 0.9 and view["field"][1, 1, 3] == 0x1234
    view["active"][2] = 1
    assert t[2].active == 1


def test_device_arena_first_fit_and_coalescing(monkeypatch):
    """DeviceArena bookkeeping with the device allocation faked: first fit,
    256-byte rounding, oversize chunks, and merging of released neighbours."""
    from finitevolumegpu_amd import Ensemble as E
    base = [1 << 40]

    def fake_call(name, *args):
        assert name == "fvmi_malloc"
        ptr, size = args
        ptr._obj.value = base[0]
        base[0] += size + (1 << 20)
        return 0
    monkeypatch.setattr(E._lib, "call", fake_call)
    ar = E.DeviceArena(chunk_bytes=4096)
    a = ar.alloc(100)            # -> 256 bytes at offset 0 of chunk 0
    b = ar.alloc(1000)           # -> 1024 at 256
    c = ar.alloc(2816)           # -> fills chunk 0 exactly (256 + 1024 + 2816 = 4096)
    assert [h[:2] for h in (a, b, c)] == [(0, 0), (0, 256), (0, 1280)] and a[2] == 256 and b[2] == 1024
    assert ar.free[0] == [] and ar.used == 4096
    d = ar.alloc(10000)          # larger than a chunk: a chunk of its own size
    assert d[:2] == (1, 0) and ar.chunks[1][1] == 10240 and ar.free[1] == []
    ar.release(b)
    ar.release(a)                # merges with b's block
    assert ar.free[0] == [[0, 1280]]
    e = ar.alloc(512)            # first fit inside the merged hole
    assert e[:2] == (0, 0) and ar.free[0] == [[512, 768]]
    ar.release(c)
    ar.release(e)                # everything merges back into one block
    assert ar.free[0] == [[0, 4096]] and ar.used == 10240
    assert ar.address(d) == ar.chunks[1][0].value
    assert ar.reserved() == 4096 + 10240


def test_disk_generator_matches_direct_mask():
    """The row-run disk of the ensemble/serving workloads equals the direct
    cell-centre test (x-cx)^2 < r^2 - (y-cy)^2, including disks clipped by the
    domain and empty ones."""
    from finitevolumegpu_amd.initial_conditions import _centres, _disk
    rng = np.random.default_rng(0)
    for _ in range(200):
        n = int(rng.integers(2, 160))
        cx, cy = rng.uniform(-5, n + 5, size=2)
        r = rng.uniform(0, n)
        x = _centres(n)
        want = np.where((x[None, :] - cx) ** 2 < r * r - (x[:, None] - cy) ** 2, 2.5, 1.0).astype(np.float32)
        got = _disk(n, cx, cy, r, 2.5)
        assert got.dtype == np.float32 and np.array_equal(got, want), (n, cx, cy, r)


def test_prefetcher_close_skips_cancelled_futures():
    """Closing the request prefetcher while materialisations are still queued
    (serve() stopped early) cancels them without raising, and returns the
    staging buffers of the ones that finished."""
    import threading
    from finitevolumegpu_amd.Serving import _Prefetcher

    class FakePinned:
        def __init__(self):
            self.returned = []

        def put(self, buf):
            self.returned.append(buf)

    pinned = FakePinned()
    p = _Prefetcher(workers=1, pinned=pinned)
    gate = threading.Event()
    p.futures[0] = p.pool.submit(lambda: (gate.wait(5.0), {"_staging": ["buf0"]})[1])
    for i in range(1, 4):  # queued behind the busy worker
        p.futures[i] = p.pool.submit(lambda: {"_staging": ["never"]})
    threading.Timer(0.05, gate.set).start()
    p.close()
    assert pinned.returned == ["buf0"]
    assert p.futures == {}
