"""Batched ensemble of independent simulations of one scheme.

The reference advances one grid per simulator object, each substep a separate
kernel launch and every 100 steps a device-to-host read of the CFL minimum
(Simulator.py:176-228).  For many small grids that loop is launch- and
host-bound.  ``Ensemble`` keeps every instance's state, dt policy and time in
a device table (``fvmi_instance_t``) and advances all instances together:

* one work list of wavefront segments covers every active instance, balanced
  so that small and large grids share the device;
* a full step is four launches (prepare, substep 0, substep 1, advance), and
  ``steps_per_graph`` steps are captured once in a hipGraph and replayed;
* between replays the host reads one integer (instances still running) and
  rebuilds the work list when finished instances leave too many idle segments.

Each instance follows BaseSimulator.simulate(t_end) exactly (dt refreshed from
its own CFL every 100 steps, last step shortened, time in double precision)
with the single-grid kernel arithmetic, so its result is bitwise the result of
the plain simulator run alone with the same arithmetic policy.
"""

import ctypes
import logging

import numpy as np

from . import _lib
from .Common import HipArray2D
from .HipContext import Stream
from .Simulator import BoundaryCondition, swe_initial_cfl

STRIP_COLS = 60


def _bc_code(bc):
    if isinstance(bc, BoundaryCondition):
        return int(bc.asCodedInt())
    if isinstance(bc, str):
        T = BoundaryCondition.Type
        if bc == "ns_reflective_ew_periodic":  # the Rayleigh-Taylor column
            return int(BoundaryCondition({"north": T.Reflective, "south": T.Reflective, "east": T.Periodic,
                                          "west": T.Periodic}).asCodedInt())
        kind = {"reflective": T.Reflective, "outflow": T.Outflow, "periodic": T.Periodic}[bc]
        return int(BoundaryCondition.uniform(kind).asCodedInt())
    return int(bc)


def plan_segments(ids, nx, ny, seg):
    """Work list of the instances ``ids`` (sizes ``nx``, ``ny``): one int4
    (instance, strip, first row, rows) per run of at most ``seg`` rows of a
    60-column strip, every strip of every instance covered exactly once,
    longest segments first (ties in instance, strip, row order)."""
    ids = np.asarray(ids, dtype=np.int64)
    sx = (np.asarray(nx, dtype=np.int64) + STRIP_COLS - 1) // STRIP_COLS
    ny = np.asarray(ny, dtype=np.int64)
    per_strip = (ny + seg - 1) // seg
    count = sx * per_strip
    if count.sum() == 0:
        return np.zeros((0, 4), dtype=np.int32)
    inst = np.repeat(np.arange(ids.size), count)
    start = np.repeat(np.cumsum(count) - count, count)
    j = np.arange(count.sum()) - start
    strip = j // per_strip[inst]
    r0 = (j % per_strip[inst]) * seg
    rows = np.minimum(seg, ny[inst] - r0)
    order = np.argsort(-rows, kind="stable")
    out = np.stack([ids[inst], strip, r0, rows], axis=1)[order]
    return np.ascontiguousarray(out, dtype=np.int32)


class Ensemble:
    """``instances``: list of dicts with keys
    ``fields`` (list of (ny, nx) arrays), ``nx``, ``ny``, ``dx``, ``dy``, ``g``,
    ``t_end``, and optionally ``bc`` (BoundaryCondition, packed int or
    'reflective'/'outflow'/'periodic'), ``theta``, ``gamma``, ``cfl_scale``
    (default 0.9) and ``dt`` (fixed step, as simulate(t, dt=...))."""

    def __init__(self, context, scheme, instances=(), math="fast", steps_per_graph=8, waves_per_block=4,
                 seg_rows=None, min_seg_rows=24, waves_per_simd=4, use_graph=True, capacity=None, stream=None):
        self.logger = logging.getLogger(__name__)
        if scheme not in _lib.SCHEME_IDS:
            raise ValueError("unknown scheme %r" % scheme)
        self.context = context
        self.scheme = scheme
        self.scheme_id = _lib.SCHEME_IDS[scheme]
        self.nvars = 4 if scheme == "ee2d_kp07_dimsplit" else 3
        self.math = math
        self.launch = _lib.LaunchConfig(0, int(waves_per_block), _lib.FVMI_MATH_EXACT if math == "exact"
                                        else _lib.FVMI_MATH_FAST, 1)
        self.steps_per_graph = int(steps_per_graph)
        self.seg_rows = seg_rows
        self.min_seg_rows = int(min_seg_rows)
        self.waves_per_simd = int(waves_per_simd)
        self.use_graph = use_graph
        context.make_current()
        self.stream = stream if stream is not None else Stream()
        instances = list(instances)
        self.n = max(len(instances), int(capacity or 0), 1)
        self.buffers = [None] * self.n
        self.live = np.zeros(self.n, dtype=bool)  # slots holding an instance
        self.host_table = (_lib.Instance * self.n)()
        self.tab = np.frombuffer(self.host_table, dtype=_lib.INSTANCE_DTYPE)  # numpy view, same memory
        for i, spec in enumerate(instances):
            self._init_instance(i, spec)
            self.live[i] = True
        self.table = ctypes.c_void_p()
        _lib.call("fvmi_malloc", ctypes.byref(self.table), ctypes.sizeof(self.host_table))
        _lib.call("fvmi_memcpy_async", self.table, ctypes.addressof(self.host_table),
                  ctypes.sizeof(self.host_table), _lib.MEMCPY_HTOD, self.stream.handle)
        self.count_dev = ctypes.c_void_p()
        _lib.call("fvmi_malloc", ctypes.byref(self.count_dev), 4)
        self.count_host = ctypes.c_void_p()
        _lib.call("fvmi_host_malloc", ctypes.byref(self.count_host), 4)
        self.segs_dev = None
        self.segs_bytes = 0
        self.graph = None
        self.planned_active = None
        self.stream.synchronize()

    # ------------------------------------------------------------------
    def _init_instance(self, i, spec):
        nx, ny = int(spec["nx"]), int(spec["ny"])
        fields = [np.ascontiguousarray(f, dtype=np.float32) for f in spec["fields"]]
        if len(fields) != self.nvars:
            raise ValueError("instance %d: expected %d fields" % (i, self.nvars))
        for f in fields:
            if f.shape != (ny, nx):
                raise ValueError("instance %d: field shape %s != (%d, %d)" % (i, f.shape, ny, nx))
        bufs = [[HipArray2D(self.stream, nx, ny, 2, 2, f) for f in fields],
                [HipArray2D(self.stream, nx, ny, 2, 2, None) for _ in fields]]
        pitch = bufs[0][0].pitch
        if any(b.pitch != pitch for row in bufs for b in row):
            raise RuntimeError("fields of one instance must share a pitch")
        self.buffers[i] = bufs
        I = self.host_table[i]
        for b in range(2):
            for k in range(self.nvars):
                I.field[b][k] = bufs[b][k].data_ptr
        dx, dy = np.float32(spec["dx"]), np.float32(spec["dy"])
        g = np.float32(spec["g"])
        I.pitch, I.nx, I.ny = pitch, nx, ny
        I.bc = _bc_code(spec.get("bc", BoundaryCondition()))
        I.dx, I.dy, I.g = dx, dy, g
        default_theta = 1.8 if self.scheme == "hll2" else 1.3
        I.theta = np.float32(spec.get("theta", default_theta))
        I.gamma = np.float32(spec.get("gamma", 1.4 if self.nvars == 4 else 0.0))
        I.sqrt_g = np.sqrt(g) if g > 0 else np.float32(0.0)
        I.inv_g = np.float32(1.0) / g if g != 0 else np.float32(0.0)
        I.inv_dx = np.float32(1.0) / dx
        I.inv_dy = np.float32(1.0) / dy
        if self.nvars == 4:
            rho, ru, rv = fields[0], fields[1], fields[2]
            gam = np.float32(I.gamma)
            c0 = np.float32(min(np.min(dx / (np.abs(ru / rho) + np.sqrt(gam * rho))),
                                np.min(dy / (np.abs(rv / rho) + np.sqrt(gam * rho)))))
        else:
            c0 = swe_initial_cfl(fields[0], fields[1], fields[2], dx, dy, spec["g"])
        I.cfl[0] = np.inf
        I.cfl[1] = c0
        I.cfl_scale = float(spec.get("cfl_scale", 0.9))
        I.t = 0.0
        I.t_end = float(spec["t_end"])
        I.nt = 0
        I.active = 1
        if spec.get("dt") is not None:
            I.fixed_dt = 1
            I.dt = float(spec["dt"])
        else:
            I.fixed_dt = 0
            I.dt = 0.0
        I.dt_step = 0.0

    # ------------------------------------------------------------------
    def _download_table(self):
        _lib.call("fvmi_memcpy_async", ctypes.addressof(self.host_table), self.table,
                  ctypes.sizeof(self.host_table), _lib.MEMCPY_DTOH, self.stream.handle)
        self.stream.synchronize()

    def _running(self):
        """Live slots whose instance is still active and short of t_end (host copy)."""
        return self.live & (self.tab["active"] != 0) & (self.tab["t"] < self.tab["t_end"])

    def _active_mask(self):
        self._download_table()
        return self._running()

    def _strip_rows(self, mask):
        """Rows of 60-column strips of the instances in ``mask`` (sum over them)."""
        nx = self.tab["nx"][mask].astype(np.int64)
        ny = self.tab["ny"][mask].astype(np.int64)
        return int(((nx + STRIP_COLS - 1) // STRIP_COLS * ny).sum())

    # ------------------------------------------------------------------
    # slots
    def free_slot(self):
        free = np.flatnonzero(~self.live)
        return int(free[0]) if free.size else None

    def add(self, spec):
        """Puts a new instance in a free slot (between steps) and returns the slot."""
        i = self.free_slot()
        if i is None:
            raise RuntimeError("ensemble is full (%d slots)" % self.n)
        self._init_instance(i, spec)
        self.live[i] = True
        size = ctypes.sizeof(_lib.Instance)
        _lib.call("fvmi_memcpy_async", self.table.value + i * size, ctypes.addressof(self.host_table[i]), size,
                  _lib.MEMCPY_HTOD, self.stream.handle)
        self.stream.synchronize()
        self.planned_active = None  # work list rebuilt at the next advance
        return i

    def remove(self, i):
        """Frees slot i (its instance must have finished or be abandoned)."""
        if not self.live[i]:
            return
        self.host_table[i].active = 0
        size = ctypes.sizeof(_lib.Instance)
        _lib.call("fvmi_memcpy_async", self.table.value + i * size, ctypes.addressof(self.host_table[i]), size,
                  _lib.MEMCPY_HTOD, self.stream.handle)
        self.stream.synchronize()
        for row in self.buffers[i]:
            for b in row:
                b.free()
        self.buffers[i] = None
        self.live[i] = False
        if self.planned_active is not None and self.planned_active[i]:
            self.planned_active = None

    def finished(self):
        """Live slots whose instance reached t_end."""
        self._download_table()
        return [int(i) for i in np.flatnonzero(self.live & ~self._running())]

    def bytes_of(self, spec):
        """Device memory an instance would take (two padded buffers per field)."""
        nx, ny = int(spec["nx"]), int(spec["ny"])
        pitch = ((nx + 4) * 4 + 255) // 256 * 256
        return 2 * self.nvars * (ny + 4) * pitch

    def plan(self, active):
        """Work list for the instances in ``active``: segments of at most seg_rows
        rows, never crossing a 60-column strip, longest first."""
        active = np.asarray(active, dtype=bool) & self.live
        total = self._strip_rows(active)
        if self.seg_rows:
            seg = int(self.seg_rows)
        else:
            slots = self.context.cu_count * 4 * self.waves_per_simd
            seg = max(self.min_seg_rows, -(-total // max(1, slots)))
        arr = plan_segments(np.flatnonzero(active), self.tab["nx"][active], self.tab["ny"][active], seg)
        if self.segs_dev is None or arr.nbytes > self.segs_bytes:
            # grow only (a free would synchronise the device under other pools' work)
            if self.segs_dev is not None:
                _lib.call("fvmi_free", self.segs_dev)
            self.segs_bytes = max(4096, 2 * arr.nbytes)
            self.segs_dev = ctypes.c_void_p()
            _lib.call("fvmi_malloc", ctypes.byref(self.segs_dev), self.segs_bytes)
        if arr.size:
            _lib.call("fvmi_memcpy_async", self.segs_dev, arr.ctypes.data, arr.nbytes, _lib.MEMCPY_HTOD,
                      self.stream.handle)
        self.num_segs = len(arr)
        self.planned_active = active.copy()
        self.planned_rows = total
        self.stream.synchronize()
        if self.graph is not None:
            _lib.call("fvmi_graph_destroy", self.graph)
            self.graph = None

    def _enqueue_steps(self, k):
        for _ in range(k):
            _lib.call("fvmi_ensemble_prepare", self.table, self.n, self.stream.handle)
            for step in (0, 1):
                _lib.call("fvmi_ensemble_substep", self.scheme_id, self.table, self.segs_dev, self.num_segs, step,
                          self.launch, self.stream.handle)
            _lib.call("fvmi_ensemble_advance", self.table, self.n, self.stream.handle)

    def _capture(self):
        _lib.call("fvmi_stream_begin_capture", self.stream.handle)
        try:
            self._enqueue_steps(self.steps_per_graph)
        finally:
            g = ctypes.c_void_p()
            rc = _lib.lib().fvmi_stream_end_capture(self.stream.handle, ctypes.byref(g))
        _lib.check(rc, "fvmi_stream_end_capture")
        self.graph = g

    def count_active(self):
        _lib.call("fvmi_ensemble_count_active", self.table, self.n, self.count_dev, self.stream.handle)
        _lib.call("fvmi_memcpy_async", self.count_host, self.count_dev, 4, _lib.MEMCPY_DTOH, self.stream.handle)
        self.stream.synchronize()
        return ctypes.c_int.from_address(self.count_host.value).value

    def advance(self, replan_fraction=0.75, wait=True):
        """Runs one block of ``steps_per_graph`` steps; returns the number of
        instances still running afterwards (``wait=False``: only enqueues and
        returns None)."""
        if self.planned_active is None:
            self.plan(self._active_mask())
        if self.num_segs == 0:
            return self.count_active() if wait else None
        if self.use_graph:
            if self.graph is None:
                self._capture()
            _lib.call("fvmi_graph_launch", self.graph, self.stream.handle)
        else:
            self._enqueue_steps(self.steps_per_graph)
        if not wait:
            return None
        left = self.count_active()
        if left and left < self.planned_active.sum():
            active = self._active_mask()
            if self._strip_rows(active) < replan_fraction * self.planned_rows:
                self.plan(active)
        return left

    def run(self, max_blocks=None):
        """Advances until every instance reached its t_end."""
        blocks = 0
        while True:
            left = self.advance()
            blocks += 1
            if left == 0 or (max_blocks is not None and blocks >= max_blocks):
                return blocks

    def state(self, i):
        """(t, nt) of instance i."""
        self._download_table()
        return self.host_table[i].t, self.host_table[i].nt

    def download(self, i):
        return [b.download(self.stream) for b in self.buffers[i][0]]

    def cell_steps(self):
        """Sum over live instances of nx*ny*steps taken."""
        self._download_table()
        m = self.live
        return int((self.tab["nx"][m].astype(np.int64) * self.tab["ny"][m] * self.tab["nt"][m]).sum())

    def close(self):
        if self.graph is not None:
            _lib.lib().fvmi_graph_destroy(self.graph)
            self.graph = None
        for p in ("segs_dev", "table", "count_dev"):
            ptr = getattr(self, p, None)
            if ptr is not None and ptr.value:
                _lib.lib().fvmi_free(ptr)
                setattr(self, p, None)
        if getattr(self, "count_host", None) is not None and self.count_host.value:
            _lib.lib().fvmi_host_free(self.count_host)
            self.count_host = None
        for bufs in self.buffers:
            if bufs is None:
                continue
            for row in bufs:
                for b in row:
                    b.free()
        self.buffers = [None] * self.n
        self.live[:] = False

    def __del__(self):
        if _lib._lib is not None:
            try:
                self.close()
            except Exception:  # noqa: BLE001 - interpreter shutdown
                pass
