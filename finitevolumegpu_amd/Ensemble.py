"""Batched ensemble of independent simulations of one scheme.

The reference advances one grid per simulator object, each substep a separate
kernel launch and every 100 steps a device-to-host read of the CFL minimum
(Simulator.py:176-228).  For many small grids that loop is launch- and
host-bound.  ``Ensemble`` keeps every instance's state, dt policy and time in
a device table (``fvmi_instance_t``) and advances all instances together:

* one work list of wavefront segments covers every active instance, balanced
  so that small and large grids share the device;
* a full step is four launches (prepare, substep 0, substep 1, advance), and
  ``steps_per_graph`` steps are captured once in a hipGraph and replayed; the
  substep kernels read the work list and its length on the device, so the
  graph survives instances joining and leaving (it is captured again only when
  the list outgrows its buffer);
* instance fields live in a ``DeviceArena`` (large device allocations carved
  up on the host), so adding or retiring an instance never calls hipFree,
  which would wait for the whole device;
* between replays the host reads one integer (instances still running) and
  rebuilds the work list when finished instances leave too many idle segments.

Each instance follows BaseSimulator.simulate(t_end) exactly (dt refreshed from
its own CFL every 100 steps, last step shortened, time in double precision)
with the single-grid kernel arithmetic, so its result is bitwise the result of
the plain simulator run alone with the same arithmetic policy.
"""

import bisect
import collections
import ctypes
import logging
import threading

import numpy as np

from . import _lib
from .HipContext import Stream
from .Simulator import BoundaryCondition, resolve_math, swe_initial_cfl

def strip_cols(scheme, math):
    """Output columns of one strip of a scheme's march (strip_cols<kC> in
    fv_substep.h): 124 where it marches two columns per lane
    (EE2D_KP07_dimsplit in fast arithmetic), else 60."""
    return 124 if (scheme == "ee2d_kp07_dimsplit" and math == "fast") else 60


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


def plan_segments(ids, nx, ny, seg, strip_cols=124):
    """Work list of the instances ``ids`` (sizes ``nx``, ``ny``): one int4
    (instance, strip, first row, rows) per run of at most ``seg`` rows of a
    ``strip_cols``-column strip, every strip of every instance covered exactly
    once, longest segments first (ties in instance, strip, row order)."""
    ids = np.asarray(ids, dtype=np.int64)
    sx = (np.asarray(nx, dtype=np.int64) + strip_cols - 1) // strip_cols
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


class DeviceArena:
    """First-fit sub-allocator over large device allocations (``chunk_bytes``
    each, or one of exactly the size of a larger request).  ``alloc`` returns a
    handle (chunk, offset, bytes); ``release`` is host bookkeeping only, with
    neighbouring free blocks merged.  Chunks go back to the device in
    ``close``."""

    ALIGN = 256

    def __init__(self, chunk_bytes=256 << 20):
        self.chunk_bytes = int(chunk_bytes)
        self.chunks = []  # [c_void_p, bytes]
        self.free = []    # per chunk: [offset, bytes] blocks sorted by offset
        self.used = 0

    def alloc(self, nbytes):
        n = -(-int(nbytes) // self.ALIGN) * self.ALIGN
        for c, blocks in enumerate(self.free):
            for j, (off, size) in enumerate(blocks):
                if size >= n:
                    if size == n:
                        del blocks[j]
                    else:
                        blocks[j] = [off + n, size - n]
                    self.used += n
                    return (c, off, n)
        size = max(self.chunk_bytes, n)
        ptr = ctypes.c_void_p()
        _lib.call("fvmi_malloc", ctypes.byref(ptr), size)
        self.chunks.append([ptr, size])
        self.free.append([[n, size - n]] if size > n else [])
        self.used += n
        return (len(self.chunks) - 1, 0, n)

    def address(self, handle):
        return self.chunks[handle[0]][0].value + handle[1]

    def release(self, handle):
        c, off, n = handle
        blocks = self.free[c]
        i = bisect.bisect_left(blocks, [off, 0])
        blocks.insert(i, [off, n])
        if i + 1 < len(blocks) and blocks[i][0] + blocks[i][1] == blocks[i + 1][0]:
            blocks[i][1] += blocks[i + 1][1]
            del blocks[i + 1]
        if i > 0 and blocks[i - 1][0] + blocks[i - 1][1] == blocks[i][0]:
            blocks[i - 1][1] += blocks[i][1]
            del blocks[i]
        self.used -= n

    def reserved(self):
        return sum(size for _, size in self.chunks)

    def close(self):
        for ptr, _ in self.chunks:
            if ptr.value:
                _lib.lib().fvmi_free(ptr)
        self.chunks, self.free, self.used = [], [], 0


class PinnedPool:
    """Page-locked host buffers in power-of-two size classes (at least 1 MiB),
    reused.  Copies from page-locked memory are asynchronous, so an instance's
    upload can queue behind a running block without holding the host thread
    until the block ends.  Thread-safe (requests are staged on worker
    threads)."""

    MIN_BYTES = 1 << 20

    def __init__(self):
        self.lock = threading.Lock()
        self.free = collections.defaultdict(list)
        self.owned = []

    def get(self, nbytes):
        size = self.MIN_BYTES
        while size < nbytes:
            size *= 2
        with self.lock:
            if self.free[size]:
                return self.free[size].pop()
        ptr = ctypes.c_void_p()
        _lib.call("fvmi_host_malloc", ctypes.byref(ptr), size)
        buf = (ptr, size)
        with self.lock:
            self.owned.append(buf)
        return buf

    def put(self, buf):
        with self.lock:
            self.free[buf[1]].append(buf)

    def view(self, buf, shape):
        """float32 array of ``shape`` over the start of ``buf``."""
        count = int(np.prod(shape))
        arr = np.ctypeslib.as_array((ctypes.c_float * count).from_address(buf[0].value))
        return arr.reshape(shape)

    def close(self):
        with self.lock:
            for ptr, _ in self.owned:
                if ptr.value:
                    _lib.lib().fvmi_host_free(ptr)
            self.owned, self.free = [], collections.defaultdict(list)


def instance_initial_cfl(nvars, fields, dx, dy, g, gamma):
    """Host CFL of an instance's initial data, float32 NumPy arithmetic as the
    plain simulators compute it (KP07_dimsplit.py:103-105,
    EE2D_KP07_dimsplit.py:106-108)."""
    dx, dy = np.float32(dx), np.float32(dy)
    if nvars == 4:
        rho, ru, rv = fields[0], fields[1], fields[2]
        gam = np.float32(gamma)
        return np.float32(min(np.min(dx / (np.abs(ru / rho) + np.sqrt(gam * rho))),
                               np.min(dy / (np.abs(rv / rho) + np.sqrt(gam * rho)))))
    return swe_initial_cfl(fields[0], fields[1], fields[2], dx, dy, g)


def instance_pitch(nx):
    """Row pitch in bytes of an instance's padded fields (2 ghost cells a side)."""
    return ((int(nx) + 4) * 4 + 255) // 256 * 256


class Ensemble:
    """``instances``: list of dicts with keys
    ``fields`` (list of (ny, nx) arrays), ``nx``, ``ny``, ``dx``, ``dy``, ``g``,
    ``t_end``, and optionally ``bc`` (BoundaryCondition, packed int or
    'reflective'/'outflow'/'periodic'/'ns_reflective_ew_periodic'), ``theta``,
    ``gamma``, ``cfl_scale`` (default 0.9), ``dt`` (fixed step, as
    simulate(t, dt=...)) and ``cfl0`` (the initial CFL as
    ``instance_initial_cfl`` computes it, when already known).  Fields in
    page-locked memory upload without blocking the host."""

    def __init__(self, context, scheme, instances=(), math=None, steps_per_graph=8, waves_per_block=4,
                 seg_rows=None, min_seg_rows=24, waves_per_simd=4, use_graph=True, capacity=None, stream=None,
                 arena=None):
        self.logger = logging.getLogger(__name__)
        if scheme not in _lib.SCHEME_IDS:
            raise ValueError("unknown scheme %r" % scheme)
        self.context = context
        self.scheme = scheme
        self.scheme_id = _lib.SCHEME_IDS[scheme]
        self.nvars = 4 if scheme == "ee2d_kp07_dimsplit" else 3
        self.math = math = resolve_math(scheme, math)  # None: the scheme's default policy
        self.launch = _lib.LaunchConfig(0, int(waves_per_block), _lib.FVMI_MATH_EXACT if math == "exact"
                                        else _lib.FVMI_MATH_FAST, 1)
        self.steps_per_graph = int(steps_per_graph)
        self.seg_rows = seg_rows
        self.min_seg_rows = int(min_seg_rows)
        self.waves_per_simd = int(waves_per_simd)
        self.use_graph = use_graph
        context.make_current()
        self.stream = stream if stream is not None else Stream()
        self.arena = arena if arena is not None else DeviceArena()
        self.own_arena = arena is None
        self._host_refs = []  # host arrays of copies not yet known to be complete
        instances = list(instances)
        self.n = max(len(instances), int(capacity or 0), 1)
        self.buffers = [None] * self.n  # arena handle of each live slot
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
        self.nsegs_dev = ctypes.c_void_p()
        _lib.call("fvmi_malloc", ctypes.byref(self.nsegs_dev), 4)
        self.segs_dev = None
        self.segs_cap = 0     # segments the device list can hold
        self.segs_host = None  # pinned staging: count + list
        self.num_segs = 0
        self.graph = None
        self.captures = 0
        self.planned_active = None
        self._sync()

    # ------------------------------------------------------------------
    def _init_instance(self, i, spec):
        nx, ny = int(spec["nx"]), int(spec["ny"])
        fields = [np.ascontiguousarray(f, dtype=np.float32) for f in spec["fields"]]
        if len(fields) != self.nvars:
            raise ValueError("instance %d: expected %d fields" % (i, self.nvars))
        for f in fields:
            if f.shape != (ny, nx):
                raise ValueError("instance %d: field shape %s != (%d, %d)" % (i, f.shape, ny, nx))
        # one arena block: [buffer][variable] padded fields of (ny+4) rows each,
        # zeroed (ghost memory included) and the initial state copied in
        pitch = instance_pitch(nx)
        rows = ny + 4
        if rows * pitch >= 1 << 31:  # the kernels' buffer descriptors cover 2 GiB per field
            raise ValueError("instance %d: (ny+4)*pitch must be below 2 GiB" % i)
        handle = self.arena.alloc(2 * self.nvars * rows * pitch)
        base = self.arena.address(handle)
        self.buffers[i] = handle
        _lib.call("fvmi_fill_f32", base, pitch, nx + 4, 2 * self.nvars * rows, 0.0, self.stream.handle)
        I = self.host_table[i]
        for b in range(2):
            for k in range(self.nvars):
                I.field[b][k] = base + (b * self.nvars + k) * rows * pitch
        for k, f in enumerate(fields):
            _lib.call("fvmi_memcpy2d_async", I.field[0][k] + 2 * pitch + 8, pitch, f.ctypes.data, nx * 4, nx * 4, ny,
                      _lib.MEMCPY_HTOD, self.stream.handle)
            self._host_refs.append(f)
        dx, dy = np.float32(spec["dx"]), np.float32(spec["dy"])
        g = np.float32(spec["g"])
        I.pitch, I.nx, I.ny = pitch, nx, ny
        I.bc = _bc_code(spec.get("bc", BoundaryCondition()))
        I.dx, I.dy, I.g = dx, dy, g
        default_theta = 1.8 if self.scheme == "hll2" else 1.3
        I.theta = np.float32(spec.get("theta", default_theta))
        gamma = spec.get("gamma")
        I.gamma = np.float32((1.4 if self.nvars == 4 else 0.0) if gamma is None else gamma)
        I.sqrt_g = np.sqrt(g) if g > 0 else np.float32(0.0)
        I.inv_g = np.float32(1.0) / g if g != 0 else np.float32(0.0)
        I.inv_dx = np.float32(1.0) / dx
        I.inv_dy = np.float32(1.0) / dy
        c0 = spec.get("cfl0")  # precomputed off the admitting thread (Serving.request_spec)
        if c0 is None:
            c0 = instance_initial_cfl(self.nvars, fields, dx, dy, spec["g"], I.gamma)
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
    def _sync(self):
        self.stream.synchronize()
        self._host_refs = []

    def _download_table(self):
        _lib.call("fvmi_memcpy_async", ctypes.addressof(self.host_table), self.table,
                  ctypes.sizeof(self.host_table), _lib.MEMCPY_DTOH, self.stream.handle)
        self._sync()

    def _running(self):
        """Live slots whose instance is still active and short of t_end (host copy)."""
        return self.live & (self.tab["active"] != 0) & (self.tab["t"] < self.tab["t_end"])

    def _active_mask(self):
        self._download_table()
        return self._running()

    def _strip_rows(self, mask):
        """Rows of strips of the instances in ``mask`` (sum over them)."""
        nx = self.tab["nx"][mask].astype(np.int64)
        ny = self.tab["ny"][mask].astype(np.int64)
        w = strip_cols(self.scheme, self.math)
        return int(((nx + w - 1) // w * ny).sum())

    # ------------------------------------------------------------------
    # slots
    def free_slot(self):
        free = np.flatnonzero(~self.live)
        return int(free[0]) if free.size else None

    def add(self, spec):
        """Puts a new instance in a free slot and returns the slot.  The copies
        are queued on the ensemble's stream behind any block in flight, so an
        instance can join between replays without waiting for the device."""
        i = self.free_slot()
        if i is None:
            raise RuntimeError("ensemble is full (%d slots)" % self.n)
        self._init_instance(i, spec)
        self.live[i] = True
        size = ctypes.sizeof(_lib.Instance)
        _lib.call("fvmi_memcpy_async", self.table.value + i * size, ctypes.addressof(self.host_table[i]), size,
                  _lib.MEMCPY_HTOD, self.stream.handle)
        self.planned_active = None  # work list rebuilt at the next advance
        return i

    def remove(self, i):
        """Frees slot i (its instance must have finished or be abandoned).  Waits
        for the ensemble's stream before its memory can be handed out again."""
        if not self.live[i]:
            return
        self.host_table[i].active = 0
        size = ctypes.sizeof(_lib.Instance)
        _lib.call("fvmi_memcpy_async", self.table.value + i * size, ctypes.addressof(self.host_table[i]), size,
                  _lib.MEMCPY_HTOD, self.stream.handle)
        self._sync()
        self.arena.release(self.buffers[i])
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
        return 2 * self.nvars * (int(spec["ny"]) + 4) * instance_pitch(spec["nx"])

    def plan(self, active):
        """Work list for the instances in ``active``: segments of at most seg_rows
        rows, never crossing a strip, longest first.  The list and its
        length are written to the device on the ensemble's stream; the graph
        is captured again only if the list outgrows its buffer."""
        active = np.asarray(active, dtype=bool) & self.live
        total = self._strip_rows(active)
        if self.seg_rows:
            seg = int(self.seg_rows)
        else:
            slots = self.context.cu_count * 4 * self.waves_per_simd
            seg = max(self.min_seg_rows, -(-total // max(1, slots)))
        arr = plan_segments(np.flatnonzero(active), self.tab["nx"][active], self.tab["ny"][active], seg,
                            strip_cols(self.scheme, self.math))
        need = len(arr)
        self._sync()  # the pinned staging buffer may still feed the previous upload
        if self.segs_dev is None or need > self.segs_cap:
            for name in ("segs_dev",):
                if getattr(self, name) is not None:
                    _lib.call("fvmi_free", getattr(self, name))
            if self.segs_host is not None:
                _lib.call("fvmi_host_free", self.segs_host)
            self.segs_cap = max(256, 2 * need)
            self.segs_dev = ctypes.c_void_p()
            _lib.call("fvmi_malloc", ctypes.byref(self.segs_dev), 16 * self.segs_cap)
            self.segs_host = ctypes.c_void_p()
            _lib.call("fvmi_host_malloc", ctypes.byref(self.segs_host), 16 * self.segs_cap + 16)
            self._drop_graph()
        staging = np.ctypeslib.as_array(ctypes.cast(self.segs_host, ctypes.POINTER(ctypes.c_int32)),
                                        shape=(4 * self.segs_cap + 4,))
        staging[0] = need
        if need:
            staging[4:4 + 4 * need] = arr.reshape(-1)
            _lib.call("fvmi_memcpy_async", self.segs_dev, self.segs_host.value + 16, 16 * need, _lib.MEMCPY_HTOD,
                      self.stream.handle)
        _lib.call("fvmi_memcpy_async", self.nsegs_dev, self.segs_host, 4, _lib.MEMCPY_HTOD, self.stream.handle)
        self.num_segs = need
        self.planned_active = active.copy()
        self.planned_rows = total

    def _waves(self):
        """Wavefronts of a substep launch: one per slot of the work list's buffer
        (at most twice the segments when it was sized), so every segment gets
        its own wave and the hardware hands the longest-first list to SIMDs as
        they free up; waves past the list's length exit at once."""
        return int(self.segs_cap)

    def _enqueue_steps(self, k):
        waves = self._waves()
        for _ in range(k):
            _lib.call("fvmi_ensemble_prepare", self.table, self.n, self.stream.handle)
            for step in (0, 1):
                _lib.call("fvmi_ensemble_substep", self.scheme_id, self.table, self.segs_dev, self.nsegs_dev, waves,
                          step, self.launch, self.stream.handle)
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
        self.captures += 1

    def _drop_graph(self):
        if self.graph is not None:
            _lib.call("fvmi_graph_destroy", self.graph)
            self.graph = None

    def count_active(self):
        _lib.call("fvmi_ensemble_count_active", self.table, self.n, self.count_dev, self.stream.handle)
        _lib.call("fvmi_memcpy_async", self.count_host, self.count_dev, 4, _lib.MEMCPY_DTOH, self.stream.handle)
        self._sync()
        return ctypes.c_int.from_address(self.count_host.value).value

    def advance(self, replan_fraction=0.75, wait=True):
        """Runs one block of ``steps_per_graph`` steps; returns the number of
        instances still running afterwards (``wait=False``: only enqueues and
        returns None)."""
        if self.planned_active is None:
            # the host copy of the table can only be behind the device, so its
            # running set contains every instance that still has to step
            self.plan(self._running())
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
        """Interior of every field of instance i (float32 (ny, nx) arrays)."""
        I = self.host_table[i]
        nx, ny, pitch = int(I.nx), int(I.ny), int(I.pitch)
        out = []
        for k in range(self.nvars):
            a = np.empty((ny, nx), dtype=np.float32)
            _lib.call("fvmi_memcpy2d_async", a.ctypes.data, nx * 4, I.field[0][k] + 2 * pitch + 8, pitch, nx * 4, ny,
                      _lib.MEMCPY_DTOH, self.stream.handle)
            out.append(a)
        self._sync()
        return out

    def nonfinite_counts(self, slots):
        """Number of non-finite interior values of each instance in ``slots``
        (every field of its state buffer), counted on the device; one copy
        back for all of them."""
        slots = [int(i) for i in slots]
        if not slots:
            return np.zeros(0, dtype=np.int64)
        need = 16 * len(slots) * self.nvars
        if getattr(self, "stats_cap", 0) < need:
            if getattr(self, "stats_dev", None) is not None:
                self._sync()
                _lib.call("fvmi_free", self.stats_dev)
            self.stats_dev = ctypes.c_void_p()
            _lib.call("fvmi_malloc", ctypes.byref(self.stats_dev), need)
            self.stats_cap = need
        base = self.stats_dev.value
        _lib.call("fvmi_fill_f32", base, need, need // 4, 1, 0.0, self.stream.handle)
        for j, i in enumerate(slots):
            I = self.host_table[i]
            for k in range(self.nvars):
                out = base + 16 * (j * self.nvars + k)
                _lib.call("fvmi_field_stats", int(I.nx), int(I.ny), 2, 2, I.field[0][k], int(I.pitch), out, out + 8,
                          self.stream.handle)
        host = np.zeros(need // 8, dtype=np.float64)
        _lib.call("fvmi_memcpy_async", host.ctypes.data, base, need, _lib.MEMCPY_DTOH, self.stream.handle)
        self._sync()
        bad = host.view(np.uint32)[2::4].reshape(len(slots), self.nvars)
        return bad.sum(axis=1).astype(np.int64)

    def cell_steps(self):
        """Sum over live instances of nx*ny*steps taken."""
        self._download_table()
        m = self.live
        return int((self.tab["nx"][m].astype(np.int64) * self.tab["ny"][m] * self.tab["nt"][m]).sum())

    def close(self):
        if getattr(self, "graph", None) is not None:
            _lib.lib().fvmi_graph_destroy(self.graph)
            self.graph = None
        if getattr(self, "stream", None) is not None:
            self.stream.synchronize()
        for p in ("segs_dev", "nsegs_dev", "table", "count_dev", "stats_dev"):
            ptr = getattr(self, p, None)
            if ptr is not None and ptr.value:
                _lib.lib().fvmi_free(ptr)
            setattr(self, p, None)
        for p in ("count_host", "segs_host"):
            ptr = getattr(self, p, None)
            if ptr is not None and ptr.value:
                _lib.lib().fvmi_host_free(ptr)
            setattr(self, p, None)
        arena = getattr(self, "arena", None)
        if arena is not None:
            if self.own_arena:
                arena.close()
            else:
                for handle in self.buffers:
                    if handle is not None:
                        arena.release(handle)
        self.arena = None
        self.buffers = [None] * self.n
        self.live[:] = False

    def __del__(self):
        if _lib._lib is not None:
            try:
                self.close()
            except Exception:  # noqa: BLE001 - interpreter shutdown
                pass
