"""Boundary conditions and the simulator base class (GPUSimulators/Simulator.py).

``BoundaryCondition`` and ``BaseSimulator`` keep the reference's public API:
``step(dt)``, ``simulate(t, dt=None)``, ``download(variables=None)``,
``synchronize()``, ``simTime()``, ``simSteps()``, ``getExtent()``,
``setBoundaryConditions()``, ``getBoundaryConditions()``, ``check()``,
``computeDt()`` and the dt policy of Simulator.py:176-228 (dt refreshed from the
device CFL minimum every 100 steps, last step shortened to hit ``t_end``,
simulation time accumulated in double precision).

``DimsplitSimulator`` holds what the four scheme wrappers share
(KP07_dimsplit.py:55-137 and its HLL2 / WAF / EE2D siblings): two padded states
``u0``/``u1`` swapped after every substep, the device CFL slots, the initial
CFL from the host arrays, and the ``substep(dt, i)`` launch with ``dt*0.5``.
"""

import ctypes
import logging
from enum import IntEnum

import numpy as np

from . import _lib
from . import Common
from .HipContext import Stream


class BoundaryCondition:
    """Boundary conditions of the four domain edges (Simulator.py:38-100)."""

    class Type(IntEnum):
        """Codes shared with the kernels (common.h:109-114).  ``Outflow`` (4) is a
        zero-gradient ghost copy with no counterpart in the reference."""
        Dirichlet = 0
        Neumann = 1
        Periodic = 2
        Reflective = 3
        Outflow = 4

    def __init__(self, types=None):
        if types is None:
            types = {"north": BoundaryCondition.Type.Reflective, "south": BoundaryCondition.Type.Reflective,
                     "east": BoundaryCondition.Type.Reflective, "west": BoundaryCondition.Type.Reflective}
        self.north = BoundaryCondition.Type(types["north"])
        self.south = BoundaryCondition.Type(types["south"])
        self.east = BoundaryCondition.Type(types["east"])
        self.west = BoundaryCondition.Type(types["west"])
        if BoundaryCondition.Type.Neumann in (self.north, self.south, self.east, self.west):
            raise NotImplementedError("Neumann boundary condition not supported")

    def __str__(self):
        return "[north={:s}, south={:s}, east={:s}, west={:s}]".format(
            str(self.north), str(self.south), str(self.east), str(self.west))

    def asCodedInt(self):
        """Packs the four types into N<<24 | S<<16 | E<<8 | W (Simulator.py:78-92)."""
        bc = 0
        bc |= (int(self.north) & 0xF) << 24
        bc |= (int(self.south) & 0xF) << 16
        bc |= (int(self.east) & 0xF) << 8
        bc |= (int(self.west) & 0xF) << 0
        return np.int32(bc)

    @staticmethod
    def getTypes(bc):
        bc = int(bc)
        return {"north": BoundaryCondition.Type((bc >> 24) & 0xF),
                "south": BoundaryCondition.Type((bc >> 16) & 0xF),
                "east": BoundaryCondition.Type((bc >> 8) & 0xF),
                "west": BoundaryCondition.Type(bc & 0xF)}

    @staticmethod
    def uniform(kind):
        t = BoundaryCondition.Type(kind)
        return BoundaryCondition({"north": t, "south": t, "east": t, "west": t})


class BaseSimulator:
    """Stepping loop and dt policy (Simulator.py:109-266)."""

    def __init__(self, context, nx, ny, dx, dy, boundary_conditions, cfl_scale, num_substeps,
                 block_width, block_height):
        self.logger = logging.getLogger(__name__ + "." + self.__class__.__name__)
        self.context = context
        self.nx = np.int32(nx)
        self.ny = np.int32(ny)
        self.dx = np.float32(dx)
        self.dy = np.float32(dy)
        self.setBoundaryConditions(boundary_conditions)
        self.cfl_scale = cfl_scale
        self.num_substeps = num_substeps
        # Accepted for the reference's signature; the gfx950 kernels are shaped by
        # tile_rows / waves_per_block instead (see DimsplitSimulator).
        self.block_size = (int(block_width), int(block_height), 1)
        self.grid_size = (int(np.ceil(self.nx / float(block_width))), int(np.ceil(self.ny / float(block_height))))
        context.make_current()
        self.stream = Stream()
        self.internal_stream = Stream()
        self.t = 0.0
        self.nt = 0
        self.dt = None

    def __str__(self):
        return "{:s} [{:d}x{:d}]".format(self.__class__.__name__, int(self.nx), int(self.ny))

    def simulate(self, t, dt=None):
        """Advances ``t`` time units (Simulator.py:176-217)."""
        printer = Common.ProgressPrinter(t)
        t_start = self.simTime()
        t_end = t_start + t
        update_dt = True
        if dt is not None:
            update_dt = False
            self.dt = dt
        while self.simTime() < t_end:
            if update_dt and (self.simSteps() % 100 == 0):
                self.dt = self.computeDt() * self.cfl_scale
            current_dt = np.float32(min(self.dt, t_end - self.simTime()))
            if current_dt <= 0.0:
                self.logger.warning("step %d: no time left to advance (dt %g), stopping", self.simSteps(),
                                    float(current_dt))
                break
            self.step(current_dt)
            print_string = printer.getPrintString(self.simTime() - t_start)
            if print_string:
                self.logger.info("%s: %s", self, print_string)
                try:
                    self.check()
                except AssertionError as e:
                    # where the check failed, as "Step=<n>, time=<t>" (callers of the
                    # reference read this suffix)
                    e.args = e.args + ("Step=%d, time=%f" % (self.simSteps(), self.simTime()),)
                    raise

    def step(self, dt):
        """One full step of ``num_substeps`` substeps (Simulator.py:220-228)."""
        for i in range(self.num_substeps):
            self.substep(dt, i)
        self.t += float(dt)
        self.nt += 1

    def download(self, variables=None):
        return self.getOutput().download(self.stream, variables)

    def synchronize(self):
        self.stream.synchronize()

    def simTime(self):
        return self.t

    def simSteps(self):
        return self.nt

    def getExtent(self):
        return [0, 0, self.nx * self.dx, self.ny * self.dy]

    def setBoundaryConditions(self, boundary_conditions):
        self.logger.debug("Boundary conditions set to {:s}".format(str(boundary_conditions)))
        self.boundary_conditions = boundary_conditions.asCodedInt()

    def getBoundaryConditions(self):
        return BoundaryCondition(BoundaryCondition.getTypes(self.boundary_conditions))

    def substep(self, dt, step_number):
        raise NotImplementedError("Needs to be implemented in subclass")

    def getOutput(self):
        raise NotImplementedError("Needs to be implemented in subclass")

    def check(self):
        self.logger.warning("check() is not implemented - please implement")

    def computeDt(self):
        raise NotImplementedError("Needs to be implemented in subclass")


_MATH = {"fast": _lib.FVMI_MATH_FAST, "exact": _lib.FVMI_MATH_EXACT}

# Arithmetic policy of each scheme when the caller passes math=None: the fast
# policy where it stays within relative L1 1e-5 per field of the float32
# reference arithmetic, the exact one (bitwise equal to it) where it does not.
# WAF's limiter turns rounding differences in its upwind selections and ratios
# into O(1) changes of the limiter weights (the float32 and float64 oracles
# already differ by 2e-5 in the momenta of the 512^2 dam break), and the Euler
# momenta (~1e-2 against a pressure of 2.5 in the Rayleigh-Taylor and other
# near-rest states) move by 1e-5 of their size for one rounding of the
# pressure flux, so both default to exact arithmetic.
DEFAULT_MATH = {"kp07_dimsplit": "fast", "hll2": "fast", "waf": "exact", "ee2d_kp07_dimsplit": "exact",
                "lxf": "fast", "hll": "fast", "force": "fast", "kp07": "fast"}


def resolve_math(scheme, math):
    """The policy a simulator of ``scheme`` runs with: ``math`` when given,
    else the scheme's default (DEFAULT_MATH)."""
    if math is None:
        math = DEFAULT_MATH[scheme]
    if math not in _MATH:
        raise ValueError("math must be 'fast', 'exact' or None")
    return math

class DimsplitSimulator(BaseSimulator):
    """Shared body of the four dimensionally split scheme wrappers."""

    scheme = None  # "kp07_dimsplit" | "hll2" | "waf" | "ee2d_kp07_dimsplit"
    nvars = 3
    # default launch shape when neither the caller nor the autotuner gives one
    default_waves_per_block = 4

    def _init_dimsplit(self, fields, cfl_host, tile_rows, waves_per_block, math, xcd_remap):
        if self.context.autotuner is not None and tile_rows is None and waves_per_block is None:
            peak = self.context.autotuner.get_peak_performance(self.__class__)
            tile_rows = int(peak["tile_rows"])
            waves_per_block = int(peak["waves_per_block"])
            self.logger.debug("Autotuned launch shape: %d rows x %d waves", tile_rows, waves_per_block)
        math = resolve_math(self.scheme, math)
        self.math = math
        # tile_rows = rows per wavefront segment; 0 lets the library balance the
        # segments over the device's SIMDs
        self.launch = _lib.LaunchConfig(int(tile_rows or 0),
                                        int(waves_per_block or self.default_waves_per_block),
                                        _MATH[math], 1 if xcd_remap else 2)
        self.gc_x = 2
        self.gc_y = 2
        self.u0 = Common.ArakawaA2D(self.stream, self.nx, self.ny, 2, 2, fields)
        self.u1 = Common.ArakawaA2D(self.stream, self.nx, self.ny, 2, 2, [None] * len(fields))
        # Two-slot device CFL minimum (see fvmi.h): slot 0 = +inf, slot 1 = initial value.
        ptr = ctypes.c_void_p()
        _lib.call("fvmi_malloc", ctypes.byref(ptr), 8)
        self.cfl_data = ptr
        init = np.array([np.inf, cfl_host], dtype=np.float32)
        _lib.call("fvmi_memcpy_async", self.cfl_data, init.ctypes.data, 8, _lib.MEMCPY_HTOD, self.stream.handle)
        self.stream.synchronize()

    def cfl_pointer(self):
        return self.cfl_data.value

    def substep(self, dt, step_number):
        self.substepDimsplit(dt * 0.5, step_number)

    def substepDimsplit(self, dt, substep):
        self._launch(np.float32(dt), int(substep), self.u0, self.u1, self.cfl_pointer())
        self.u0, self.u1 = self.u1, self.u0

    def _field_args(self, state):
        args = []
        for k in range(self.nvars):
            args += [state[k].data_ptr, state[k].pitch]
        return args

    def getOutput(self):
        return self.u0

    def check(self):
        self.u0.check(self.stream)
        self.u1.check(self.stream)

    def computeDt(self):
        """0.5 * CFL minimum of the last substep (KP07_dimsplit.py:135-137)."""
        host = np.zeros(2, dtype=np.float32)
        _lib.call("fvmi_memcpy_async", host.ctypes.data, self.cfl_data, 8, _lib.MEMCPY_DTOH, self.stream.handle)
        self.stream.synchronize()
        # float32 minimum promoted to double as the reference's 0-d array times 0.5
        return float(host[1]) * 0.5

    def __del__(self):
        ptr = getattr(self, "cfl_data", None)
        if ptr is not None and ptr.value and _lib._lib is not None:
            _lib.lib().fvmi_free(ptr)
            self.cfl_data = None


class UnsplitSimulator(BaseSimulator):
    """Shared body of the LxF, HLL, FORCE and KP07 wrappers: fields with
    ``gc`` ghost cells, one launch per substep with the full dt (LxF.py:106-
    121, KP07.py:108-126), swap after each.  The two-float CFL buffer
    alternates slots between launches (see fvmi.h); computeDt is
    ``dt_factor`` times the last launch's minimum."""

    scheme = None  # "lxf" | "hll" | "force" | "kp07"
    gc = 1
    nvars = 3
    dt_factor = 0.5
    default_waves_per_block = 4

    def _init_unsplit(self, fields, cfl_host, tile_rows, waves_per_block, math, xcd_remap):
        if self.context.autotuner is not None and tile_rows is None and waves_per_block is None:
            peak = self.context.autotuner.get_peak_performance(self.__class__)
            tile_rows = int(peak["tile_rows"])
            waves_per_block = int(peak["waves_per_block"])
        math = resolve_math(self.scheme, math)
        self.math = math
        self.launch = _lib.LaunchConfig(int(tile_rows or 0),
                                        int(waves_per_block or self.default_waves_per_block),
                                        _MATH[math], 1 if xcd_remap else 2)
        self.gc_x = self.gc
        self.gc_y = self.gc
        self.u0 = Common.ArakawaA2D(self.stream, self.nx, self.ny, self.gc, self.gc, fields)
        self.u1 = Common.ArakawaA2D(self.stream, self.nx, self.ny, self.gc, self.gc, [None] * len(fields))
        ptr = ctypes.c_void_p()
        _lib.call("fvmi_malloc", ctypes.byref(ptr), 8)
        self.cfl_data = ptr
        self.cfl_last = 1  # slot holding the latest minimum: the initial value
        init = np.array([np.inf, cfl_host], dtype=np.float32)
        _lib.call("fvmi_memcpy_async", self.cfl_data, init.ctypes.data, 8, _lib.MEMCPY_HTOD, self.stream.handle)
        self.stream.synchronize()

    def cfl_pointer(self):
        return self.cfl_data.value

    def substep(self, dt, step_number):
        slot = 1 - self.cfl_last
        self._launch(np.float32(dt), int(step_number), self.u0, self.u1, self.cfl_pointer(), slot)
        self.cfl_last = slot
        self.u0, self.u1 = self.u1, self.u0

    def _field_args(self, state):
        args = []
        for k in range(self.nvars):
            args += [state[k].data_ptr, state[k].pitch]
        return args

    def getOutput(self):
        return self.u0

    def check(self):
        self.u0.check(self.stream)
        self.u1.check(self.stream)

    def computeDt(self):
        """``dt_factor`` x the CFL minimum of the last launch (LxF.py:123-125,
        HLL.py:122-124, FORCE.py:127-129, KP07.py:136-138)."""
        host = np.zeros(2, dtype=np.float32)
        _lib.call("fvmi_memcpy_async", host.ctypes.data, self.cfl_data, 8, _lib.MEMCPY_DTOH, self.stream.handle)
        self.stream.synchronize()
        return float(host[self.cfl_last]) * self.dt_factor

    def __del__(self):
        ptr = getattr(self, "cfl_data", None)
        if ptr is not None and ptr.value and _lib._lib is not None:
            _lib.lib().fvmi_free(ptr)
            self.cfl_data = None


def swe_initial_cfl(h0, hu0, hv0, dx, dy, g):
    """Host CFL of the initial data (KP07_dimsplit.py:103-105), float32."""
    h0 = np.asarray(h0)
    hu0 = np.asarray(hu0)
    hv0 = np.asarray(hv0)
    dt_x = np.min(dx / (np.abs(hu0 / h0) + np.sqrt(g * h0)))
    dt_y = np.min(dy / (np.abs(hv0 / h0) + np.sqrt(g * h0)))
    return np.float32(min(dt_x, dt_y))
