"""Unsplit second-order Kurganov-Petrova shallow-water scheme with TVD
Runge-Kutta time stepping (GPUSimulators/KP07.py; kernel
GPUSimulators/cuda/SWE2D_KP07.cu)."""

import numpy as np

from . import _lib
from .Simulator import BoundaryCondition, UnsplitSimulator, swe_initial_cfl


def stepOrderToCodedInt(step, order):
    """Simulator.stepOrderToCodedInt: step << 16 | order."""
    return (int(step) << 16) | (int(order) & 0xFFFF)


class KP07(UnsplitSimulator):
    """Constructor signature of KP07.__init__ (KP07.py:54-106): ``order`` RK
    stages per step (1 or 2), computeDt = min(cfl) * 0.5**(order-1)."""

    scheme = "kp07"
    gc = 2
    nvars = 3

    def __init__(self, context, h0, hu0, hv0, nx, ny, dx, dy, g, theta=1.3, cfl_scale=0.9, order=2,
                 boundary_conditions=BoundaryCondition(), block_width=16, block_height=16,
                 tile_rows=None, waves_per_block=None, math=None, xcd_remap=True):
        if order not in (1, 2):
            raise ValueError("order must be 1 or 2")
        super().__init__(context, nx, ny, dx, dy, boundary_conditions, cfl_scale, order, block_width, block_height)
        self.g = np.float32(g)
        self.theta = np.float32(theta)
        self.order = order
        self.dt_factor = 0.5 ** (order - 1)
        self._init_unsplit([h0, hu0, hv0], swe_initial_cfl(h0, hu0, hv0, self.dx, self.dy, g),
                           tile_rows, waves_per_block, math, xcd_remap)

    def substepRK(self, dt, substep):
        self.substep(dt, substep)

    def _launch(self, dt, substep, u0, u1, cfl, slot):
        _lib.call("fvmi_swe_kp07_substep", int(self.nx), int(self.ny), float(self.dx), float(self.dy), float(dt),
                  float(self.g), float(self.theta), stepOrderToCodedInt(substep, self.order),
                  int(self.boundary_conditions), *self._field_args(u0), *self._field_args(u1), cfl, slot,
                  self.launch, self.stream.handle)
