"""Second-order HLL shallow-water scheme, dimensionally split
(GPUSimulators/HLL2.py; kernel GPUSimulators/cuda/SWE2D_HLL2.cu)."""

import numpy as np

from . import _lib
from .Simulator import BoundaryCondition, DimsplitSimulator, swe_initial_cfl


class HLL2(DimsplitSimulator):
    """Constructor signature of HLL2.__init__ (HLL2.py, theta default 1.8)."""

    scheme = "hll2"
    nvars = 3

    def __init__(self, context, h0, hu0, hv0, nx, ny, dx, dy, g, theta=1.8, cfl_scale=0.9,
                 boundary_conditions=BoundaryCondition(), block_width=16, block_height=16,
                 tile_rows=None, waves_per_block=None, math=None, xcd_remap=True):
        super().__init__(context, nx, ny, dx, dy, boundary_conditions, cfl_scale, 2, block_width, block_height)
        self.g = np.float32(g)
        self.theta = np.float32(theta)
        self._init_dimsplit([h0, hu0, hv0], swe_initial_cfl(h0, hu0, hv0, self.dx, self.dy, g),
                            tile_rows, waves_per_block, math, xcd_remap)

    def _launch(self, dt, substep, u0, u1, cfl):
        _lib.call("fvmi_swe_hll2_substep", int(self.nx), int(self.ny), float(self.dx), float(self.dy),
                  float(dt), float(self.g), float(self.theta), substep, int(self.boundary_conditions),
                  *self._field_args(u0), *self._field_args(u1), cfl, self.launch, self.stream.handle)
