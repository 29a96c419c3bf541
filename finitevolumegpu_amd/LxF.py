"""Lax-Friedrichs shallow-water scheme, first order, unsplit
(GPUSimulators/LxF.py; kernel GPUSimulators/cuda/SWE2D_LxF.cu)."""

import numpy as np

from . import _lib
from .Simulator import BoundaryCondition, UnsplitSimulator, swe_initial_cfl


class LxF(UnsplitSimulator):
    """Constructor signature of LxF.__init__ (LxF.py:52-104)."""

    scheme = "lxf"
    gc = 1
    nvars = 3
    dt_factor = 0.5

    def __init__(self, context, h0, hu0, hv0, nx, ny, dx, dy, g, cfl_scale=0.9,
                 boundary_conditions=BoundaryCondition(), block_width=16, block_height=16,
                 tile_rows=None, waves_per_block=None, math=None, xcd_remap=True):
        super().__init__(context, nx, ny, dx, dy, boundary_conditions, cfl_scale, 1, block_width, block_height)
        self.g = np.float32(g)
        self._init_unsplit([h0, hu0, hv0], swe_initial_cfl(h0, hu0, hv0, self.dx, self.dy, g),
                           tile_rows, waves_per_block, math, xcd_remap)

    def _launch(self, dt, substep, u0, u1, cfl, slot):
        _lib.call("fvmi_swe_lxf_substep", int(self.nx), int(self.ny), float(self.dx), float(self.dy), float(dt),
                  float(self.g), int(self.boundary_conditions), *self._field_args(u0), *self._field_args(u1), cfl,
                  slot, self.launch, self.stream.handle)
