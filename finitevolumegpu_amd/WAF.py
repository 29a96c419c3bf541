"""Weighted average flux (WAF) shallow-water scheme, dimensionally split
(GPUSimulators/WAF.py; kernel GPUSimulators/cuda/SWE2D_WAF.cu)."""

import numpy as np

from . import _lib
from .Simulator import BoundaryCondition, DimsplitSimulator, swe_initial_cfl


class WAF(DimsplitSimulator):
    """Constructor signature of WAF.__init__ (WAF.py:50-58).

    The reference kernel writes no CFL value, so its ``computeDt`` returns half
    the initial CFL forever (WAF.py:82, 115; SWE2D_WAF.cu:103-119).  Here the
    kernel reports the CFL like the other schemes; ``waf_static_dt=True``
    reproduces the reference behaviour."""

    scheme = "waf"
    nvars = 3

    def __init__(self, context, h0, hu0, hv0, nx, ny, dx, dy, g, cfl_scale=0.9,
                 boundary_conditions=BoundaryCondition(), block_width=16, block_height=16,
                 tile_rows=None, waves_per_block=None, math=None, xcd_remap=True, waf_static_dt=False):
        super().__init__(context, nx, ny, dx, dy, boundary_conditions, cfl_scale, 2, block_width, block_height)
        self.g = np.float32(g)
        self.waf_static_dt = bool(waf_static_dt)
        self._init_dimsplit([h0, hu0, hv0], swe_initial_cfl(h0, hu0, hv0, self.dx, self.dy, g),
                            tile_rows, waves_per_block, math, xcd_remap)

    def cfl_pointer(self):
        return None if self.waf_static_dt else self.cfl_data.value

    def _launch(self, dt, substep, u0, u1, cfl):
        _lib.call("fvmi_swe_waf_substep", int(self.nx), int(self.ny), float(self.dx), float(self.dy),
                  float(dt), float(self.g), substep, int(self.boundary_conditions),
                  *self._field_args(u0), *self._field_args(u1), cfl, self.launch, self.stream.handle)
