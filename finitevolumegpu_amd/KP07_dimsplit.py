"""Dimensionally split Kurganov-Petrova (2007) shallow-water scheme
(GPUSimulators/KP07_dimsplit.py; kernel GPUSimulators/cuda/SWE2D_KP07_dimsplit.cu).

Second order in space (minmod-theta slopes), half-step predictor, central-upwind
flux; two substeps per step, x-y then y-x.
"""

import numpy as np

from . import _lib
from .Simulator import BoundaryCondition, DimsplitSimulator, swe_initial_cfl


class KP07_dimsplit(DimsplitSimulator):
    """Constructor signature of KP07_dimsplit.__init__ (KP07_dimsplit.py:55-64), plus
    the gfx950 launch shape (``tile_rows``, ``waves_per_block``), the arithmetic
    policy ``math`` ('fast' | 'exact' | None for the scheme's default,
    Simulator.DEFAULT_MATH) and ``xcd_remap``."""

    scheme = "kp07_dimsplit"
    nvars = 3

    def __init__(self, context, h0, hu0, hv0, nx, ny, dx, dy, g, theta=1.3, cfl_scale=0.9,
                 boundary_conditions=BoundaryCondition(), block_width=16, block_height=16,
                 tile_rows=None, waves_per_block=None, math=None, xcd_remap=True):
        super().__init__(context, nx, ny, dx, dy, boundary_conditions, cfl_scale, 2, block_width, block_height)
        self.g = np.float32(g)
        self.theta = np.float32(theta)
        self._init_dimsplit([h0, hu0, hv0], swe_initial_cfl(h0, hu0, hv0, self.dx, self.dy, g),
                            tile_rows, waves_per_block, math, xcd_remap)

    def _launch(self, dt, substep, u0, u1, cfl):
        _lib.call("fvmi_swe_kp07_dimsplit_substep", int(self.nx), int(self.ny), float(self.dx), float(self.dy),
                  float(dt), float(self.g), float(self.theta), substep, int(self.boundary_conditions),
                  *self._field_args(u0), *self._field_args(u1), cfl, self.launch, self.stream.handle)
