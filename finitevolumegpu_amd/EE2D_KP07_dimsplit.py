"""Dimensionally split KP07-type scheme for the Euler equations with gravity
(GPUSimulators/EE2D_KP07_dimsplit.py; kernel GPUSimulators/cuda/EE2D_KP07_dimsplit.cu).

x fluxes are HLL, y fluxes central-upwind, and gravity acts after both sweeps,
as in the reference kernel.  The reference's predictor evaluates the flux with
gamma in the pressure argument and its CFL uses sqrt(gamma*rho); both are kept
so that results match the reference arithmetic.

The plain reference wrapper never swaps its buffers (EE2D_KP07_dimsplit.py:112-138);
this class swaps after every substep, which is what MPISimulator does
(MPISimulator.py:326) and what a time integrator needs.
"""

import numpy as np

from . import _lib
from .Simulator import BoundaryCondition, DimsplitSimulator


def euler_initial_cfl(rho, rho_u, rho_v, dx, dy, gamma):
    """Host CFL of the initial data (EE2D_KP07_dimsplit.py:106-108), float32."""
    rho = np.asarray(rho)
    dt_x = np.min(dx / (np.abs(np.asarray(rho_u) / rho) + np.sqrt(gamma * rho)))
    dt_y = np.min(dy / (np.abs(np.asarray(rho_v) / rho) + np.sqrt(gamma * rho)))
    return np.float32(min(dt_x, dt_y))


class EE2D_KP07_dimsplit(DimsplitSimulator):
    """Constructor signature of EE2D_KP07_dimsplit.__init__ (EE2D_KP07_dimsplit.py:57-67)."""

    scheme = "ee2d_kp07_dimsplit"
    nvars = 4

    def __init__(self, context, rho, rho_u, rho_v, E, nx, ny, dx, dy, g, gamma, theta=1.3, cfl_scale=0.9,
                 boundary_conditions=BoundaryCondition(), block_width=16, block_height=8,
                 tile_rows=None, waves_per_block=None, math=None, xcd_remap=True):
        super().__init__(context, nx, ny, dx, dy, boundary_conditions, cfl_scale, 2, block_width, block_height)
        self.g = np.float32(g)
        self.gamma = np.float32(gamma)
        self.theta = np.float32(theta)
        self._init_dimsplit([rho, rho_u, rho_v, E], euler_initial_cfl(rho, rho_u, rho_v, self.dx, self.dy, gamma),
                            tile_rows, waves_per_block, math, xcd_remap)

    def substep(self, dt, step_number, external=True, internal=True):
        """``external``/``internal`` select the edge strips / the interior, as the
        reference's overlap split does (EE2D_KP07_dimsplit.py:112-266).  Both
        launches apply the boundary conditions at the true domain edges.  The
        buffers are swapped only by a full (external and internal) substep; a
        split caller swaps with ``swapBuffers`` after both halves."""
        self.substepDimsplit(0.5 * dt, step_number, external, internal)

    def substepDimsplit(self, dt, substep, external=True, internal=True):
        dt = np.float32(dt)
        cfl = self.cfl_pointer()
        nx, ny = int(self.nx), int(self.ny)
        if external and internal:
            self._launch(dt, substep, self.u0, self.u1, cfl, 0, 0, nx, ny)
            self.u0, self.u1 = self.u1, self.u0
            return
        gx, gy = self.gc_x, self.gc_y
        if external:
            # north, south rows then west, east columns (corners once)
            self._launch(dt, substep, self.u0, self.u1, cfl, 0, ny - gy, nx, ny)
            self._launch(dt, substep, self.u0, self.u1, cfl, 0, 0, nx, gy)
            self._launch(dt, substep, self.u0, self.u1, cfl, 0, gy, gx, ny - gy)
            self._launch(dt, substep, self.u0, self.u1, cfl, nx - gx, gy, nx, ny - gy)
        if internal:
            self._launch(dt, substep, self.u0, self.u1, cfl, gx, gy, nx - gx, ny - gy)

    def swapBuffers(self):
        self.u0, self.u1 = self.u1, self.u0

    def _launch(self, dt, substep, u0, u1, cfl, x0=0, y0=0, x1=0, y1=0):
        _lib.call("fvmi_euler_kp07_dimsplit_substep", int(self.nx), int(self.ny), float(self.dx), float(self.dy),
                  float(dt), float(self.g), float(self.gamma), float(self.theta), substep,
                  int(self.boundary_conditions), *self._field_args(u0), *self._field_args(u1), cfl,
                  x0, y0, x1, y1, self.launch, self.stream.handle)
